"""The C-ABI library loads, exports every entry point include/seevr.h
declares, and its defaults equal the reference constants.  No device work."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    text = (ROOT / "include" / "seevr.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sv_\w+)\s*\(", text)))


def test_library_exports_every_declared_entry_point():
    from seevr_amd import _lib
    lib = _lib.load()
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIB_PATH)], capture_output=True, text=True)
    exported = set(re.findall(r"\bT (sv_\w+)", out.stdout))
    assert set(names) <= exported
    assert set(_lib.EXPORTS) == set(names)


def test_defaults_equal_reference_constants():
    from seevr_amd import default_caps, default_params
    c = default_caps()
    assert (c.max_bodies, c.max_contacts, c.max_pairs) == (256, 4000, 10000)  # data.zig:13-16
    assert (c.max_vertices, c.max_triangles, c.max_hull_verts) == (10_000_000,) * 3  # data.zig:11-17
    p = default_params()
    assert p.pgs_iterations == 4  # physics.comp:68
    assert (p.baumgarte, p.slop) == pytest.approx((0.2, 0.005))  # physics.comp:66-67
    assert (p.linear_damping, p.angular_damping) == pytest.approx((0.999, 0.998))  # physics.comp:422-427


def test_buffer_ids_follow_binding_tables():
    from seevr_amd._lib import BUF
    # physics.zig:37-68 binding order
    order = ["BODY_POS", "BODY_QUAT", "BODY_VEL", "BODY_OMEGA", "BODY_INV_MASS", "BODY_INV_INERTIA",
             "BODY_VERT_START", "BODY_VERT_COUNT", "GEOM_BODY_IDX", "GEOM_DATA", "GEOM_FRICTION",
             "GEOM_RESTITUTION", "HULL_VERTS", "VERTICES", "ORIGINAL_VERTICES", "CONTACT_POS", "CONTACT_NORMAL",
             "CONTACT_PENETRATION", "CONTACT_GEOM_A", "CONTACT_GEOM_B", "CONTACT_LAMBDA_N", "ATOMIC_COUNTERS",
             "BODY_AABB_MIN", "BODY_AABB_MAX", "COLLISION_PAIRS", "AABB_TEMP_MIN", "AABB_TEMP_MAX",
             "AABB_TEMP_BODY", "SUPPORT_LOOKUP_OFFSET", "SUPPORT_LOOKUP"]
    assert [BUF[k] for k in order] == list(range(30))
    header = (ROOT / "include" / "seevr.h").read_text()
    for k, v in BUF.items():
        assert re.search(rf"SV_{k}\s*=\s*{v}\b", header), k


def test_errors_are_reported_not_raised():
    from seevr_amd import _lib
    lib = _lib.load()
    h = C.c_void_p()
    caps = _lib.Caps()
    lib.sv_default_caps(C.byref(caps))
    caps.max_bodies = 0
    assert lib.sv_create(C.byref(caps), None, C.byref(h)) == -1
    assert b"max_bodies" in lib.sv_last_error()
    assert lib.sv_physics_step(None, 1, 1, 1, 1, C.c_float(0.1), (C.c_float * 3)()) == -1
    assert lib.sv_upload(None, 0, None, 0, 0) == -1
    assert lib.sv_destroy(None) == 0
