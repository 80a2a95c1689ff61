"""The library's host scene builder (data.zig addMaterial / addBox /
addSphere / buildSupportLookup / upload derivations) against the oracle's
restatement (bit for bit) and against closed forms for mass, inertia,
counts and the support lookup."""
import ctypes as C

import numpy as np
import pytest

from harness import OrScene, oracle

SHAPES = [
    ("box", (0, -0.5, 0), (10, 0.5, 10), 0.0, 0.5, 0.0),
    ("box", (0.02, 2.5, 0), (0.5, 0.5, 0.5), 7800.0, 0.5, 0.0),
    ("sphere", (1.5, 3, 0), 0.4, 16, 1200.0, 0.2, 0.37),
    ("box", (3, 1, 0), (0.75, 0.2, 0.6), 8900.0, 0.3, 0.3),
    ("sphere", (-1.0, 3, 0), 0.5, 5, 1100.0, 0.6, 0.4),
    ("sphere", (-2.0, 1, 1), 0.3, 4, 1000.0, 0.6, 0.4),
]


def build_both(shapes=SHAPES):
    from seevr_amd import Scene, default_caps
    sc = Scene(default_caps(max_bodies=64, max_hull_verts=10000, max_vertices=10000, max_triangles=20000,
                            max_support_entries=64 * 193))
    lib = oracle()
    os_ = OrScene()
    lib.scene_init(C.byref(os_), 64, 10000, 20000, 10000)
    mats_p, mats_o = {}, {}
    f3 = C.c_float * 3
    for s in shapes:
        d = s[3] if s[0] == "box" else s[4]
        if d not in mats_p:
            mats_p[d] = sc.add_material(d)
            mats_o[d] = lib.scene_add_material(C.byref(os_), d)
        if s[0] == "box":
            _, c, h, d, fr, re = s
            sc.add_box(c, h, fr, re, mats_p[d])
            lib.scene_add_box(C.byref(os_), f3(*c), f3(*h), fr, re, mats_o[d])
        else:
            _, c, r, segs, d, fr, re = s
            sc.add_sphere(c, r, segs, fr, re, mats_p[d])
            lib.scene_add_sphere(C.byref(os_), f3(*c), r, segs, fr, re, mats_o[d])
    lib.scene_derive(C.byref(os_))
    return sc, os_


def oarr(os_, field, count, dtype):
    p = getattr(os_, field)
    buf = (C.c_char * (count * 4)).from_address(p)
    return np.frombuffer(bytes(buf), dtype=dtype)


def test_scene_matches_oracle_restatement_bitwise():
    sc, os_ = build_both()
    cnt = sc.counts()
    assert cnt["num_bodies"] == os_.num_bodies == len(SHAPES)
    assert cnt["num_vertices"] == os_.num_vertices
    assert cnt["num_hull_verts"] == os_.num_hull_verts
    assert cnt["num_triangles"] == os_.num_triangles
    assert cnt["num_support_entries"] == os_.num_support_entries
    n, nv, nh, nt = os_.num_bodies, os_.num_vertices, os_.num_hull_verts, os_.num_triangles
    pairs = [("BODY_POS", "body_pos", 3 * n, np.float32), ("BODY_QUAT", "body_quat", 4 * n, np.float32),
             ("BODY_INV_MASS", "body_inv_mass", n, np.float32),
             ("BODY_INV_INERTIA", "body_inv_inertia", 3 * n, np.float32),
             ("BODY_VERT_START", "body_vert_start", n, np.uint32),
             ("BODY_VERT_COUNT", "body_vert_count", n, np.uint32),
             ("GEOM_BODY_IDX", "geom_body_idx", n, np.uint32), ("GEOM_DATA", "geom_data", 8 * n, np.float32),
             ("GEOM_FRICTION", "geom_friction", n, np.float32),
             ("GEOM_RESTITUTION", "geom_restitution", n, np.float32),
             ("HULL_VERTS", "hull_verts", 3 * nh, np.float32), ("VERTICES", "vertices", 3 * nv, np.float32),
             ("ORIGINAL_VERTICES", "orig_vertices", 3 * nv, np.float32),
             ("INDICES", "indices", 3 * nt, np.uint32),
             ("SUPPORT_LOOKUP_OFFSET", "support_offset", n, np.uint32),
             ("SUPPORT_LOOKUP", "support_lookup", os_.num_support_entries, np.uint32)]
    for name, field, count, dt in pairs:
        got = sc.array(name)
        want = oarr(os_, field, count, dt)
        assert got.size == count, name
        assert (got.view(np.uint32) == want.view(np.uint32)).all(), name


def test_mass_inertia_counts_closed_form():
    sc, _ = build_both()
    im = sc.array("BODY_INV_MASS")
    ii = sc.array("BODY_INV_INERTIA").reshape(-1, 3)
    gd = sc.array("GEOM_DATA").reshape(-1, 8)
    start = 0
    for b, s in enumerate(SHAPES):
        if s[0] == "box":
            _, c, h, d, fr, re = s
            m = d * 8 * h[0] * h[1] * h[2]
            w = 2 * np.array(h, float)
            I = m / 12 * np.array([w[1] ** 2 + w[2] ** 2, w[0] ** 2 + w[2] ** 2, w[0] ** 2 + w[1] ** 2])
            nverts = 8
        else:
            _, c, r, segs, d, fr, re = s
            m = d * 4 / 3 * np.pi * r ** 3
            I = np.full(3, 0.4 * m * r * r)
            nverts = (segs + 1) ** 2
        if m == 0:
            assert im[b] == 0 and (ii[b] == 0).all()
        else:
            assert im[b] == pytest.approx(1 / m, rel=2e-6)
            assert ii[b] == pytest.approx(1 / I, rel=2e-6)
        assert gd[b, 0] == start and gd[b, 1] == nverts and (gd[b, 2:] == 0).all()
        start += nverts
    c = sc.counts()
    assert c["num_triangles"] == 12 * 3 + sum(2 * s[3] ** 2 for s in SHAPES if s[0] == "sphere")


def test_support_lookup_offsets_and_argmax():
    sc, _ = build_both()
    off = sc.array("SUPPORT_LOOKUP_OFFSET")
    lk = sc.array("SUPPORT_LOOKUP")
    hull = sc.array("HULL_VERTS").reshape(-1, 3).astype(np.float64)
    gd = sc.array("GEOM_DATA").reshape(-1, 8)
    expect_off, k = [], 0
    for s in SHAPES:
        nverts = 8 if s[0] == "box" else (s[3] + 1) ** 2
        if nverts > 32:
            expect_off.append(1 + 192 * k)
            k += 1
        else:
            expect_off.append(0)
    assert off.tolist() == expect_off
    perps = [((0, 1, 0), (0, 0, 1)), ((1, 0, 0), (0, 0, 1)), ((1, 0, 0), (0, 1, 0))]
    checked = 0
    for b, o in enumerate(off):
        if o == 0:
            continue
        s, cnt = int(gd[b, 0]), int(gd[b, 1])
        pts = hull[s:s + cnt]
        for ax in range(3):
            p1, p2 = np.array(perps[ax][0], float), np.array(perps[ax][1], float)
            for bin_ in range(64):
                a = bin_ / 64 * 2 * np.pi
                d = np.cos(a) * p1 + np.sin(a) * p2
                dots = pts @ d
                top = np.sort(dots)[-2:]
                got = int(lk[o + ax * 64 + bin_])
                # exact argmax where the best vertex is not a near-tie
                if top[1] - top[0] > 1e-5:
                    assert got == int(np.argmax(dots))
                    checked += 1
                else:
                    assert dots[got] >= top[1] - 1e-5
    assert checked > 100


def test_capacity_and_argument_errors():
    from seevr_amd import Scene, SeevrError, default_caps
    sc = Scene(default_caps(max_bodies=2, max_hull_verts=100, max_vertices=100, max_triangles=200,
                            max_support_entries=10))
    m = sc.add_material(1000.0)
    sc.add_box((0, 0, 0), (1, 1, 1), 0.5, 0.0, m)
    with pytest.raises(SeevrError, match="SV_E_CAPACITY"):
        sc.add_sphere((0, 0, 0), 1.0, 8, 0.5, 0.0, m)  # 81 verts > 32 needs a lookup slot of 192
    with pytest.raises(SeevrError, match="SV_E_INVALID"):
        sc.add_box((0, 0, 0), (1, 1, 1), 0.5, 0.0, 7)  # unknown material
    sc.add_box((0, 0, 0), (1, 1, 1), 0.5, 0.0, m)
    with pytest.raises(SeevrError, match="SV_E_CAPACITY"):
        sc.add_box((0, 0, 0), (1, 1, 1), 0.5, 0.0, m)  # body capacity 2
