"""Test harness: ctypes binding of the CPU oracle (oracle/_build) and a host
World that holds a full copy of the physics buffers, runs the oracle on them
and moves them to and from a device Context.

Only tests, the smoke check and the benchmark's CPU-baseline leg use this.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
ORACLE_DIR = ROOT / "oracle"
ORACLE_LIB = ORACLE_DIR / "_build" / "libsvoracle.so"
ORACLE_LIB64 = ORACLE_DIR / "_build" / "libsvoracle64.so"
ORACLE_LIB_MT = ORACLE_DIR / "_build" / "libsvoracle_mt.so"  # OpenMP stages, same results

vp = C.c_void_p


class OrBufs(C.Structure):
    _fields_ = [(n, vp) for n in (
        "body_pos", "body_quat", "body_vel", "body_omega", "body_inv_mass", "body_inv_inertia",
        "body_vert_start", "body_vert_count", "geom_body_idx", "geom_data", "geom_friction",
        "geom_restitution", "hull_verts", "vertices", "orig_vertices", "contact_pos", "contact_normal",
        "contact_pen", "contact_geom_a", "contact_geom_b", "lambda_n", "counters", "aabb_min", "aabb_max",
        "pairs", "support_offset", "support_lookup", "hull_range")] + [
        ("max_pairs", C.c_uint32), ("max_contacts", C.c_uint32), ("pgs_iterations", C.c_uint32),
        ("baumgarte", C.c_float), ("slop", C.c_float), ("lin_damping", C.c_float), ("ang_damping", C.c_float)]


class OrBufs64(C.Structure):
    _fields_ = OrBufs._fields_[:-4] + [("baumgarte", C.c_double), ("slop", C.c_double),
                                       ("lin_damping", C.c_double), ("ang_damping", C.c_double)]


class OrBvh(C.Structure):
    _fields_ = [(n, vp) for n in (
        "vertices", "indices", "centroids", "morton", "sort_idx", "bounds", "node_min", "node_max",
        "node_left", "node_right", "node_count", "node_parent", "prim_idx", "flags")]


class OrScene(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in (
        "cap_bodies", "cap_verts", "cap_tris", "cap_hull", "cap_lookup", "cap_mat", "num_vertices",
        "num_triangles", "num_bodies", "num_geoms", "num_hull_verts", "num_materials",
        "num_support_entries")] + [(n, vp) for n in (
        "vertices", "orig_vertices", "indices", "body_pos", "body_quat", "body_vel", "body_omega",
        "body_inv_mass", "body_inertia", "body_inv_inertia", "body_vert_start", "body_vert_count",
        "geom_body_idx", "geom_data", "geom_friction", "geom_restitution", "geom_vert_start",
        "geom_vert_count", "hull_verts", "support_offset", "support_lookup", "material_density")]


_or = {}


def build_oracle() -> None:
    subprocess.run(["make", "-s", "-C", str(ORACLE_DIR)], check=True)


def oracle(double: bool = False, mt: bool = False) -> C.CDLL:
    key = "f64" if double else ("f32mt" if mt else "f32")
    if key in _or:
        return _or[key]
    path = ORACLE_LIB64 if double else (ORACLE_LIB_MT if mt else ORACLE_LIB)
    if not path.exists():
        build_oracle()
    lib = C.CDLL(str(path))
    p = "or64_" if double else "or_"
    R = C.c_double if double else C.c_float
    B = C.POINTER(OrBufs64 if double else OrBufs)
    T = C.POINTER(OrBvh)
    u32 = C.c_uint32
    sig = {
        "atan2f": (C.c_float, [C.c_float, C.c_float]),
        "threads": (C.c_int, []),
        "step_gravity": (None, [B, u32, R, C.POINTER(R)]),
        "step_compute_aabb": (None, [B, u32]),
        "step_broad_phase": (None, [B, u32]),
        "step_narrow_phase": (None, [B]),
        "step_clear": (None, [B]),
        "step_solve": (None, [B, R]),
        "step_integrate": (None, [B, u32, R]),
        "step_update_verts": (None, [B, u32, u32]),
        "step_clear_counters": (None, [B]),
        "physics_step": (None, [B, u32, u32, u32, u32, C.c_float, C.POINTER(C.c_float)]),
        "step_solve_warm": (None, [B, R]),
        "step_solve_in_order": (None, [B, R, C.c_int, vp]),
        "warm_lookup": (None, [B, vp, vp, u32, R]),
        "warm_save": (u32, [B, vp, vp]),
        "physics_step_warm": (None, [B, u32, u32, u32, C.c_float, C.POINTER(C.c_float), R, vp, vp,
                                     C.POINTER(u32)]),
        "support": (None, [B, u32, C.POINTER(R), C.POINTER(R)]),
        "mpr": (C.c_int, [B, u32, u32, C.POINTER(R), C.POINTER(R), C.POINTER(R)]),
        "expand_bits": (u32, [u32]),
        "morton3d": (u32, [R, R, R]),
        "clz32": (C.c_int, [u32]),
        "bvh_centroids": (None, [T, u32]),
        "bvh_bounds": (None, [T, u32]),
        "bvh_morton": (None, [T, u32]),
        "bvh_radix_sort": (None, [T, u32]),
        "bvh_hierarchy": (None, [T, u32]),
        "bvh_leaf_aabb": (None, [T, u32]),
        "bvh_propagate": (None, [T, u32]),
        "bvh_build": (None, [T, u32]),
        "bvh_build_bodies": (None, [T, vp, vp, u32]),
        "scene_init": (C.c_int, [C.POINTER(OrScene), u32, u32, u32, u32]),
        "scene_free": (None, [C.POINTER(OrScene)]),
        "scene_add_material": (u32, [C.POINTER(OrScene), C.c_float]),
        "scene_add_box": (u32, [C.POINTER(OrScene), C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_float,
                                C.c_float, u32]),
        "scene_add_sphere": (u32, [C.POINTER(OrScene), C.POINTER(C.c_float), C.c_float, u32, C.c_float,
                                   C.c_float, u32]),
        "scene_derive": (None, [C.POINTER(OrScene)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, p + name)
        fn.restype = res
        fn.argtypes = args
        setattr(lib, name, fn)
    _or[key] = lib
    return lib


def ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(vp)


# Names of the per-body / per-geom state a World holds, with (scene buffer
# name, dtype, components per body).
STATE = ["BODY_POS", "BODY_QUAT", "BODY_VEL", "BODY_OMEGA"]
STATIC = ["BODY_INV_MASS", "BODY_INV_INERTIA", "BODY_VERT_START", "BODY_VERT_COUNT", "GEOM_BODY_IDX",
          "GEOM_DATA", "GEOM_FRICTION", "GEOM_RESTITUTION", "HULL_VERTS", "VERTICES", "ORIGINAL_VERTICES",
          "SUPPORT_LOOKUP_OFFSET", "SUPPORT_LOOKUP", "GEOM_HULL_RANGE", "INDICES"]


class World:
    """Host copy of every physics buffer in the reference layout."""

    def __init__(self, arrays: dict, counts: dict, max_pairs: int, max_contacts: int, pgs_iterations: int = 30,
                 double: bool = False):
        self.double = double
        self.rt = np.float64 if double else np.float32
        self.a = {}
        for k, v in arrays.items():
            v = np.array(v, copy=True)
            if v.dtype == np.float32 and double:
                v = v.astype(np.float64)
            self.a[k] = v
        if self.a["SUPPORT_LOOKUP"].size == 0:
            self.a["SUPPORT_LOOKUP"] = np.zeros(1, np.uint32)
        self.n = counts["num_bodies"]
        self.nv = counts["num_vertices"]
        self.nh = counts["num_hull_verts"]
        self.nt = counts["num_triangles"]
        self.max_pairs = max_pairs
        self.max_contacts = max_contacts
        self.pgs_iterations = pgs_iterations
        rt = self.rt
        self.a["CONTACT_POS"] = np.zeros(3 * max_contacts + 3, rt)
        self.a["CONTACT_NORMAL"] = np.zeros(3 * max_contacts + 3, rt)
        self.a["CONTACT_PENETRATION"] = np.zeros(max_contacts + 1, rt)
        self.a["CONTACT_GEOM_A"] = np.zeros(max_contacts + 1, np.uint32)
        self.a["CONTACT_GEOM_B"] = np.zeros(max_contacts + 1, np.uint32)
        self.a["CONTACT_LAMBDA_N"] = np.zeros(max_contacts + 1, rt)
        self.a["ATOMIC_COUNTERS"] = np.zeros(4, np.uint32)
        self.a["BODY_AABB_MIN"] = np.zeros(3 * self.n, rt)
        self.a["BODY_AABB_MAX"] = np.zeros(3 * self.n, rt)
        self.a["COLLISION_PAIRS"] = np.zeros(2 * max_pairs + 2, np.uint32)
        self.use_hull_range = False
        self._bufs = None

    @classmethod
    def from_scene(cls, scene, max_pairs: int, max_contacts: int, pgs_iterations: int = 30, double=False):
        arrays = {k: scene.array(k) for k in STATE + STATIC}
        return cls(arrays, scene.counts(), max_pairs, max_contacts, pgs_iterations, double)

    def copy(self) -> "World":
        w = World.__new__(World)
        w.__dict__.update(self.__dict__)
        w.a = {k: v.copy() for k, v in self.a.items()}
        for k in ("warm_keys", "warm_lam"):
            if getattr(self, k, None) is not None:
                setattr(w, k, getattr(self, k).copy())
        w._bufs = None
        return w

    # -- oracle ------------------------------------------------------------
    def bufs(self):
        B = OrBufs64 if self.double else OrBufs
        b = B()
        a = self.a
        names = {
            "body_pos": "BODY_POS", "body_quat": "BODY_QUAT", "body_vel": "BODY_VEL",
            "body_omega": "BODY_OMEGA", "body_inv_mass": "BODY_INV_MASS",
            "body_inv_inertia": "BODY_INV_INERTIA", "body_vert_start": "BODY_VERT_START",
            "body_vert_count": "BODY_VERT_COUNT", "geom_body_idx": "GEOM_BODY_IDX", "geom_data": "GEOM_DATA",
            "geom_friction": "GEOM_FRICTION", "geom_restitution": "GEOM_RESTITUTION",
            "hull_verts": "HULL_VERTS", "vertices": "VERTICES", "orig_vertices": "ORIGINAL_VERTICES",
            "contact_pos": "CONTACT_POS", "contact_normal": "CONTACT_NORMAL",
            "contact_pen": "CONTACT_PENETRATION", "contact_geom_a": "CONTACT_GEOM_A",
            "contact_geom_b": "CONTACT_GEOM_B", "lambda_n": "CONTACT_LAMBDA_N", "counters": "ATOMIC_COUNTERS",
            "aabb_min": "BODY_AABB_MIN", "aabb_max": "BODY_AABB_MAX", "pairs": "COLLISION_PAIRS",
            "support_offset": "SUPPORT_LOOKUP_OFFSET", "support_lookup": "SUPPORT_LOOKUP",
        }
        for f, k in names.items():
            setattr(b, f, a[k].ctypes.data)
        b.hull_range = a["GEOM_HULL_RANGE"].ctypes.data if self.use_hull_range else None
        b.max_pairs = self.max_pairs
        b.max_contacts = self.max_contacts
        b.pgs_iterations = self.pgs_iterations
        b.baumgarte, b.slop, b.lin_damping, b.ang_damping = 0.2, 0.005, 0.999, 0.998
        if not self.double:
            # float constants exactly as the shader literals round them
            b.baumgarte = float(np.float32(0.2))
            b.slop = float(np.float32(0.005))
            b.lin_damping = float(np.float32(0.999))
            b.ang_damping = float(np.float32(0.998))
        for k, v in getattr(self, "param_over", {}).items():
            setattr(b, k, v)
        self._bufs = b
        return b

    def lib(self, mt: bool = False):
        return oracle(self.double, mt and not self.double)

    def oracle_step(self, substeps=10, dt=1.0 / 60.0, g=(0.0, -9.81, 0.0), mt: bool = False):
        """physics_step of the oracle; mt=True runs the OpenMP build (same results)."""
        b = self.bufs()
        gg = (C.c_float * 3)(*g)
        self.lib(mt).physics_step(C.byref(b), self.n, self.nv, self.nh, substeps, C.c_float(dt), gg)

    def oracle_step_warm(self, factor: float, substeps=10, dt=1.0 / 60.0, g=(0.0, -9.81, 0.0)):
        """physics_step with warm start; the cache lives on the World."""
        if getattr(self, "warm_keys", None) is None:
            self.warm_reset()
        b = self.bufs()
        R = C.c_double if self.double else C.c_float
        n = C.c_uint32(self.warm_n)
        self.lib().physics_step_warm(C.byref(b), self.n, self.nv, substeps, C.c_float(dt), (C.c_float * 3)(*g),
                                     R(factor), ptr(self.warm_keys), ptr(self.warm_lam), C.byref(n))
        self.warm_n = n.value

    def warm_reset(self):
        self.warm_keys = np.zeros(max(self.max_contacts, 1), np.uint64)
        self.warm_lam = np.zeros(max(self.max_contacts, 1), np.float64 if self.double else np.float32)
        self.warm_n = 0

    def oracle_stage(self, step: int, count: int = None, dt: float = 0.0, g=(0.0, 0.0, 0.0), mt: bool = False):
        """One stage by its reference step id; mt=True runs the OpenMP build (same results)."""
        b = self.bufs()
        lib = self.lib(mt)
        R = C.c_double if self.double else C.c_float
        count = self.n if count is None else count
        if step == 0:
            lib.step_gravity(C.byref(b), count, R(dt), (R * 3)(*g))
        elif step in (1, 9):
            lib.step_compute_aabb(C.byref(b), count)
        elif step == 2:
            lib.step_broad_phase(C.byref(b), count)
        elif step == 3:
            lib.step_narrow_phase(C.byref(b))
        elif step == 4:
            lib.step_clear(C.byref(b))
        elif step == 5:
            lib.step_solve(C.byref(b), R(dt))
        elif step == 6:
            lib.step_integrate(C.byref(b), count, R(dt))
        elif step == 7:
            lib.step_update_verts(C.byref(b), count, self.n)
        elif step == 8:
            lib.step_clear_counters(C.byref(b))
        else:
            raise ValueError(step)

    # -- views ---------------------------------------------------------------
    def pairs(self) -> np.ndarray:
        n = min(int(self.a["ATOMIC_COUNTERS"][0]), self.max_pairs)
        return self.a["COLLISION_PAIRS"][: 2 * n].reshape(-1, 2)

    def ncontacts(self) -> int:
        return min(int(self.a["ATOMIC_COUNTERS"][1]), self.max_contacts)

    def contacts(self) -> dict:
        nc = self.ncontacts()
        a = self.a
        return dict(pos=a["CONTACT_POS"][: 3 * nc], normal=a["CONTACT_NORMAL"][: 3 * nc],
                    pen=a["CONTACT_PENETRATION"][:nc], ga=a["CONTACT_GEOM_A"][:nc], gb=a["CONTACT_GEOM_B"][:nc],
                    lam=a["CONTACT_LAMBDA_N"][:nc])

    # -- device transfer -----------------------------------------------------------
    def upload(self, ctx, names=None) -> None:
        names = names or (STATE + STATIC + ["CONTACT_POS", "CONTACT_NORMAL", "CONTACT_PENETRATION",
                                            "CONTACT_GEOM_A", "CONTACT_GEOM_B", "CONTACT_LAMBDA_N",
                                            "ATOMIC_COUNTERS", "BODY_AABB_MIN", "BODY_AABB_MAX",
                                            "COLLISION_PAIRS"])
        for k in names:
            v = self.a[k]
            nbytes = ctx.buffer_bytes(k)
            cnt = min(v.size, nbytes // 4)
            if cnt:
                ctx.upload(k, v[:cnt].astype(np.float32 if v.dtype.kind == "f" else np.uint32))

    def download_state(self, ctx) -> dict:
        n = self.n
        return {"BODY_POS": ctx.download("BODY_POS", 3 * n), "BODY_QUAT": ctx.download("BODY_QUAT", 4 * n),
                "BODY_VEL": ctx.download("BODY_VEL", 3 * n), "BODY_OMEGA": ctx.download("BODY_OMEGA", 3 * n)}


def device_contacts(ctx, max_contacts: int) -> dict:
    cnt = ctx.download("ATOMIC_COUNTERS", 4)
    nc = min(int(cnt[1]), max_contacts)
    return dict(pos=ctx.download("CONTACT_POS", 3 * nc), normal=ctx.download("CONTACT_NORMAL", 3 * nc),
                pen=ctx.download("CONTACT_PENETRATION", nc), ga=ctx.download("CONTACT_GEOM_A", nc),
                gb=ctx.download("CONTACT_GEOM_B", nc), lam=ctx.download("CONTACT_LAMBDA_N", nc))


def device_pairs(ctx, max_pairs: int) -> np.ndarray:
    cnt = ctx.download("ATOMIC_COUNTERS", 4)
    n = min(int(cnt[0]), max_pairs)
    return ctx.download("COLLISION_PAIRS", 2 * n).reshape(-1, 2)


def bits(a: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a.view(np.uint32)


def assert_bitwise(got: np.ndarray, want: np.ndarray, what: str) -> None:
    got = np.asarray(got)
    want = np.asarray(want)
    assert got.shape == want.shape, f"{what}: shape {got.shape} != {want.shape}"
    if got.dtype.kind == "f":
        g, w = bits(got), bits(want)
        same = (g == w) | (np.isnan(got) & np.isnan(want))
    else:
        same = got == want
    if not same.all():
        idx = np.flatnonzero(~same)[:8]
        raise AssertionError(f"{what}: {(~same).sum()} of {same.size} differ; first {idx.tolist()}: "
                             f"got {got.ravel()[idx].tolist()} want {want.ravel()[idx].tolist()}")


def rel_close(got: np.ndarray, want: np.ndarray, rtol: float, atol: float = 0.0) -> tuple[bool, float]:
    got = np.asarray(got, np.float64)
    want = np.asarray(want, np.float64)
    err = np.abs(got - want) / np.maximum(np.abs(want), 1e-30)
    ok = np.abs(got - want) <= rtol * np.abs(want) + atol
    return bool(ok.all()), float(np.max(np.where(np.abs(want) > 0, err, np.abs(got - want)), initial=0.0))


def make_scene(workload: int, count: int, seed: int = 1, segments: int = 0, extra_caps: dict | None = None):
    """Build a seeded workload scene with capacities that fit it."""
    from seevr_amd import Scene, default_caps
    big = default_caps(max_bodies=count + 16, max_hull_verts=(count + 16) * 300, max_vertices=(count + 16) * 300,
                       max_triangles=(count + 16) * 600, max_support_entries=(count + 16) * 193 + 1)
    sc = Scene(big)
    sc.build_workload(workload, count, segments, seed)
    return sc


class HipMem:
    """Caller-owned device memory and streams for tests of entry points that
    take device pointers, straight from the HIP runtime the library loaded
    (torch would bring its own copy of the runtime)."""

    _hip = None

    def __init__(self):
        if HipMem._hip is None:
            from seevr_amd import _lib
            _lib.load()  # the library's runtime is the process's libamdhip64
            h = C.CDLL("libamdhip64.so.7")
            for name, args in {"hipMalloc": [C.POINTER(C.c_void_p), C.c_size_t], "hipFree": [C.c_void_p],
                               "hipMemcpy": [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int],
                               "hipStreamCreate": [C.POINTER(C.c_void_p)], "hipStreamDestroy": [C.c_void_p],
                               "hipDeviceSynchronize": []}.items():
                fn = getattr(h, name)
                fn.restype = C.c_int
                fn.argtypes = args
            HipMem._hip = h
        self.hip = HipMem._hip
        self.blocks = []
        self.streams = []

    def alloc(self, nbytes: int, fill=None) -> int:
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), max(nbytes, 16)) == 0, "hipMalloc"
        self.blocks.append(p.value)
        if fill is not None:
            self.upload(p.value, np.full(max(nbytes, 16) // 4, fill, np.int32))
        return p.value

    def upload(self, dev: int, arr: np.ndarray) -> None:
        a = np.ascontiguousarray(arr)
        assert self.hip.hipMemcpy(C.c_void_p(dev), a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0  # H2D

    def download(self, dev: int, count: int, dtype) -> np.ndarray:
        out = np.empty(count, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(dev), out.nbytes, 2) == 0  # D2H
        return out

    def stream(self) -> int:
        s = C.c_void_p()
        assert self.hip.hipStreamCreate(C.byref(s)) == 0
        self.streams.append(s.value)
        return s.value

    def close(self) -> None:
        self.hip.hipDeviceSynchronize()
        for s in self.streams:
            self.hip.hipStreamDestroy(C.c_void_p(s))
        for p in self.blocks:
            self.hip.hipFree(C.c_void_p(p))
        self.blocks, self.streams = [], []
