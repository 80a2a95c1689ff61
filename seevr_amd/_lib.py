"""ctypes binding of libseevr_hip.so (declarations of include/seevr.h).

The library must have been built in-tree (seevr_amd/build.py).  Loading
fails loudly when it is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libseevr_hip.so"

# sv_buf ids (include/seevr.h)
BUF = dict(
    BODY_POS=0, BODY_QUAT=1, BODY_VEL=2, BODY_OMEGA=3, BODY_INV_MASS=4, BODY_INV_INERTIA=5,
    BODY_VERT_START=6, BODY_VERT_COUNT=7, GEOM_BODY_IDX=8, GEOM_DATA=9, GEOM_FRICTION=10,
    GEOM_RESTITUTION=11, HULL_VERTS=12, VERTICES=13, ORIGINAL_VERTICES=14, CONTACT_POS=15,
    CONTACT_NORMAL=16, CONTACT_PENETRATION=17, CONTACT_GEOM_A=18, CONTACT_GEOM_B=19,
    CONTACT_LAMBDA_N=20, ATOMIC_COUNTERS=21, BODY_AABB_MIN=22, BODY_AABB_MAX=23,
    COLLISION_PAIRS=24, AABB_TEMP_MIN=25, AABB_TEMP_MAX=26, AABB_TEMP_BODY=27,
    SUPPORT_LOOKUP_OFFSET=28, SUPPORT_LOOKUP=29, INDICES=30, BVH_CENTROIDS=31, BVH_MORTON=32,
    BVH_MORTON_TEMP=33, BVH_SORT_INDICES=34, BVH_SORT_TEMP=35, BVH_SCENE_BOUNDS=36,
    BVH_AABB_MIN=37, BVH_AABB_MAX=38, BVH_LEFT=39, BVH_RIGHT=40, BVH_COUNT=41, BVH_PARENT=42,
    BVH_PRIM_INDICES=43, BVH_FLAGS=44, BVH_TEMP=45, GEOM_HULL_RANGE=46,
)
BUF_COUNT = 47

# element dtype of each buffer ('f' float32, 'u' uint32)
BUF_DTYPE = {k: "f" for k in BUF}
for _k in ("BODY_VERT_START", "BODY_VERT_COUNT", "GEOM_BODY_IDX", "CONTACT_GEOM_A", "CONTACT_GEOM_B",
           "ATOMIC_COUNTERS", "COLLISION_PAIRS", "AABB_TEMP_BODY", "SUPPORT_LOOKUP_OFFSET",
           "SUPPORT_LOOKUP", "INDICES", "BVH_MORTON", "BVH_MORTON_TEMP", "BVH_SORT_INDICES",
           "BVH_SORT_TEMP", "BVH_LEFT", "BVH_RIGHT", "BVH_COUNT", "BVH_PARENT", "BVH_PRIM_INDICES",
           "BVH_FLAGS", "BVH_TEMP", "GEOM_HULL_RANGE"):
    BUF_DTYPE[_k] = "u"

STATUS_PAIR_OVERFLOW = 1
STATUS_CONTACT_OVERFLOW = 2
STATUS_SOLVER_TIMEOUT = 4

ERRORS = {-1: "SV_E_INVALID", -2: "SV_E_CAPACITY", -3: "SV_E_DEVICE", -4: "SV_E_NOMEM",
          -5: "SV_E_TIMEOUT", -6: "SV_E_RANGE"}


class Caps(C.Structure):
    _fields_ = [("max_bodies", C.c_uint32), ("max_hull_verts", C.c_uint32), ("max_vertices", C.c_uint32),
                ("max_triangles", C.c_uint32), ("max_pairs", C.c_uint32), ("max_contacts", C.c_uint32),
                ("max_support_entries", C.c_uint32)]


class Params(C.Structure):
    _fields_ = [("pgs_iterations", C.c_uint32), ("baumgarte", C.c_float), ("slop", C.c_float),
                ("linear_damping", C.c_float), ("angular_damping", C.c_float),
                ("hull_range_u32", C.c_uint32)]


class BatchCaps(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("max_scenes", "max_bodies", "max_hull_verts", "max_vertices",
                                          "max_triangles", "max_support_entries", "max_scene_bodies", "max_scene_pairs",
                                          "max_scene_contacts")]


class SceneCounts(C.Structure):
    _fields_ = [("num_vertices", C.c_uint32), ("num_triangles", C.c_uint32), ("num_bodies", C.c_uint32),
                ("num_geoms", C.c_uint32), ("num_hull_verts", C.c_uint32), ("num_materials", C.c_uint32),
                ("num_support_entries", C.c_uint32)]


class BvhArrays(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("centroids", "morton", "sort_indices", "prim_indices", "scene_bounds",
                                          "node_min", "node_max", "left", "right", "count", "parent")]


NOALLOC_MAX_BODIES = 1 << 20

EXPORTS = [
    "sv_default_caps", "sv_default_params", "sv_create", "sv_destroy", "sv_get_params", "sv_set_params",
    "sv_buffer_ptr", "sv_bind_buffer", "sv_upload", "sv_download", "sv_physics_step", "sv_read_counters",
    "sv_physics_dispatch", "sv_bvh_build", "sv_bvh_build_bodies", "sv_bvh_build_bodies_noalloc", "sv_bvh_dispatch", "sv_debug_atan2", "sv_debug_radix_sort", "sv_debug_scan", "sv_debug_sort_rows",
    "sv_debug_copy_bandwidth", "sv_debug_broadphase_scratch", "sv_debug_seq_streams", "sv_sweep_depth", "sv_debug_handoff", "sv_debug_xcc_ids", "sv_debug_set_sweep_variant", "sv_last_error", "sv_device_info", "sv_stage_times", "sv_enable_stage_timing", "sv_stage_timing_only", "sv_set_solver", "sv_solver_residual", "sv_debug_color_counts", "sv_set_broadphase", "sv_broadphase_method", "sv_set_warm_start", "sv_scene_create",
    "sv_scene_destroy", "sv_scene_add_material", "sv_scene_add_box", "sv_scene_add_sphere",
    "sv_scene_set_body_state", "sv_scene_get_counts", "sv_scene_array", "sv_scene_upload",
    "sv_scene_build_workload", "sv_scene_build_batch", "sv_scene_serialize", "sv_scene_deserialize",
    "sv_scene_save", "sv_scene_load", "sv_scene_read_state", "sv_batch_create", "sv_batch_destroy", "sv_batch_context",
    "sv_batch_set_scenes", "sv_batch_step", "sv_batch_read_counters", "sv_batch_set_config", "sv_batch_autotune",
    "sv_batch_set_pair_tile", "sv_batch_pair_tile",
    "sv_batch_set_phase_timing", "sv_batch_phase_times",
    "sv_batch_diff_reserve", "sv_batch_diff_forward", "sv_batch_diff_backward",
]

_lib = None


class SeevrError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


def load(path: Path | str | None = None) -> C.CDLL:
    """Load libseevr_hip.so and declare its signatures (cached)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"{p} is missing: build it with `python -m seevr_amd.build` (HIP is required)")
    lib = C.CDLL(str(p))
    vp, u32, f32, sz = C.c_void_p, C.c_uint32, C.c_float, C.c_size_t
    fp = C.POINTER(C.c_float)
    sig = {
        "sv_default_caps": (None, [C.POINTER(Caps)]),
        "sv_default_params": (None, [C.POINTER(Params)]),
        "sv_create": (C.c_int, [C.POINTER(Caps), C.POINTER(Params), C.POINTER(vp)]),
        "sv_destroy": (C.c_int, [vp]),
        "sv_get_params": (C.c_int, [vp, C.POINTER(Params)]),
        "sv_set_params": (C.c_int, [vp, C.POINTER(Params)]),
        "sv_buffer_ptr": (C.c_int, [vp, C.c_int, C.POINTER(vp), C.POINTER(sz)]),
        "sv_bind_buffer": (C.c_int, [vp, C.c_int, vp, sz]),
        "sv_upload": (C.c_int, [vp, C.c_int, vp, sz, sz]),
        "sv_download": (C.c_int, [vp, C.c_int, vp, sz, sz]),
        "sv_physics_step": (C.c_int, [vp, u32, u32, u32, u32, f32, fp]),
        "sv_read_counters": (C.c_int, [vp, C.POINTER(u32)]),
        "sv_physics_dispatch": (C.c_int, [vp, u32, u32, f32, fp, u32]),
        "sv_bvh_build": (C.c_int, [vp, u32]),
        "sv_bvh_build_bodies": (C.c_int, [vp, u32]),
        "sv_bvh_dispatch": (C.c_int, [vp, u32, u32]),
        "sv_bvh_build_bodies_noalloc": (C.c_int, [vp, vp, u32, C.POINTER(BvhArrays), vp]),
        "sv_debug_atan2": (C.c_int, [vp, fp, fp, fp, u32]),
        "sv_debug_copy_bandwidth": (C.c_int, [vp, sz, u32, C.POINTER(C.c_double)]),
        "sv_debug_xcc_ids": (C.c_int, [vp, u32, C.POINTER(u32)]),
        "sv_debug_radix_sort": (C.c_int, [vp, C.POINTER(u32), C.POINTER(u32), u32, u32, u32, C.POINTER(u32),
                                          C.POINTER(u32)]),
        "sv_debug_scan": (C.c_int, [vp, C.POINTER(u32), u32, u32, C.POINTER(u32), C.POINTER(u32)]),
        "sv_debug_sort_rows": (C.c_int, [vp, C.POINTER(u32), C.POINTER(u32), u32, u32, u32, C.POINTER(u32)]),
        "sv_debug_broadphase_scratch": (C.c_int, [u32, u32, C.POINTER(C.c_uint64)]),
        "sv_debug_seq_streams": (u32, [u32]),
        "sv_sweep_depth": (C.c_int, [vp, C.POINTER(u32), C.POINTER(C.c_uint64)]),
        "sv_debug_handoff": (C.c_int, [vp, u32, C.c_int, C.POINTER(C.c_double), C.POINTER(u32)]),
        "sv_debug_set_sweep_variant": (C.c_int, [vp, C.c_int]),
        "sv_last_error": (C.c_char_p, []),
        "sv_device_info": (C.c_int, [vp, C.c_char_p, sz, C.c_char_p, sz]),
        "sv_stage_times": (C.c_int, [vp, C.c_int, C.POINTER(C.c_char_p), C.POINTER(C.c_double),
                                     C.POINTER(C.c_uint64), C.c_int]),
        "sv_enable_stage_timing": (C.c_int, [vp, C.c_int]),
        "sv_stage_timing_only": (C.c_int, [vp, C.c_char_p]),
        "sv_set_solver": (C.c_int, [vp, C.c_int, u32]),
        "sv_set_broadphase": (C.c_int, [vp, C.c_int]),
        "sv_broadphase_method": (C.c_int, [vp, u32]),
        "sv_set_warm_start": (C.c_int, [vp, f32]),
        "sv_solver_residual": (C.c_int, [vp, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(u32)]),
        "sv_debug_color_counts": (C.c_int, [vp, C.POINTER(u32), u32]),
        "sv_scene_create": (C.c_int, [C.POINTER(Caps), C.POINTER(vp)]),
        "sv_scene_destroy": (C.c_int, [vp]),
        "sv_scene_add_material": (C.c_int, [vp, fp, f32, f32, fp, f32, f32, C.POINTER(u32)]),
        "sv_scene_add_box": (C.c_int, [vp, fp, fp, fp, f32, f32, u32, C.POINTER(u32)]),
        "sv_scene_add_sphere": (C.c_int, [vp, fp, f32, fp, u32, f32, f32, u32, C.POINTER(u32)]),
        "sv_scene_set_body_state": (C.c_int, [vp, u32, fp, fp, fp]),
        "sv_scene_get_counts": (C.c_int, [vp, C.POINTER(SceneCounts)]),
        "sv_scene_array": (C.c_int, [vp, C.c_int, C.POINTER(vp), C.POINTER(sz)]),
        "sv_scene_upload": (C.c_int, [vp, vp]),
        "sv_scene_build_workload": (C.c_int, [vp, C.c_int, u32, u32, C.c_uint64]),
        "sv_scene_build_batch": (C.c_int, [vp, u32, u32, u32, u32, u32, f32, C.c_uint64, C.POINTER(u32), fp]),
        "sv_scene_serialize": (C.c_int, [vp, vp, sz, C.POINTER(sz)]),
        "sv_scene_deserialize": (C.c_int, [vp, vp, sz]),
        "sv_scene_save": (C.c_int, [vp, C.c_char_p]),
        "sv_scene_load": (C.c_int, [vp, C.c_char_p]),
        "sv_scene_read_state": (C.c_int, [vp, vp]),
        "sv_batch_create": (C.c_int, [C.POINTER(BatchCaps), C.POINTER(Params), C.POINTER(vp)]),
        "sv_batch_destroy": (C.c_int, [vp]),
        "sv_batch_context": (vp, [vp]),
        "sv_batch_set_scenes": (C.c_int, [vp, u32, C.POINTER(u32), fp]),
        "sv_batch_step": (C.c_int, [vp, u32, u32, f32]),
        "sv_batch_read_counters": (C.c_int, [vp, C.POINTER(u32)]),
        "sv_batch_set_config": (C.c_int, [vp, u32, u32, C.c_int]),
        "sv_batch_autotune": (C.c_int, [vp, u32, f32, C.POINTER(u32), C.POINTER(u32), C.POINTER(C.c_double)]),
        "sv_batch_set_pair_tile": (C.c_int, [vp, u32]),
        "sv_batch_pair_tile": (C.c_int, [vp, C.POINTER(u32)]),
        "sv_batch_set_phase_timing": (C.c_int, [vp, C.c_int]),
        "sv_batch_phase_times": (C.c_int, [vp, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]),
        "sv_batch_diff_reserve": (C.c_int, [vp, u32, u32]),
        "sv_batch_diff_forward": (C.c_int, [vp, u32, u32, f32]),
        "sv_batch_diff_backward": (C.c_int, [vp, u32, u32, f32, vp, vp, vp, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def check(rc: int) -> None:
    if rc != 0:
        lib = load()
        raise SeevrError(rc, lib.sv_last_error().decode(errors="replace"))
