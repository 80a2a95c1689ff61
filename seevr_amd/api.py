"""Python view of the C ABI: a device context (seeVR's Physics + BVH + the
device side of Data) and a host scene builder (the host side of Data).

Method names follow the reference: Physics.step / readCounters
(src/physics.zig:80,213), BVH.build (src/bvh.zig:63), Data.addMaterial /
addBox / addSphere / upload (src/data.zig:356,392,487,599).
"""
from __future__ import annotations

import ctypes as C
from typing import Iterable, Sequence

import numpy as np

from . import _lib
from ._lib import BUF, BUF_DTYPE, BatchCaps, Caps, Params, SceneCounts, check


def _fvec(vals: Iterable[float], n: int):
    arr = (C.c_float * n)(*[float(v) for v in vals])
    return arr


def default_caps(**over) -> Caps:
    lib = _lib.load()
    c = Caps()
    lib.sv_default_caps(C.byref(c))
    for k, v in over.items():
        setattr(c, k, int(v))
    return c


def default_params(**over) -> Params:
    lib = _lib.load()
    p = Params()
    lib.sv_default_params(C.byref(p))
    for k, v in over.items():
        setattr(p, k, type(getattr(p, k))(v))
    return p


def _np_dtype(name: str):
    return np.float32 if BUF_DTYPE[name] == "f" else np.uint32


class Context:
    """One sv_ctx: device buffers at fixed capacities plus the stages."""

    def __init__(self, caps: Caps | None = None, params: Params | None = None, **caps_over):
        self.lib = _lib.load()
        self.caps = caps if caps is not None else default_caps(**caps_over)
        self.params = params if params is not None else default_params()
        h = C.c_void_p()
        check(self.lib.sv_create(C.byref(self.caps), C.byref(self.params), C.byref(h)))
        self.h = h
        self.owned = True

    @classmethod
    def _view(cls, handle, caps: Caps, params: Params) -> "Context":
        """A non-owning view of a context that another object (a Batch) owns."""
        ctx = cls.__new__(cls)
        ctx.lib = _lib.load()
        ctx.caps, ctx.params = caps, params
        ctx.h = C.c_void_p(handle)
        ctx.owned = False
        return ctx

    # -- lifetime -----------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "h", None):
            if getattr(self, "owned", True):
                self.lib.sv_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- buffers ------------------------------------------------------------
    def buffer_bytes(self, name: str) -> int:
        n = C.c_size_t()
        check(self.lib.sv_buffer_ptr(self.h, BUF[name], None, C.byref(n)))
        return n.value

    def buffer_ptr(self, name: str) -> int:
        p, n = C.c_void_p(), C.c_size_t()
        check(self.lib.sv_buffer_ptr(self.h, BUF[name], C.byref(p), C.byref(n)))
        return p.value or 0

    def bind_buffer(self, name: str, dev_ptr: int, nbytes: int) -> None:
        """Use caller-owned device memory as buffer `name` (sv_bind_buffer)."""
        check(self.lib.sv_bind_buffer(self.h, BUF[name], C.c_void_p(dev_ptr), nbytes))

    def upload(self, name: str, arr, offset_elems: int = 0) -> None:
        a = np.ascontiguousarray(arr, dtype=_np_dtype(name))
        check(self.lib.sv_upload(self.h, BUF[name], a.ctypes.data_as(C.c_void_p), offset_elems * 4, a.nbytes))

    def download(self, name: str, count: int, offset_elems: int = 0) -> np.ndarray:
        out = np.empty(count, dtype=_np_dtype(name))
        if count:
            check(self.lib.sv_download(self.h, BUF[name], out.ctypes.data_as(C.c_void_p), offset_elems * 4,
                                       out.nbytes))
        return out

    def set_params(self, **kw) -> None:
        for k, v in kw.items():
            setattr(self.params, k, type(getattr(self.params, k))(v))
        check(self.lib.sv_set_params(self.h, C.byref(self.params)))

    # -- physics --------------------------------------------------------------
    def step(self, num_bodies: int, num_vertices: int, num_hull_verts: int, substeps: int = 10,
             dt: float = 1.0 / 60.0, gravity: Sequence[float] = (0.0, -9.81, 0.0)) -> None:
        """Physics.step (src/physics.zig:80-165)."""
        check(self.lib.sv_physics_step(self.h, num_bodies, num_vertices, num_hull_verts, substeps,
                                       C.c_float(dt), _fvec(gravity, 3)))

    def read_counters(self) -> tuple[int, int]:
        """Physics.readCounters (src/physics.zig:213-217)."""
        out = (C.c_uint32 * 2)()
        check(self.lib.sv_read_counters(self.h, out))
        return int(out[0]), int(out[1])

    def status(self) -> int:
        return int(self.download("ATOMIC_COUNTERS", 3)[2])

    def dispatch(self, step: int, count: int, dt: float = 0.0, gravity: Sequence[float] = (0.0, 0.0, 0.0),
                 num_bodies: int = 0) -> None:
        """One physics.comp step by its reference step id."""
        check(self.lib.sv_physics_dispatch(self.h, step, count, C.c_float(dt), _fvec(gravity, 3), num_bodies))

    # -- LBVH -----------------------------------------------------------------
    def bvh_build(self, num_tris: int) -> None:
        """BVH.build (src/bvh.zig:63-172)."""
        check(self.lib.sv_bvh_build(self.h, num_tris))

    def bvh_build_bodies(self, num_bodies: int) -> None:
        check(self.lib.sv_bvh_build_bodies(self.h, num_bodies))

    def bvh_dispatch(self, step: int, n: int) -> None:
        check(self.lib.sv_bvh_dispatch(self.h, step, n))

    # -- diagnostics ----------------------------------------------------------
    def debug_atan2(self, y: np.ndarray, x: np.ndarray) -> np.ndarray:
        y = np.ascontiguousarray(y, np.float32)
        x = np.ascontiguousarray(x, np.float32)
        out = np.empty_like(y)
        fp = C.POINTER(C.c_float)
        check(self.lib.sv_debug_atan2(self.h, y.ctypes.data_as(fp), x.ctypes.data_as(fp), out.ctypes.data_as(fp),
                                      y.size))
        return out

    def copy_bandwidth(self, nbytes: int = 1 << 30, iters: int = 10) -> float:
        """GB/s (read + write) of a plain device copy of nbytes."""
        out = C.c_double()
        check(self.lib.sv_debug_copy_bandwidth(self.h, nbytes, iters, C.byref(out)))
        return out.value

    def enable_stage_timing(self, on: bool = True, only: str | None = None) -> None:
        """HIP-event timing of the library's stages; `only` names the one stage to time."""
        check(self.lib.sv_stage_timing_only(self.h, (only or "").encode()))
        check(self.lib.sv_enable_stage_timing(self.h, int(on)))

    SOLVERS = {"auto": 0, "ordered": 1, "flow": 2, "colored": 3}

    def set_solver(self, schedule: str = "auto", workgroups: int = 0) -> None:
        """Schedule of the contact sweep: "auto", "ordered" or "flow" (all bit-identical)."""
        check(self.lib.sv_set_solver(self.h, self.SOLVERS[schedule], workgroups))

    def residual(self) -> tuple[float, float, int]:
        """(RMS, max, count) of the normal-constraint residual after step 5."""
        r, m, n = C.c_double(), C.c_double(), C.c_uint32()
        check(self.lib.sv_solver_residual(self.h, C.byref(r), C.byref(m), C.byref(n)))
        return r.value, m.value, n.value

    def color_counts(self) -> list[int]:
        out = (C.c_uint32 * 64)()
        n = self.lib.sv_debug_color_counts(self.h, out, 64)
        if n < 0:
            check(n)
        return [int(out[i]) for i in range(n)]

    BROADPHASES = {"auto": 0, "all_pairs": 1, "lbvh": 2, "sweep": 3}

    def set_broadphase(self, method: str = "auto") -> None:
        """Broadphase method: "auto", "all_pairs", "lbvh" or "sweep" (identical pair lists)."""
        check(self.lib.sv_set_broadphase(self.h, self.BROADPHASES[method]))

    def broadphase_method(self, n: int) -> str:
        """The method a broadphase over n bodies takes under the current setting."""
        rc = self.lib.sv_broadphase_method(self.h, n)
        check(min(rc, 0))
        return {v: k for k, v in self.BROADPHASES.items()}[rc]

    SWEEP_VARIANTS = {"seq": 0, "coherent": 1, "fenced": 2, "pipelined": 3, "seq_basic": 4, "paired": 5}

    def set_sweep_variant(self, variant: str) -> None:
        """Hand-off variant of the multi-workgroup sweeps (same results)."""
        check(self.lib.sv_debug_set_sweep_variant(self.h, self.SWEEP_VARIANTS[variant]))

    def debug_radix_sort(self, keys, vals, key_bits: int, n_max: int | None = None):
        """The library's stable radix sort of (keys, vals) by the low key_bits
        bits, with the count read on the device from an n_max-sized buffer."""
        k = np.ascontiguousarray(keys, np.uint32)
        v = np.ascontiguousarray(vals, np.uint32)
        n = k.size
        ko = np.zeros(n, np.uint32)
        vo = np.zeros(n, np.uint32)
        u = C.POINTER(C.c_uint32)
        check(self.lib.sv_debug_radix_sort(self.h, k.ctypes.data_as(u), v.ctypes.data_as(u), n,
                                           n if n_max is None else n_max, key_bits, ko.ctypes.data_as(u),
                                           vo.ctypes.data_as(u)))
        return ko, vo

    def debug_scan(self, values, n_max: int | None = None):
        """The library's exclusive scan of `values` (uint32, sums mod 2^32),
        with the count read on the device from an n_max-sized buffer:
        (exclusive prefixes, total)."""
        v = np.ascontiguousarray(values, np.uint32)
        n = v.size
        out = np.zeros(n, np.uint32)
        tot = C.c_uint32()
        u = C.POINTER(C.c_uint32)
        check(self.lib.sv_debug_scan(self.h, v.ctypes.data_as(u), n, n if n_max is None else n_max,
                                     out.ctypes.data_as(u), C.byref(tot)))
        return out, int(tot.value)

    def debug_sort_rows(self, cnt, entries, max_pairs: int, window_words: int = 0) -> np.ndarray:
        """The broadphase's row sort: rows of cnt[i] partners (entries in row
        order, any order within a row) to the sorted (i, j) pairs below
        max_pairs, as an (m, 2) array."""
        cn = np.ascontiguousarray(cnt, np.uint32)
        en = np.ascontiguousarray(entries, np.uint32)
        kept = min(int(cn.sum()), max_pairs)
        out = np.zeros(2 * max(kept, 1), np.uint32)
        u = C.POINTER(C.c_uint32)
        check(self.lib.sv_debug_sort_rows(self.h, cn.ctypes.data_as(u), en.ctypes.data_as(u), cn.size, max_pairs,
                                          window_words, out.ctypes.data_as(u)))
        return out[:2 * kept].reshape(-1, 2)

    def sweep_depth(self) -> tuple[int, int]:
        """(depth, items) of the last sweep's dependency graph."""
        d = C.c_uint32()
        k = C.c_uint64()
        check(self.lib.sv_sweep_depth(self.h, C.byref(d), C.byref(k)))
        return int(d.value), int(k.value)

    def handoff_us(self, rounds: int = 20000, same_xcd: bool = False) -> tuple[float, tuple[int, int]]:
        """One-way hand-off latency (us) between two waves and their XCC ids."""
        us = C.c_double()
        x = (C.c_uint32 * 2)()
        check(self.lib.sv_debug_handoff(self.h, rounds, int(same_xcd), C.byref(us), x))
        return float(us.value), (int(x[0]), int(x[1]))

    def xcc_ids(self, nblocks: int) -> np.ndarray:
        """XCD each of nblocks one-wave blocks ran on (placement probe)."""
        out = np.zeros(nblocks, np.uint32)
        check(self.lib.sv_debug_xcc_ids(self.h, nblocks, out.ctypes.data_as(C.POINTER(C.c_uint32))))
        return out

    def set_warm_start(self, factor: float) -> None:
        """Warm start beside the reference path: 0 = off (the reference's
        cleared impulses); in (0, 1], contacts of a pair that was in contact
        at the previous sweep start from factor times its impulse."""
        check(self.lib.sv_set_warm_start(self.h, C.c_float(factor)))

    def stage_times(self, reset: bool = False) -> dict:
        cap = 64
        names = (C.c_char_p * cap)()
        avg = (C.c_double * cap)()
        calls = (C.c_uint64 * cap)()
        n = self.lib.sv_stage_times(self.h, int(reset), names, avg, calls, cap)
        if n < 0:
            check(n)
        return {names[i].decode(): (avg[i], int(calls[i])) for i in range(min(n, cap))}

    def device_info(self) -> tuple[str, str]:
        name = C.create_string_buffer(256)
        arch = C.create_string_buffer(64)
        check(self.lib.sv_device_info(self.h, name, 256, arch, 64))
        return name.value.decode(), arch.value.decode()


class Scene:
    """Host scene builder with data.zig's construction and layouts."""

    def __init__(self, caps: Caps | None = None, **caps_over):
        self.lib = _lib.load()
        self.caps = caps if caps is not None else default_caps(**caps_over)
        h = C.c_void_p()
        check(self.lib.sv_scene_create(C.byref(self.caps), C.byref(h)))
        self.h = h

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.sv_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_material(self, density: float, albedo=(0.8, 0.8, 0.8), roughness=0.5, metallic=0.0,
                     emission=(0.0, 0.0, 0.0), ior=0.0) -> int:
        out = C.c_uint32()
        check(self.lib.sv_scene_add_material(self.h, _fvec(albedo, 3), roughness, metallic, _fvec(emission, 3),
                                             ior, density, C.byref(out)))
        return out.value

    def add_box(self, center, half, friction=0.5, restitution=0.0, material=0, color=(0.8, 0.8, 0.8)) -> int:
        out = C.c_uint32()
        check(self.lib.sv_scene_add_box(self.h, _fvec(center, 3), _fvec(half, 3), _fvec(color, 3), friction,
                                        restitution, material, C.byref(out)))
        return out.value

    def add_sphere(self, center, radius, segments=16, friction=0.5, restitution=0.0, material=0,
                   color=(0.8, 0.8, 0.8)) -> int:
        out = C.c_uint32()
        check(self.lib.sv_scene_add_sphere(self.h, _fvec(center, 3), radius, _fvec(color, 3), segments, friction,
                                           restitution, material, C.byref(out)))
        return out.value

    def set_body_state(self, body: int, quat=None, vel=None, omega=None) -> None:
        q = _fvec(quat, 4) if quat is not None else None
        v = _fvec(vel, 3) if vel is not None else None
        w = _fvec(omega, 3) if omega is not None else None
        check(self.lib.sv_scene_set_body_state(self.h, body, q, v, w))

    def build_workload(self, workload: int, count: int = 0, segments: int = 0, seed: int = 1) -> None:
        check(self.lib.sv_scene_build_workload(self.h, workload, count, segments, seed))

    # -- scene files ----------------------------------------------------------
    def to_bytes(self) -> bytes:
        n = C.c_size_t()
        check(self.lib.sv_scene_serialize(self.h, None, 0, C.byref(n)))
        buf = (C.c_uint8 * n.value)()
        check(self.lib.sv_scene_serialize(self.h, buf, n.value, C.byref(n)))
        return bytes(buf)

    def from_bytes(self, data: bytes) -> None:
        buf = (C.c_uint8 * len(data)).from_buffer_copy(data)
        check(self.lib.sv_scene_deserialize(self.h, buf, len(data)))

    def save(self, path) -> None:
        check(self.lib.sv_scene_save(self.h, str(path).encode()))

    def load(self, path) -> None:
        check(self.lib.sv_scene_load(self.h, str(path).encode()))

    def read_state(self, ctx: "Context") -> None:
        """Body state and render vertices of a context back into the scene."""
        check(self.lib.sv_scene_read_state(self.h, ctx.h))

    def build_batch(self, first: int, count: int, min_bodies: int = 8, max_bodies: int = 256, segments: int = 0,
                    seed: int = 1, top: float = 0.0) -> tuple[np.ndarray, np.ndarray]:
        """Scenes [first, first + count) of the batched family; returns
        (body_start[count + 1], gravity_scale[count])."""
        start = np.zeros(count + 1, np.uint32)
        gs = np.zeros(count, np.float32)
        check(self.lib.sv_scene_build_batch(self.h, first, count, min_bodies, max_bodies, segments, C.c_float(top), seed,
                                            start.ctypes.data_as(C.POINTER(C.c_uint32)),
                                            gs.ctypes.data_as(C.POINTER(C.c_float))))
        return start, gs

    def counts(self) -> dict:
        c = SceneCounts()
        check(self.lib.sv_scene_get_counts(self.h, C.byref(c)))
        return {k: int(getattr(c, k)) for k, _ in SceneCounts._fields_}

    def array(self, name: str) -> np.ndarray:
        p, n = C.c_void_p(), C.c_size_t()
        check(self.lib.sv_scene_array(self.h, BUF[name], C.byref(p), C.byref(n)))
        dt = _np_dtype(name)
        cnt = n.value // 4
        if cnt == 0:
            return np.zeros(0, dtype=dt)
        buf = (C.c_char * n.value).from_address(p.value)
        return np.frombuffer(bytes(buf), dtype=dt).copy()

    def upload(self, ctx: Context) -> None:
        """Data.upload (src/data.zig:599-655)."""
        check(self.lib.sv_scene_upload(self.h, ctx.h))


def caps_for(scene: Scene, pairs: int, contacts: int, triangles: int | None = None) -> Caps:
    """Capacities that fit a built scene."""
    c = scene.counts()
    return default_caps(max_bodies=max(c["num_bodies"], 1), max_hull_verts=max(c["num_hull_verts"], 1),
                        max_vertices=max(c["num_vertices"], 1),
                        max_triangles=max(c["num_triangles"] if triangles is None else triangles, 1),
                        max_pairs=pairs, max_contacts=contacts,
                        max_support_entries=max(c["num_support_entries"], 1))


class Batch:
    """Batched engine (sv_batch_*): independent scenes stepped together."""

    def __init__(self, caps: BatchCaps, params: Params | None = None):
        self.lib = _lib.load()
        self.caps = caps
        self.params = params if params is not None else default_params()
        h = C.c_void_p()
        check(self.lib.sv_batch_create(C.byref(caps), C.byref(self.params), C.byref(h)))
        self.h = h
        cc = default_caps(max_bodies=caps.max_bodies, max_hull_verts=caps.max_hull_verts,
                          max_vertices=caps.max_vertices, max_triangles=max(caps.max_triangles, 1),
                          max_pairs=caps.max_scenes * caps.max_scene_pairs,
                          max_contacts=caps.max_scenes * caps.max_scene_contacts,
                          max_support_entries=caps.max_support_entries)
        self.ctx = Context._view(self.lib.sv_batch_context(h), cc, self.params)
        self.num_scenes = 0
        self.num_bodies = 0

    def close(self) -> None:
        if getattr(self, "h", None):
            self.ctx.close()
            self.lib.sv_batch_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_scenes(self, body_start, gravity) -> None:
        bs = np.ascontiguousarray(body_start, np.uint32)
        g = np.ascontiguousarray(gravity, np.float32).reshape(-1)
        n = bs.size - 1
        assert g.size == 3 * n
        check(self.lib.sv_batch_set_scenes(self.h, n, bs.ctypes.data_as(C.POINTER(C.c_uint32)),
                                           g.ctypes.data_as(C.POINTER(C.c_float))))
        self.num_scenes = n
        self.num_bodies = int(bs[-1])

    def step(self, num_vertices: int, substeps: int = 10, dt: float = 1.0 / 60.0) -> None:
        check(self.lib.sv_batch_step(self.h, num_vertices, substeps, C.c_float(dt)))

    def counters(self) -> np.ndarray:
        out = np.zeros(3 * max(self.num_scenes, 1), np.uint32)
        check(self.lib.sv_batch_read_counters(self.h, out.ctypes.data_as(C.POINTER(C.c_uint32))))
        return out[: 3 * self.num_scenes].reshape(-1, 3)

    def set_config(self, threads: int = 256, workgroups_per_cu: int = 2, use_graph: bool = True) -> None:
        check(self.lib.sv_batch_set_config(self.h, threads, workgroups_per_cu, int(use_graph)))

    def set_pair_tile(self, lanes_per_row: int = 1) -> None:
        """Broadphase lanes per row (1, 2, 4, ..., 64); the pairs are the same for every tile."""
        check(self.lib.sv_batch_set_pair_tile(self.h, lanes_per_row))

    @property
    def pair_tile(self) -> int:
        g = C.c_uint32()
        check(self.lib.sv_batch_pair_tile(self.h, C.byref(g)))
        return g.value

    def autotune(self, substeps: int = 10, dt: float = 1.0 / 60.0) -> tuple[int, int, float]:
        t, w, us = C.c_uint32(), C.c_uint32(), C.c_double()
        check(self.lib.sv_batch_autotune(self.h, substeps, C.c_float(dt), C.byref(t), C.byref(w), C.byref(us)))
        return t.value, w.value, us.value

    PHASES = ("next scene", "gravity+aabb", "broadphase", "narrowphase", "sweep setup", "sweep",
              "write-back+integrate")

    def set_phase_timing(self, on: bool = True) -> None:
        """Device wall-clock timing of the batched kernel's phases (see phase_times)."""
        check(self.lib.sv_batch_set_phase_timing(self.h, int(on)))

    def phase_times(self) -> tuple[dict, int]:
        """Microseconds each phase took, summed over workgroups, since the last
        call, and the scenes processed; zeroes the totals."""
        us = (C.c_double * len(self.PHASES))()
        n = C.c_uint64()
        check(self.lib.sv_batch_phase_times(self.h, us, C.byref(n)))
        return {k: us[i] for i, k in enumerate(self.PHASES)}, int(n.value)

    # -- differentiable rollout ---------------------------------------------
    def diff_reserve(self, max_frames: int, substeps: int = 10) -> None:
        check(self.lib.sv_batch_diff_reserve(self.h, max_frames, substeps))

    def diff_forward(self, frames: int, substeps: int = 10, dt: float = 1.0 / 60.0) -> None:
        check(self.lib.sv_batch_diff_forward(self.h, frames, substeps, C.c_float(dt)))

    def diff_backward(self, frames: int, grad_state, substeps: int = 10, dt: float = 1.0 / 60.0,
                      out=None):
        """grad_state: adjoint of the final packed state (13N floats; numpy
        array, device address as int, or None).  Returns (adjoint of the
        initial packed state, d/dfriction, d/drestitution) as numpy arrays, or
        writes into out = (state_addr, fric_addr, rest_addr) device addresses."""
        n = self.ctx.caps.max_bodies
        gs = None
        if isinstance(grad_state, int):
            gs = C.c_void_p(grad_state)
        elif grad_state is not None:
            ga = np.ascontiguousarray(grad_state, np.float32)
            gs = ga.ctypes.data_as(C.c_void_p)
        if out is not None:
            check(self.lib.sv_batch_diff_backward(self.h, frames, substeps, C.c_float(dt), gs,
                                                  *(C.c_void_p(a) for a in out)))
            return None
        nb = self.num_bodies
        so = np.zeros(13 * nb, np.float32)
        fo = np.zeros(nb, np.float32)
        ro = np.zeros(nb, np.float32)
        check(self.lib.sv_batch_diff_backward(self.h, frames, substeps, C.c_float(dt), gs,
                                              so.ctypes.data_as(C.c_void_p), fo.ctypes.data_as(C.c_void_p),
                                              ro.ctypes.data_as(C.c_void_p)))
        return so, fo, ro

    def packed_state(self) -> np.ndarray:
        nb = self.num_bodies
        return np.concatenate([self.ctx.download(k, w * nb) for k, w in
                               (("BODY_POS", 3), ("BODY_QUAT", 4), ("BODY_VEL", 3), ("BODY_OMEGA", 3))])

    def set_packed_state(self, x) -> None:
        nb = self.num_bodies
        x = np.ascontiguousarray(x, np.float32)
        off = 0
        for k, w in (("BODY_POS", 3), ("BODY_QUAT", 4), ("BODY_VEL", 3), ("BODY_OMEGA", 3)):
            self.ctx.upload(k, x[off:off + w * nb])
            off += w * nb
