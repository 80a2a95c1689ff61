"""ctypes binding of the CPU restatement of the reference tracer's
closest-hit query (oracle/sv_trace.c, raytrace.comp:56-133), and seeded
ray sets aimed at a scene.  Test infrastructure only."""
from __future__ import annotations

import ctypes as C

import numpy as np

from harness import ORACLE_DIR, build_oracle, ptr

LIB = ORACLE_DIR / "_build" / "libsvtrace.so"
_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build_oracle()
        _lib = C.CDLL(str(LIB))
        vp, u32 = C.c_void_p, C.c_uint32
        _lib.ort_trace.argtypes = [vp] * 8 + [u32, vp, vp, vp, vp, vp, vp]
        _lib.ort_trace.restype = None
        _lib.ort_brute.argtypes = [vp, vp, u32, u32, vp, vp, vp, vp]
        _lib.ort_brute.restype = None
        _lib.ort_ray_tri.argtypes = [vp, vp, vp, vp, vp]
        _lib.ort_ray_tri.restype = C.c_float
    return _lib


def trace(verts, idx, bvh, ro, rd):
    """Closest hit through BVH arrays (dict with node_min/max, node_left/right,
    node_count, prim_idx).  Returns (t, tri, dropped, node tests)."""
    n = ro.shape[0]
    t = np.zeros(n, np.float32)
    tri = np.zeros(n, np.uint32)
    dropped = np.zeros(n, np.uint32)
    visited = C.c_uint64(0)
    arrs = [np.ascontiguousarray(bvh[k]) for k in ("node_min", "node_max", "node_left", "node_right",
                                                    "node_count", "prim_idx")]
    lib().ort_trace(ptr(verts), ptr(idx), *[ptr(a) for a in arrs], n, ptr(ro), ptr(rd), ptr(t), ptr(tri),
                    ptr(dropped), C.byref(visited))
    return t, tri, dropped.astype(bool), visited.value


def brute(verts, idx, n_tris, ro, rd):
    n = ro.shape[0]
    t = np.zeros(n, np.float32)
    tri = np.zeros(n, np.uint32)
    lib().ort_brute(ptr(verts), ptr(idx), n_tris, n, ptr(ro), ptr(rd), ptr(t), ptr(tri))
    return t, tri


def ray_tri(ro, rd, v0, v1, v2) -> float:
    a = [np.ascontiguousarray(x, np.float32) for x in (ro, rd, v0, v1, v2)]
    return float(lib().ort_ray_tri(*[ptr(x) for x in a]))


def rays_at(verts, idx, n_rays, seed, miss_frac=0.2):
    """Rays from points on a sphere of radius 2x the scene's extent around its
    centre, aimed at random triangle centroids (hits) or in random directions
    (mostly misses); directions normalised in float32."""
    rng = np.random.default_rng(seed)
    v = verts.reshape(-1, 3).astype(np.float64)
    lo, hi = v.min(0), v.max(0)
    ctr, rad = (lo + hi) / 2, float(np.linalg.norm(hi - lo)) + 1.0
    d = rng.normal(size=(n_rays, 3))
    ro = ctr + 2.0 * rad * d / np.linalg.norm(d, axis=1, keepdims=True)
    tris = idx.reshape(-1, 3)[rng.integers(0, idx.size // 3, n_rays)]
    w = rng.dirichlet((1.0, 1.0, 1.0), n_rays)
    tgt = (v[tris] * w[:, :, None]).sum(1)
    rd = tgt - ro
    k = rng.random(n_rays) < miss_frac
    rd[k] = rng.normal(size=(int(k.sum()), 3))
    rd /= np.linalg.norm(rd, axis=1, keepdims=True)
    return np.ascontiguousarray(ro, np.float32), np.ascontiguousarray(rd, np.float32)


def check_hits(verts, idx, bvh, ro, rd, n_tris, tie_rtol=1e-6, tie_frac=0.01):
    """Traversal hits equal brute-force closest hits for every ray the stack
    guard dropped nothing on, except near-ties: a hit at a box boundary that
    the tracer's float slab test (raytrace.comp:74-82) rounds outside, so the
    traversal returns the next hit within tie_rtol of it.  Those may be at
    most tie_frac of the rays.  Exact ties (equal t) may name either
    triangle, as long as the traversal's triangle gives that t.  Returns
    (hit count, dropped count)."""
    t, tri, dropped, _ = trace(verts, idx, bvh, ro, rd)
    bt, btri = brute(verts, idx, n_tris, ro, rd)
    ok = ~dropped
    # a dropped subtree or a rounded-out box can only lose hits, never invent a closer one
    assert np.all(t >= bt), "traversal found a hit closer than every triangle"
    differ = ok & (t != bt)
    near = differ & (bt < 1e10) & (t - bt <= tie_rtol * bt)
    assert not np.any(differ & ~near), f"closest t differs from brute force on {int((differ & ~near).sum())} rays"
    assert near.sum() <= tie_frac * len(t), f"{int(near.sum())} near-ties"
    same = ok & ~differ & (tri != btri)
    for r in np.nonzero(same)[0]:
        i0, i1, i2 = idx[3 * tri[r]: 3 * tri[r] + 3]
        tt = ray_tri(ro[r], rd[r], verts[3 * i0: 3 * i0 + 3], verts[3 * i1: 3 * i1 + 3], verts[3 * i2: 3 * i2 + 3])
        assert np.float32(tt) == t[r], f"ray {r}: triangle {tri[r]} does not give the closest t"
    return int((bt < 1e10).sum()), int(dropped.sum())
