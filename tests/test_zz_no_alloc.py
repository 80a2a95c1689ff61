"""sv_bvh_build_bodies_noalloc (the allocation-free body LBVH build, whose
scratch is a static device array) against sv_bvh_build_bodies on a context:
every output bit for bit, on two calls in a row, the second with fewer
bodies (the first call's scratch contents must not leak into it), on the
null stream and on a caller stream; and the size limit."""
import ctypes as C

import numpy as np
import pytest

from harness import HipMem, assert_bitwise, make_scene

pytestmark = pytest.mark.gpu

FIELDS = ("centroids", "morton", "sort_indices", "prim_indices", "scene_bounds", "node_min", "node_max",
          "left", "right", "count", "parent")
BUF_OF = dict(centroids="BVH_CENTROIDS", morton="BVH_MORTON", sort_indices="BVH_SORT_INDICES",
              prim_indices="BVH_PRIM_INDICES", scene_bounds="BVH_SCENE_BOUNDS", node_min="BVH_AABB_MIN",
              node_max="BVH_AABB_MAX", left="BVH_LEFT", right="BVH_RIGHT", count="BVH_COUNT", parent="BVH_PARENT")


def sizes(n):
    m = 2 * n - 1
    return dict(centroids=3 * n, morton=n, sort_indices=n, prim_indices=n, scene_bounds=6, node_min=3 * m,
                node_max=3 * m, left=m, right=m, count=m, parent=m)


def dtype(f):
    return np.float32 if f in ("centroids", "scene_bounds", "node_min", "node_max") else np.uint32


def settled_boxes(n, seed):
    """Body boxes of a seeded pile after a few frames on the device."""
    from seevr_amd import Context, caps_for, default_params
    sc = make_scene(1, n, seed=seed)
    cnt = sc.counts()
    ctx = Context(caps_for(sc, 16 * n + 1024, 8 * n + 1024), default_params(pgs_iterations=30))
    sc.upload(ctx)
    for _ in range(30):
        ctx.step(cnt["num_bodies"], cnt["num_vertices"], cnt["num_hull_verts"], 10, 1 / 60, (0.0, -9.81, 0.0))
    return sc, ctx, cnt["num_bodies"]


def reference_build(ctx, n):
    """The existing entry point: sv_bvh_build_bodies on the context's boxes."""
    ctx.bvh_build_bodies(n)
    return {f: ctx.download(BUF_OF[f], sizes(n)[f]) for f in FIELDS}


def noalloc_build(hm, amin, amax, n, stream=None):
    from seevr_amd import _lib
    lib = _lib.load()
    dev_min = hm.alloc(amin.nbytes)
    dev_max = hm.alloc(amax.nbytes)
    hm.upload(dev_min, amin)
    hm.upload(dev_max, amax)
    out = {f: hm.alloc(4 * sizes(n)[f], fill=-7) for f in FIELDS}
    arr = _lib.BvhArrays(**out)
    rc = lib.sv_bvh_build_bodies_noalloc(C.c_void_p(dev_min), C.c_void_p(dev_max), n, C.byref(arr),
                                         C.c_void_p(stream))
    _lib.check(rc)
    return {f: hm.download(out[f], sizes(n)[f], dtype(f)) for f in FIELDS}


def compare(got, want, n, what):
    for f in FIELDS:
        g, w = got[f], want[f]
        if f == "right":  # leaves' right child is never written (bvh.comp:304-308)
            g, w = g[: n - 1], w[: n - 1]
        assert_bitwise(g, w, f"{what}: {f}")


def test_noalloc_equals_context_build_twice_shrinking():
    big, small = 4000, 900
    sc1, ctx1, n1 = settled_boxes(big, seed=21)
    sc2, ctx2, n2 = settled_boxes(small, seed=22)
    hm = HipMem()
    try:
        for ctx, n, what in ((ctx1, n1, "first call"), (ctx2, n2, "second call, fewer bodies")):
            amin = ctx.download("BODY_AABB_MIN", 3 * n)
            amax = ctx.download("BODY_AABB_MAX", 3 * n)
            want = reference_build(ctx, n)
            compare(noalloc_build(hm, amin, amax, n), want, n, what)
        # a third call on a caller stream, after the scratch held the larger build
        amin = ctx2.download("BODY_AABB_MIN", 3 * n2)
        amax = ctx2.download("BODY_AABB_MAX", 3 * n2)
        compare(noalloc_build(hm, amin, amax, n2, stream=hm.stream()), reference_build(ctx2, n2), n2,
                "caller stream")
    finally:
        ctx1.close()
        ctx2.close()
        hm.close()


def test_noalloc_limits():
    from seevr_amd import _lib
    lib = _lib.load()
    arr = _lib.BvhArrays()
    rc = lib.sv_bvh_build_bodies_noalloc(None, None, _lib.NOALLOC_MAX_BODIES + 1, C.byref(arr), None)
    assert rc == -2 and b"exceed" in lib.sv_last_error()
    assert lib.sv_bvh_build_bodies_noalloc(None, None, 10, C.byref(arr), None) == -1  # null arrays
    assert lib.sv_bvh_build_bodies_noalloc(None, None, 0, C.byref(arr), None) == 0    # nothing to build
