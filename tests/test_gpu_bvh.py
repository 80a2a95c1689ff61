"""LBVH parity (triangles and body AABBs) of libseevr_hip.so with the oracle:
integer outputs (Morton codes, sorted order, topology) bit-exact, node boxes
bit-exact (min/max are exact)."""
import ctypes as C

import numpy as np
import pytest

from harness import OrBvh, assert_bitwise, make_scene, oracle, ptr

pytestmark = pytest.mark.gpu


def oracle_bvh(verts, idx, n, bodies=None):
    m = 2 * n - 1
    out = dict(centroids=np.zeros(3 * n, np.float32), morton=np.zeros(n, np.uint32),
               sort_idx=np.zeros(n, np.uint32), bounds=np.zeros(6, np.float32),
               node_min=np.zeros(3 * m, np.float32), node_max=np.zeros(3 * m, np.float32),
               node_left=np.zeros(m, np.uint32), node_right=np.zeros(m, np.uint32),
               node_count=np.zeros(m, np.uint32), node_parent=np.zeros(m, np.uint32),
               prim_idx=np.zeros(n, np.uint32), flags=np.zeros(n, np.uint32))
    t = OrBvh(vertices=ptr(verts), indices=ptr(idx), **{k: ptr(v) for k, v in out.items()})
    lib = oracle()
    if bodies is None:
        lib.bvh_build(C.byref(t), n)
    else:
        lib.bvh_build_bodies(C.byref(t), ptr(bodies[0]), ptr(bodies[1]), n)
    return out


def device_bvh(ctx, n):
    m = 2 * n - 1
    return dict(centroids=ctx.download("BVH_CENTROIDS", 3 * n), morton=ctx.download("BVH_MORTON", n),
                sort_idx=ctx.download("BVH_SORT_INDICES", n), bounds=ctx.download("BVH_SCENE_BOUNDS", 6),
                node_min=ctx.download("BVH_AABB_MIN", 3 * m), node_max=ctx.download("BVH_AABB_MAX", 3 * m),
                node_left=ctx.download("BVH_LEFT", m), node_right=ctx.download("BVH_RIGHT", m),
                node_count=ctx.download("BVH_COUNT", m), node_parent=ctx.download("BVH_PARENT", m),
                prim_idx=ctx.download("BVH_PRIM_INDICES", n))


def compare(got, want, n):
    for k in ("centroids", "bounds", "morton", "sort_idx", "prim_idx", "node_left", "node_count",
              "node_parent", "node_min", "node_max"):
        assert_bitwise(got[k], want[k], k)
    # node_right of leaves is never written by the reference (bvh.comp:304-308)
    assert_bitwise(got["node_right"][: n - 1], want["node_right"][: n - 1], "node_right internal")


@pytest.mark.parametrize("bodies", [6, 40])
def test_triangle_bvh_bitwise(bodies):
    from seevr_amd import Context, caps_for
    sc = make_scene(1, bodies, seed=11)
    ctx = Context(caps_for(sc, 64, 64))
    sc.upload(ctx)
    n = sc.counts()["num_triangles"]
    ctx.bvh_build(n)
    verts, idx = sc.array("VERTICES"), sc.array("INDICES")
    compare(device_bvh(ctx, n), oracle_bvh(verts, idx, n), n)
    ctx.close()


@pytest.mark.parametrize("n", [1, 2, 3, 257, 5000])
def test_body_bvh_bitwise_with_duplicates(n):
    from seevr_amd import Context, default_caps
    rng = np.random.default_rng(n)
    # quantised centres make many equal Morton codes (index tie-break path)
    c = np.round(rng.uniform(-5, 5, (n, 3)) * 2) / 2
    h = rng.uniform(0.1, 0.5, (n, 3))
    amin = (c - h).astype(np.float32).ravel()
    amax = (c + h).astype(np.float32).ravel()
    ctx = Context(default_caps(max_bodies=n, max_hull_verts=8, max_vertices=8, max_triangles=1))
    ctx.upload("BODY_AABB_MIN", amin)
    ctx.upload("BODY_AABB_MAX", amax)
    ctx.bvh_build_bodies(n)
    got = device_bvh(ctx, n)
    want = oracle_bvh(None, None, n, bodies=(amin, amax))
    compare(got, want, n)
    ctx.close()
