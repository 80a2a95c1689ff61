"""The tracer-side BVH contract on the CPU: the restated closest-hit query of
raytrace.comp:86-133 over the oracle's LBVH (bvh.comp) returns the
brute-force closest hit; planted layout mistakes (a swapped child, an off-by-one
prim_idx, leaf boxes missing a vertex) make it fail."""
import ctypes as C

import numpy as np
import pytest

from harness import OrBvh, make_scene, oracle, ptr
from trace_oracle import check_hits, ray_tri, rays_at, trace


def oracle_bvh(verts, idx, n):
    m = 2 * n - 1
    out = dict(centroids=np.zeros(3 * n, np.float32), morton=np.zeros(n, np.uint32),
               sort_idx=np.zeros(n, np.uint32), bounds=np.zeros(6, np.float32),
               node_min=np.zeros(3 * m, np.float32), node_max=np.zeros(3 * m, np.float32),
               node_left=np.zeros(m, np.uint32), node_right=np.zeros(m, np.uint32),
               node_count=np.zeros(m, np.uint32), node_parent=np.zeros(m, np.uint32),
               prim_idx=np.zeros(n, np.uint32), flags=np.zeros(n, np.uint32))
    t = OrBvh(vertices=ptr(verts), indices=ptr(idx), **{k: ptr(v) for k, v in out.items()})
    oracle().bvh_build(C.byref(t), n)
    return out


def test_ray_tri_closed_forms():
    v0, v1, v2 = (0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (0.0, 0.0, 1.0)
    assert ray_tri((0.25, 2.0, 0.25), (0.0, -1.0, 0.0), v0, v1, v2) == 2.0
    assert ray_tri((0.25, -3.0, 0.25), (0.0, 1.0, 0.0), v0, v1, v2) == 3.0  # back face hits too
    assert ray_tri((0.75, 2.0, 0.75), (0.0, -1.0, 0.0), v0, v1, v2) == -1.0  # u + v > 1
    assert ray_tri((0.25, 2.0, 0.25), (1.0, 0.0, 0.0), v0, v1, v2) == -1.0  # parallel
    assert ray_tri((0.25, 0.0005, 0.25), (0.0, -1.0, 0.0), v0, v1, v2) == -1.0  # t below 0.001


@pytest.fixture(scope="module")
def scene():
    sc = make_scene(1, 60, seed=5)
    n = sc.counts()["num_triangles"]
    verts, idx = sc.array("VERTICES").copy(), sc.array("INDICES").copy()
    return verts, idx, n, oracle_bvh(verts, idx, n)


def test_traversal_equals_brute_force(scene):
    verts, idx, n, bvh = scene
    ro, rd = rays_at(verts, idx, 3000, seed=1)
    hits, dropped = check_hits(verts, idx, bvh, ro, rd, n)
    assert hits > 2000 and dropped == 0


@pytest.mark.parametrize("plant", ["swap_child", "prim_idx", "leaf_box"])
def test_planted_layout_mistakes_fail(scene, plant):
    verts, idx, n, bvh = scene
    ro, rd = rays_at(verts, idx, 3000, seed=2)
    bad = {k: v.copy() for k, v in bvh.items()}
    if plant == "swap_child":
        # root's children exchanged with a grandchild's: subtrees under the wrong boxes
        r = int(bad["node_left"][0])
        bad["node_left"][0], bad["node_left"][r] = bad["node_left"][r], bad["node_left"][0]
    elif plant == "prim_idx":
        # off by one in the sorted-position -> triangle map
        bad["prim_idx"] = np.roll(bad["prim_idx"], 1)
    else:
        # leaf boxes from the first two vertices only
        tri = verts.reshape(-1, 3)[idx.reshape(-1, 3)[bad["prim_idx"]]][:, :2]
        bad["node_min"][3 * (n - 1):] = tri.min(1).reshape(-1)
        bad["node_max"][3 * (n - 1):] = tri.max(1).reshape(-1)
    with pytest.raises(AssertionError):
        check_hits(verts, idx, bad, ro, rd, n)


def test_deep_tree_drops_are_reported():
    # Centroids whose Morton keys peel off one at a time from the left (each
    # key one bit short of the top corner) over a cluster of 64 equal keys
    # split by index: a right spine 37 deep.  A ray into the cluster passes
    # every spine box, so the left siblings pile up past the tracer's
    # 31-entry guard (raytrace.comp:127-128) and pushes are dropped.
    cents = []
    for k in range(10):
        q = 1023 - 2 ** (9 - k)
        for ax in range(3):
            c = [1023.5] * 3
            c[ax] = q + 0.5
            cents.append(c)
    cents = np.array(cents + [[1023.5] * 3] * 64 + [[0.5] * 3]) / 1024.0
    k, s = len(cents), 1e-4
    tri = np.zeros((k, 3, 3))
    tri[:, 0], tri[:, 1], tri[:, 2] = cents + [-s, 0, 0], cents + [s, 0, -s], cents + [0, 0, s]
    verts = tri.astype(np.float32).reshape(-1)
    idx = np.arange(3 * k, dtype=np.uint32)
    bvh = oracle_bvh(verts, idx, k)
    par = bvh["node_parent"]
    depth = np.zeros(2 * k - 1, np.int64)
    for v in range(1, 2 * k - 1):
        d, u = 0, v
        while u != 0:
            u, d = int(par[u]), d + 1
        depth[v] = d
    assert depth.max() > 31
    ro = np.array([[1023.5 / 1024, 2.0, 1023.5 / 1024], [0.5 / 1024, 2.0, 0.5 / 1024]], np.float32)
    rd = np.array([[0.0, -1.0, 0.0]] * 2, np.float32)
    t, hit, dropped, _ = trace(verts, idx, bvh, ro, rd)
    assert dropped.tolist() == [True, False]
    hits, nd = check_hits(verts, idx, bvh, ro, rd, k)
    assert hits == 2 and nd == 1
