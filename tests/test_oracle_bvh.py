"""Pins for the oracle's LBVH (bvh.comp restatement): the radix sort against
numpy's stable argsort, the Karras hierarchy against a brute-force top-down
construction over index-augmented keys, and the refit against the union of
each subtree's leaves."""
import ctypes as C

import numpy as np
import pytest

from harness import OrBvh, oracle, ptr


def run(n, keys=None, verts=None, idx=None, bodies=None):
    m = 2 * n - 1
    out = dict(centroids=np.zeros(3 * n, np.float32), morton=np.zeros(n, np.uint32),
               sort_idx=np.zeros(n, np.uint32), bounds=np.zeros(6, np.float32),
               node_min=np.zeros(3 * m, np.float32), node_max=np.zeros(3 * m, np.float32),
               node_left=np.zeros(m, np.uint32), node_right=np.zeros(m, np.uint32),
               node_count=np.zeros(m, np.uint32), node_parent=np.zeros(m, np.uint32),
               prim_idx=np.zeros(n, np.uint32), flags=np.zeros(n, np.uint32))
    t = OrBvh(vertices=ptr(verts), indices=ptr(idx), **{k: ptr(v) for k, v in out.items()})
    lib = oracle()
    if keys is not None:
        out["morton"][:] = keys
        out["sort_idx"][:] = np.arange(n, dtype=np.uint32)
        lib.bvh_radix_sort(C.byref(t), n)
        lib.bvh_hierarchy(C.byref(t), n)
    elif bodies is not None:
        lib.bvh_build_bodies(C.byref(t), ptr(bodies[0]), ptr(bodies[1]), n)
    else:
        lib.bvh_build(C.byref(t), n)
    return out


@pytest.mark.parametrize("n,span", [(1, 10), (7, 4), (1000, 1 << 30), (5000, 64)])
def test_radix_sort_is_stable_sort(n, span):
    rng = np.random.default_rng(n)
    keys = rng.integers(0, span, n, dtype=np.uint64).astype(np.uint32)
    if n > 1:
        keys[-1] = 0xFFFFFFFF  # full 32-bit key
    out = run(n, keys=keys)
    order = np.argsort(keys, kind="stable")
    assert (out["sort_idx"] == order).all()
    assert (out["morton"] == keys[order]).all()


def karras_brute(sorted_keys):
    """Top-down construction: node range [f, l] splits after the last index s
    whose augmented key (key, index) shares more leading bits with f than l
    does.  Internal node ids follow Karras: the root is 0 and each child that
    is a range [s+1, l] or [f, s] takes the id of its end at the split."""
    n = len(sorted_keys)
    aug = [(int(k) << 32) | i for i, k in enumerate(sorted_keys)]

    def prefix(a, b):
        x = aug[a] ^ aug[b]
        return 64 - x.bit_length()

    left, right = {}, {}
    stack = [(0, 0, n - 1)]
    while stack:
        node, f, l = stack.pop()
        dn = prefix(f, l)
        s = max(k for k in range(f, l) if prefix(f, k) > dn) if l > f + 1 else f
        lc = s + n - 1 if s == f else s
        rc = s + 1 + n - 1 if s + 1 == l else s + 1
        left[node], right[node] = lc, rc
        if s != f:
            stack.append((s, f, s))
        if s + 1 != l:
            stack.append((s + 1, s + 1, l))
    return left, right


@pytest.mark.parametrize("n,span", [(2, 4), (3, 2), (17, 8), (300, 1 << 30), (600, 16)])
def test_karras_matches_top_down_construction(n, span):
    rng = np.random.default_rng(n + span)
    keys = np.sort(rng.integers(0, span, n, dtype=np.uint64).astype(np.uint32))
    out = run(n, keys=keys)
    left, right = karras_brute(keys)
    assert len(left) == n - 1
    for i in range(n - 1):
        assert out["node_left"][i] == left[i] and out["node_right"][i] == right[i], i
        assert out["node_count"][i] == 0
        assert out["node_parent"][left[i]] == i and out["node_parent"][right[i]] == i
    assert out["node_parent"][0] == 0xFFFFFFFF
    for i in range(n):
        assert out["node_left"][n - 1 + i] == i and out["node_count"][n - 1 + i] == 1


def subtree_leaves(out, node, n):
    if node >= n - 1:
        return [node - (n - 1)]
    return subtree_leaves(out, int(out["node_left"][node]), n) + subtree_leaves(out, int(out["node_right"][node]), n)


def test_triangle_build_refit_and_morton():
    rng = np.random.default_rng(12)
    nt = 400
    verts = rng.uniform(-3, 7, (3 * nt, 3)).astype(np.float32)
    idx = rng.permutation(3 * nt).astype(np.uint32)
    out = run(nt, verts=verts.ravel(), idx=idx)
    tri = verts[idx.reshape(-1, 3)]
    cent = ((tri[:, 0] + tri[:, 1]) + tri[:, 2]) / np.float32(3.0)
    assert (out["centroids"].reshape(-1, 3) == cent).all()
    assert (out["bounds"] == np.concatenate([cent.min(0), cent.max(0)])).all()
    # Morton codes: quantise each normalised centroid to 10 bits and interleave
    bmin, ext = cent.min(0), cent.max(0) - cent.min(0)
    q = np.clip((cent - bmin) * (np.float32(1) / ext) * np.float32(1024), 0, 1023).astype(np.uint64)
    code = np.zeros(nt, np.uint64)
    for b in range(10):
        for a in range(3):
            code |= ((q[:, a] >> b) & 1) << (3 * b + 2 - a)
    order = np.argsort(code.astype(np.uint32), kind="stable")
    assert (out["sort_idx"] == order).all()
    assert (out["morton"] == code[order]).all()
    assert (out["prim_idx"] == out["sort_idx"]).all()
    tmin, tmax = tri.min(1), tri.max(1)
    nmin = out["node_min"].reshape(-1, 3)
    nmax = out["node_max"].reshape(-1, 3)
    for node in list(range(nt - 1)) + [nt - 1 + 5, 2 * nt - 2]:
        leaves = subtree_leaves(out, node, nt)
        prims = out["prim_idx"][leaves]
        assert (nmin[node] == tmin[prims].min(0)).all(), node
        assert (nmax[node] == tmax[prims].max(0)).all(), node


def test_body_build_single_and_flat():
    out = run(1, bodies=(np.array([0, 0, 0], np.float32), np.array([1, 1, 1], np.float32)))
    assert out["node_parent"][0] == 0xFFFFFFFF and out["node_count"][0] == 1
    assert (out["node_min"][:3] == 0).all() and (out["node_max"][:3] == 1).all()
    # all centres on one plane: the zero extent axis gets inverse extent 0
    n = 64
    c = np.stack([np.arange(n) % 8, np.zeros(n), np.arange(n) // 8], 1).astype(np.float32)
    out = run(n, bodies=((c - 0.25).ravel(), (c + 0.25).ravel()))
    assert out["bounds"][1] == out["bounds"][4] == 0
    assert len(set(out["morton"].tolist())) == n
