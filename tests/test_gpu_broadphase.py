"""LBVH-traversal and sweep-and-prune broadphases against the all-pairs loop
of physics.comp:451-472.

Every device method must give the oracle's pair list bit for bit: the same
set, in lexicographic (i, j) order, the first max_pairs of them kept and
counters[0] holding the total.  Boxes are placed on an integer grid so that
many of them touch exactly (touching counts as overlap), some are
duplicates and about a quarter are static (static-static pairs are skipped).
"""
import numpy as np
import pytest

from harness import World, assert_bitwise, device_pairs, make_scene

pytestmark = pytest.mark.gpu


def box_world(n, seed, max_pairs, static_frac=0.25, spread=None):
    """A World of n bodies whose AABBs are overwritten with random grid boxes."""
    sc = make_scene(1, max(n - 1, 1), seed=seed, segments=4)
    w = World.from_scene(sc, max_pairs=max_pairs, max_contacts=16)
    assert w.n == n
    rng = np.random.default_rng(seed)
    spread = 1.6 * n ** (1.0 / 3.0) if spread is None else spread
    lo = np.floor(rng.uniform(0, spread, (n, 3))).astype(np.float32)
    ext = rng.integers(1, 3, (n, 3)).astype(np.float32)
    dup = rng.random(n) < 0.05
    src = rng.integers(0, n, n)
    lo[dup] = lo[src[dup]]
    ext[dup] = ext[src[dup]]
    w.a["BODY_AABB_MIN"][:] = lo.ravel()
    w.a["BODY_AABB_MAX"][:] = (lo + ext).ravel()
    inv_mass = w.a["BODY_INV_MASS"]
    inv_mass[:] = np.where(rng.random(n) < static_frac, 0.0, 1.0).astype(np.float32)
    return sc, w


def device_broadphase(sc, w, method):
    from seevr_amd import Context, caps_for, default_params
    ctx = Context(caps_for(sc, w.max_pairs, w.max_contacts), default_params())
    sc.upload(ctx)
    ctx.set_broadphase(method)
    w.upload(ctx)
    ctx.dispatch(8, 1)
    ctx.dispatch(2, w.n)
    cnt = ctx.download("ATOMIC_COUNTERS", 3)
    pairs = device_pairs(ctx, min(int(cnt[0]), w.max_pairs))
    ctx.close()
    return cnt, pairs


def oracle_broadphase(w):
    w.a["ATOMIC_COUNTERS"][:] = 0
    w.oracle_stage(8)
    w.oracle_stage(2)
    return w.a["ATOMIC_COUNTERS"].copy(), w.pairs()


@pytest.mark.parametrize("method", ["all_pairs", "lbvh", "sweep"])
@pytest.mark.parametrize("n,seed", [(2, 1), (3, 2), (64, 3), (700, 4), (3000, 5)])
def test_pairs_match_oracle(n, seed, method):
    sc, w = box_world(n, seed, max_pairs=400000)
    want_cnt, want = oracle_broadphase(w)
    got_cnt, got = device_broadphase(sc, w, method)
    assert int(got_cnt[0]) == int(want_cnt[0])
    if n >= 64:
        assert int(want_cnt[0]) > n // 2, "scene should be dense"
    assert_bitwise(got, want, f"{method} pairs n={n}")


@pytest.mark.parametrize("method", ["all_pairs", "lbvh", "sweep"])
@pytest.mark.parametrize("cap", [1, 37, 500])
def test_truncation_keeps_lexicographic_prefix(cap, method):
    """A row straddles the cap: only its smallest partners are kept."""
    sc, w = box_world(300, 11, max_pairs=cap)
    want_cnt, want = oracle_broadphase(w)
    assert int(want_cnt[0]) > cap
    got_cnt, got = device_broadphase(sc, w, method)
    assert int(got_cnt[0]) == int(want_cnt[0])
    assert int(got_cnt[2]) & 1, "pair overflow flag"
    assert_bitwise(got, want, f"{method} truncated pairs cap={cap}")


@pytest.mark.parametrize("method", ["lbvh", "sweep"])
def test_all_static_and_two_bodies(method):
    sc, w = box_world(200, 12, max_pairs=1000, static_frac=1.0)
    got_cnt, got = device_broadphase(sc, w, method)
    assert int(got_cnt[0]) == 0 and got.size == 0
    sc1 = make_scene(1, 1, seed=3, segments=4)
    w1 = World.from_scene(sc1, max_pairs=100, max_contacts=16)
    want_cnt, want = oracle_broadphase(w1.copy())
    got_cnt, got = device_broadphase(sc1, w1, method)
    assert int(got_cnt[0]) == int(want_cnt[0])
    assert_bitwise(got, want, "two-body scene")


def wide_world(n, seed, max_pairs, negative_zero=False):
    """box_world with a ground-like static slab and a few wide bodies (body
    indices 0 and 1..4, their x ranges over the whole scene) and a dense
    column of 400 bodies sharing one x range: runs of thousands of
    candidates, continued past the first wave's span."""
    sc, w = box_world(n, seed, max_pairs=max_pairs)
    lo = w.a["BODY_AABB_MIN"].reshape(-1, 3)
    hi = w.a["BODY_AABB_MAX"].reshape(-1, 3)
    span = float(hi[:, 0].max())
    lo[0] = (-1.0, -1.0, -1.0)
    hi[0] = (span + 1.0, 0.0, span + 1.0)
    w.a["BODY_INV_MASS"][0] = 0.0
    for k in range(1, 5):
        lo[k, 0], hi[k, 0] = -0.5, span + 0.5
    col = np.arange(100, n, max(1, (n - 100) // 400))[:400]
    lo[col, 0], hi[col, 0] = 3.0, 4.0
    if negative_zero:  # touching at x = 0 written as -0 and +0
        lo[col[:50], 0], hi[col[:50], 0] = -1.0, -0.0
        lo[col[50:100], 0], hi[col[50:100], 0] = 0.0, 1.0
    return sc, w


@pytest.mark.parametrize("negative_zero", [False, True])
@pytest.mark.parametrize("cap", [2_000_000, 5000])
def test_sweep_long_runs_and_touching_zero(negative_zero, cap):
    sc, w = wide_world(6000, 41, max_pairs=cap, negative_zero=negative_zero)
    want_cnt, want = oracle_broadphase(w)
    got_cnt, got = device_broadphase(sc, w, "sweep")
    assert int(got_cnt[0]) == int(want_cnt[0]) > 20000
    assert_bitwise(got, want, f"sweep pairs with wide bodies, cap={cap}, -0={negative_zero}")


def clump_world(max_pairs):
    """box_world(3000) with 60 dynamic bodies (indices from 1000) sharing one box."""
    sc, w = box_world(3000, 31, max_pairs=max_pairs)
    lo = w.a["BODY_AABB_MIN"].reshape(-1, 3)
    hi = w.a["BODY_AABB_MAX"].reshape(-1, 3)
    clump = np.arange(1000, 3000, 2000 // 60)[:60]
    lo[clump] = lo[clump[0]]
    hi[clump] = hi[clump[0]]
    w.a["BODY_INV_MASS"][clump] = 1.0
    return sc, w, clump


def test_sweep_keys_from_the_previous_launch():
    """The sweep keys a launch with a map from the previous launch's lower x
    bounds: launches on one context whose boxes move outside that range
    (clamped keys) and back must still give the oracle's list."""
    from seevr_amd import Context, caps_for, default_params
    sc, w = wide_world(6000, 43, max_pairs=2_000_000)
    ctx = Context(caps_for(sc, w.max_pairs, w.max_contacts), default_params())
    sc.upload(ctx)
    ctx.set_broadphase("sweep")
    base_lo = w.a["BODY_AABB_MIN"].copy()
    base_hi = w.a["BODY_AABB_MAX"].copy()
    rng = np.random.default_rng(5)
    for k, (shift, scale) in enumerate([(0.0, 1.0), (0.0, 1.0), (-7.0, 1.0), (3.0, 2.5), (0.0, 0.25), (0.0, 1.0)]):
        lo = base_lo.reshape(-1, 3).copy()
        hi = base_hi.reshape(-1, 3).copy()
        jitter = rng.integers(-1, 2, len(lo)).astype(np.float32)
        lo[:, 0] = (lo[:, 0] + jitter) * np.float32(scale) + np.float32(shift)
        hi[:, 0] = (hi[:, 0] + jitter) * np.float32(scale) + np.float32(shift)
        w.a["BODY_AABB_MIN"][:] = lo.ravel()
        w.a["BODY_AABB_MAX"][:] = hi.ravel()
        want_cnt, want = oracle_broadphase(w)
        w.upload(ctx)
        ctx.dispatch(8, 1)
        ctx.dispatch(2, w.n)
        cnt = ctx.download("ATOMIC_COUNTERS", 3)
        got = device_pairs(ctx, min(int(cnt[0]), w.max_pairs))
        assert int(cnt[0]) == int(want_cnt[0])
        assert_bitwise(got, want, f"sweep launch {k} (shift {shift}, scale {scale})")
    ctx.close()


@pytest.mark.parametrize("cut", [False, True])
def test_lbvh_rows_beyond_staging(cut):
    """Lanes that find more rows than the counting walk stages walk again.

    The clump's body with the largest index finds 59 rows; most other bodies
    find few.  With `cut`, the cap ends the list inside the clump's rows.
    """
    _, wf, clump = clump_world(400000)
    _, full = oracle_broadphase(wf)
    per_j = np.bincount(full[:, 1], minlength=wf.n)
    assert per_j.max() >= 59 and (per_j[per_j > 0] <= 32).mean() > 0.5 and (per_j > 32).sum() > 0
    cap = 400000
    if cut:
        cap = int(np.searchsorted(full[:, 0], clump[len(clump) // 2])) + 7
    sc, w, _ = clump_world(cap)
    want_cnt, want = oracle_broadphase(w)
    got_cnt, got = device_broadphase(sc, w, "lbvh")
    assert int(got_cnt[0]) == int(want_cnt[0])
    assert_bitwise(got, want, f"lbvh pairs with a clump, cap={cap}")


def test_lbvh_and_sweep_equal_all_pairs_at_scale():
    """40k bodies: the three device methods agree (the oracle pins them above)."""
    sc, w = box_world(40000, 21, max_pairs=2_000_000)
    a_cnt, a = device_broadphase(sc, w, "all_pairs")
    b_cnt, b = device_broadphase(sc, w, "lbvh")
    c_cnt, c = device_broadphase(sc, w, "sweep")
    assert int(a_cnt[0]) == int(b_cnt[0]) == int(c_cnt[0]) > 40000
    assert_bitwise(b, a, "lbvh vs all-pairs at 40k")
    assert_bitwise(c, a, "sweep vs all-pairs at 40k")
