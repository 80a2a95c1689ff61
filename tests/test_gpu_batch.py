"""Batched engine against the plain path: every scene of a batch must equal
the same scene stepped alone by sv_physics_step, bit for bit (state, pair
list, contacts, counters), with and without the captured hipGraph, after the
autotuner, and with per-scene overflow.  The plain path itself is pinned to
the oracle by test_gpu_parity.py.
"""
import numpy as np
import pytest

from harness import assert_bitwise

pytestmark = pytest.mark.gpu

SEG = 8          # 81-vertex spheres: the support-lookup path
PAIRS, CONTACTS = 4096, 2048
STATE = (("BODY_POS", 3), ("BODY_QUAT", 4), ("BODY_VEL", 3), ("BODY_OMEGA", 3))


def build_batch(n_scenes, seed=3, ps=PAIRS, cs=CONTACTS):
    from seevr_amd import Batch, Scene, default_caps, default_params
    from seevr_amd._lib import BatchCaps
    big = default_caps(max_bodies=n_scenes * 257 + 16, max_hull_verts=n_scenes * 257 * 82,
                       max_vertices=n_scenes * 257 * 82, max_triangles=n_scenes * 257 * 130,
                       max_support_entries=n_scenes * 257 * 193 + 1)
    sc = Scene(big)
    start, gs = sc.build_batch(0, n_scenes, 8, 256, SEG, seed)
    cnt = sc.counts()
    bc = BatchCaps(max_scenes=n_scenes, max_bodies=cnt["num_bodies"], max_hull_verts=cnt["num_hull_verts"],
                   max_vertices=cnt["num_vertices"], max_triangles=max(cnt["num_triangles"], 1),
                   max_support_entries=cnt["num_support_entries"],
                   max_scene_bodies=512, max_scene_pairs=ps, max_scene_contacts=cs)
    b = Batch(bc, default_params(pgs_iterations=30))
    sc.upload(b.ctx)
    grav = np.zeros((n_scenes, 3), np.float32)
    grav[:, 1] = np.float32(-9.81) * gs
    b.set_scenes(start, grav)
    return sc, b, start, grav, cnt["num_vertices"]


def plain_scene(k, grav, seed=3, ps=PAIRS, cs=CONTACTS):
    from seevr_amd import Context, Scene, caps_for, default_caps, default_params
    big = default_caps(max_bodies=300, max_hull_verts=300 * 82, max_vertices=300 * 82, max_triangles=300 * 130,
                       max_support_entries=300 * 193 + 1)
    sc = Scene(big)
    sc.build_batch(k, 1, 8, 256, SEG, seed)
    ctx = Context(caps_for(sc, ps, cs), default_params(pgs_iterations=30))
    sc.upload(ctx)
    cnt = sc.counts()
    return ctx, cnt["num_bodies"], cnt["num_vertices"], cnt["num_hull_verts"]


def scene_state(ctx, b0, n):
    return {k: ctx.download(k, w * n, w * b0) for k, w in STATE}


def compare_scene(batch, start, k, ctx, n, frame, ps=PAIRS, cs=CONTACTS):
    b0 = int(start[k])
    got, want = scene_state(batch.ctx, b0, n), scene_state(ctx, 0, n)
    for key in got:
        assert_bitwise(got[key], want[key], f"scene {k} {key} frame {frame}")
    bc = batch.counters()[k]
    pc = ctx.download("ATOMIC_COUNTERS", 3)
    assert int(bc[0]) == int(pc[0]) and int(bc[1]) == int(pc[1]), f"scene {k} counters {bc} vs {pc}"
    assert int(bc[2]) == int(pc[2]), f"scene {k} status bits"
    npairs = min(int(pc[0]), ps)
    gp = batch.ctx.download("COLLISION_PAIRS", 2 * npairs, 2 * k * ps).astype(np.int64) - b0
    assert_bitwise(gp.astype(np.uint32), ctx.download("COLLISION_PAIRS", 2 * npairs), f"scene {k} pairs")
    nc = min(int(pc[1]), cs)
    for key, w in (("CONTACT_POS", 3), ("CONTACT_NORMAL", 3), ("CONTACT_PENETRATION", 1),
                   ("CONTACT_LAMBDA_N", 1)):
        assert_bitwise(batch.ctx.download(key, w * nc, w * k * cs), ctx.download(key, w * nc), f"scene {k} {key}")


@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_matches_plain_path(use_graph):
    sc, b, start, grav, nv = build_batch(24)
    b.set_config(256, 2, use_graph)
    sizes = np.diff(start)
    picks = sorted({0, 7, int(np.argmax(sizes)), int(np.argmin(sizes))})
    plains = {k: plain_scene(k, grav) for k in picks}
    for f in range(50):  # bodies from 1-20 m: the lowest reach the ground within the run
        b.step(nv, 10, 1 / 60)
        for k, (ctx, n, pnv, nh) in plains.items():
            ctx.step(n, pnv, nh, 10, 1 / 60, tuple(float(x) for x in grav[k]))
            assert n == sizes[k]
            if f % 10 == 0 or f == 49:
                compare_scene(b, start, k, ctx, n, f)
    assert int(b.counters()[:, 1].max()) > 0, "scenes should be in contact"
    for ctx, *_ in plains.values():
        ctx.close()
    b.close()


def test_autotune_restores_state_and_keeps_results():
    sc, b, start, grav, nv = build_batch(16, seed=5)
    b.step(nv, 10, 1 / 60)
    n_all = int(start[-1])
    before = {k: b.ctx.download(k, w * n_all) for k, w in STATE}
    t, w, us = b.autotune(10, 1 / 60)
    assert t in (64, 128, 256, 512) and w >= 1 and us > 0
    assert b.pair_tile in (1, 4, 16, 64)
    after = {k: b.ctx.download(k, wd * n_all) for k, wd in STATE}
    for k in before:
        assert_bitwise(after[k], before[k], f"{k} restored after autotune")
    k = 3
    ctx, n, pnv, nh = plain_scene(k, grav, seed=5)
    ctx.step(n, pnv, nh, 10, 1 / 60, tuple(float(x) for x in grav[k]))
    for f in range(3):
        b.step(nv, 10, 1 / 60)
        ctx.step(n, pnv, nh, 10, 1 / 60, tuple(float(x) for x in grav[k]))
        compare_scene(b, start, k, ctx, n, f)
    ctx.close()
    b.close()


@pytest.mark.parametrize("lanes", [2, 4, 16, 64])
def test_pair_tiles_give_the_plain_path(lanes):
    """Every broadphase pair tile (a group of `lanes` lanes per row) gives a
    scene the plain path's pairs, contacts and state bit for bit, also when
    its pair slots overflow (the lexicographic prefix is kept)."""
    for ps, cs, seed in ((PAIRS, CONTACTS, 3), (40, 15, 9)):
        sc, b, start, grav, nv = build_batch(8, seed=seed, ps=ps, cs=cs)
        b.set_pair_tile(lanes)
        assert b.pair_tile == lanes
        k = int(np.argmax(np.diff(start)))
        ctx, n, pnv, nh = plain_scene(k, grav, seed=seed, ps=ps, cs=cs)
        for f in range(4):
            b.step(nv, 10, 1 / 60)
            ctx.step(n, pnv, nh, 10, 1 / 60, tuple(float(x) for x in grav[k]))
            compare_scene(b, start, k, ctx, n, f, ps=ps, cs=cs)
        if ps == 40:
            assert int(b.counters()[k, 2]) & 1, "pair overflow bit"
        for bad in (0, 3, 128):
            with pytest.raises(Exception):
                b.set_pair_tile(bad)
        ctx.close()
        b.close()


def test_batch_overflow_truncates_like_plain_path():
    ps, cs = 40, 15
    sc, b, start, grav, nv = build_batch(8, seed=9, ps=ps, cs=cs)
    k = int(np.argmax(np.diff(start)))
    ctx, n, pnv, nh = plain_scene(k, grav, seed=9, ps=ps, cs=cs)
    for f in range(4):
        b.step(nv, 10, 1 / 60)
        ctx.step(n, pnv, nh, 10, 1 / 60, tuple(float(x) for x in grav[k]))
        compare_scene(b, start, k, ctx, n, f, ps=ps, cs=cs)
    assert int(b.counters()[k, 2]) & 1, "pair overflow bit"
    ctx.close()
    b.close()


def test_phase_timing_changes_nothing_and_counts_every_scene():
    """The batched kernel's phase timing (device clock per phase) leaves the
    results bit for bit as they are, counts every scene-substep once, and
    reports positive phase times."""
    dt = 1.0 / 60.0
    sc1, b1, _, _, nv = build_batch(12)
    sc2, b2, _, _, _ = build_batch(12)
    sc1.close()
    sc2.close()
    for _ in range(10):
        b1.step(nv, 10, dt)
        b2.step(nv, 10, dt)
    b2.set_phase_timing(True)
    for _ in range(3):
        b1.step(nv, 10, dt)
        b2.step(nv, 10, dt)
    ph, n = b2.phase_times()
    assert_bitwise(b2.packed_state(), b1.packed_state(), "state with phase timing")
    assert np.array_equal(b2.counters(), b1.counters())
    assert n == 12 * 10 * 3
    assert all(v >= 0 for v in ph.values()) and ph["sweep"] > 0 and ph["narrowphase"] > 0
    again, n2 = b2.phase_times()  # read and zeroed
    assert n2 == 0 and sum(again.values()) == 0
    b2.set_phase_timing(False)
    b1.step(nv, 10, dt)
    b2.step(nv, 10, dt)
    assert b2.phase_times()[1] == 0
    assert_bitwise(b2.packed_state(), b1.packed_state(), "state after phase timing")
    b1.close()
    b2.close()
