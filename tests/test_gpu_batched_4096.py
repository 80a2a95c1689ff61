"""The 4,096-scene batched workload at its own size: scenes of
floor(exp(U[ln 8, ln 256])) bodies (16-segment spheres and boxes) plus their
own ground, gravity scaled by U[0.8, 1.2], all stepped by the persistent
batched kernel with the frame replayed as a hipGraph and the launch shape
picked by the autotuner at frame 120.  After 150 frames, 256 evenly spaced
scenes must equal the oracle stepping each scene alone for 150 frames, and
64 of them must equal the plain path (sv_physics_step) stepping the scene
alone over frames 121-150 from the batch's frame-120 state: body state and
counters bit for bit."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from harness import World, assert_bitwise, oracle

pytestmark = pytest.mark.gpu

SCENES, SEG, SEED, PS, CS = 4096, 16, 1, 4096, 2048
FRAMES, SWITCH = 150, 120
STATE = (("BODY_POS", 3), ("BODY_QUAT", 4), ("BODY_VEL", 3), ("BODY_OMEGA", 3))


def one_scene(k):
    from seevr_amd import Scene, default_caps
    sc = Scene(default_caps(max_bodies=300, max_hull_verts=300 * 290, max_vertices=300 * 290,
                            max_triangles=300 * 520, max_support_entries=300 * 193 + 1))
    start, gs = sc.build_batch(k, 1, 8, 256, SEG, SEED)
    return sc, float(np.float32(-9.81) * np.float32(gs[0]))


def oracle_run(k):
    sc, gy = one_scene(k)
    w = World.from_scene(sc, max_pairs=PS, max_contacts=CS, pgs_iterations=30)
    sc.close()
    for _ in range(FRAMES):
        w.oracle_step(10, 1 / 60, (0.0, gy, 0.0))
    return w


def test_batched_4096_scenes_against_oracle_and_plain_path():
    from seevr_amd import Batch, Context, Scene, caps_for, default_caps, default_params
    from seevr_amd._lib import BatchCaps
    big = default_caps(max_bodies=SCENES * 257 + 16, max_hull_verts=SCENES * 257 * 290,
                       max_vertices=SCENES * 257 * 290, max_triangles=SCENES * 257 * 520,
                       max_support_entries=SCENES * 257 * 193 + 1)
    sc = Scene(big)
    start, gs = sc.build_batch(0, SCENES, 8, 256, SEG, SEED)
    cnt = sc.counts()
    bc = BatchCaps(max_scenes=SCENES, max_bodies=cnt["num_bodies"], max_hull_verts=cnt["num_hull_verts"],
                   max_vertices=cnt["num_vertices"], max_triangles=max(cnt["num_triangles"], 1),
                   max_support_entries=cnt["num_support_entries"], max_scene_bodies=512, max_scene_pairs=PS,
                   max_scene_contacts=CS)
    b = Batch(bc, default_params(pgs_iterations=30))
    sc.upload(b.ctx)
    sc.close()
    grav = np.zeros((SCENES, 3), np.float32)
    grav[:, 1] = np.float32(-9.81) * gs
    b.set_scenes(start, grav)
    nv = cnt["num_vertices"]
    checked = list(range(0, SCENES, SCENES // 256))
    oracle()  # load the oracle before the worker threads use it
    with ThreadPoolExecutor(16) as pool:  # the oracle runs on the host meanwhile
        runs = pool.map(oracle_run, checked)
        for f in range(SWITCH):
            b.step(nv, 10, 1 / 60)
        snap = {k: b.ctx.download(k, w * cnt["num_bodies"]) for k, w in STATE}
        threads, wgs, us = b.autotune(10, 1 / 60)
        print(f"autotune: {threads} threads x {wgs} workgroups/CU, {us:.0f} us/substep", flush=True)
        for f in range(SWITCH, FRAMES):
            b.step(nv, 10, 1 / 60)
        counters = b.counters()
        worlds = list(runs)
    assert counters[:, 1].sum() > 50000, "the scenes should be in contact"
    for k, w in zip(checked, worlds):
        b0, n = int(start[k]), int(start[k + 1] - start[k])
        assert w.n == n
        for key, width in STATE:
            assert_bitwise(b.ctx.download(key, width * n, width * b0), w.a[key][: width * n], f"scene {k} {key}")
        assert tuple(int(v) for v in counters[k, :2]) == tuple(int(v) for v in w.a["ATOMIC_COUNTERS"][:2]), k
    print(f"{len(checked)} scenes equal the oracle after {FRAMES} frames", flush=True)
    # the plain path, each scene alone, from the batch's frame-120 state
    for k in checked[::4]:
        b0, n = int(start[k]), int(start[k + 1] - start[k])
        s1, gy = one_scene(k)
        c1 = s1.counts()
        ctx = Context(caps_for(s1, PS, CS), default_params(pgs_iterations=30))
        s1.upload(ctx)
        s1.close()
        for key, width in STATE:
            ctx.upload(key, snap[key][width * b0: width * (b0 + n)])
        for _ in range(FRAMES - SWITCH):
            ctx.step(n, c1["num_vertices"], c1["num_hull_verts"], 10, 1 / 60, (0.0, gy, 0.0))
        for key, width in STATE:
            assert_bitwise(b.ctx.download(key, width * n, width * b0), ctx.download(key, width * n),
                           f"scene {k} {key} vs plain path")
        assert tuple(int(v) for v in counters[k, :2]) == ctx.read_counters(), k
        ctx.close()
    b.close()
