#!/usr/bin/env python3
"""Benchmark of the plain physics path on one MI355X.

Workload: 32,768 bodies (half boxes with half-extents U[0.2,0.6] m, half
spheres with radius U[0.2,0.5] m, 16 segments) dropped from 1-60 m into a
30x30 m container (ground + 4 static walls), density 1000, friction 0.5,
restitution 0.1, gravity -9.81, dt 1/60 s, 10 substeps, 30 PGS iterations,
FP32.  A step is one frame: one Physics.step call (10 substeps plus the
render-vertex update) and the LBVH rebuilt over the body boxes.  Every run
first settles the pile for 240 frames, then runs W warmup frames, then times
K frames, so the timed state does not depend on W.  The metric is
body-substeps/s over the timed frames.

Prints one JSON line (see the keys below).  Usage:
  python bench.py [--gpus 1] [--steps K] [--warmup W] [--cpu-sample-substeps S]
                  [--window-file PATH]   (timed windows for seevr_amd.profsum)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

SUBSTEPS = 10
DT = 1.0 / 60.0
GRAVITY = (0.0, -9.81, 0.0)
PGS_ITERS = 30
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def build_workload(count: int, seed: int):
    from seevr_amd import Scene, default_caps
    n = count + 5
    caps = default_caps(max_bodies=n, max_hull_verts=n * 289, max_vertices=n * 289, max_triangles=n * 512,
                        max_support_entries=n * 192 + 1)
    sc = Scene(caps)
    sc.build_workload(2, count, 16, seed)
    return sc


SETTLE_FRAMES = 240  # frames every run steps before its warmup and timed frames
WINDOWS: dict[str, list[int]] = {}  # section -> CLOCK_MONOTONIC ns bracketing its timed frames


def window(name: str, t0: int, t1: int) -> None:
    WINDOWS[name] = [t0, t1]


PROFILE = ROOT / "profiles" / "c2_32k_container_kernels.md"
PROFILE_TRI = ROOT / "profiles" / "c2_triangle_bvh_kernels.md"
VALU_PEAK = 256 * 4 * 16 * 2.4e9  # non-packed FP32 lane operations per second
STAGE_FRAMES = 5  # frames after the timed ones whose stages are all timed
BROADPHASE_KERNEL = {"all_pairs": "k_broadphase<false>", "sweep": "k_sap<false>", "lbvh": "k_bb_pairs<false>"}
STAGE_KERNEL = {"phys/pgs_sweep": "k_solve_seqp2<64>", "phys/aabb": "k_aabb", "phys/gravity": "k_gravity",
                "phys/integrate": "k_integrate", "phys/broad_phase": "k_broadphase<false>",
                "bvh/morton_codes": "k_morton", "bvh/radix_sort": "k_os_one<4>"}
# kernels (and launches) behind a stage's time, for its profiled traffic
STAGE_LAUNCHES = {"bvh/radix_sort": {"k_os_one<4>": 1}}


def algorithmic_bytes(stage: str, n: int, n_hull: int, n_dyn: int, pairs: int, nc: int) -> float:
    """Bytes each stage must move at minimum per launch.

    aabb: pos 12 + quat 16 + hull range 8 per body read, 12 per hull vertex
    read, 24 per body written.  broadphase: 28 per body read (box + static
    flag) + 8 per pair written.  gravity 52, integrate 84 per dynamic body.
    LBVH radix sort: one histogram read of the keys (4) and 16 per key per
    pass (key and payload read and written), 4 passes of 8-bit digits over
    30-bit keys; Morton codes: 12 in + 4 out.
    Ordered PGS sweep: per item (iteration x contact) the two bodies'
    velocities and angular velocities read and written (96 B), the
    accumulated impulse read and written (8 B) and the contact's 20-word
    constant record read (80 B): 184 B.
    """
    if stage == "phys/aabb":
        return 36.0 * n + 12.0 * n_hull + 24.0 * n
    if stage == "phys/gravity":
        return 52.0 * n_dyn
    if stage == "phys/integrate":
        return 84.0 * n_dyn
    if stage == "phys/broad_phase":
        return 28.0 * n + 8.0 * pairs
    if stage == "phys/pgs_sweep":
        return 184.0 * PGS_ITERS * nc
    if stage == "bvh/radix_sort":  # 30-bit Morton keys, 8-bit digits: 4 passes of key + index in and out
        return 16.0 * n * 4 + 4.0 * n
    if stage == "bvh/morton_codes":
        return 16.0 * n
    return 0.0


def profiled_traffic(kernel: str, profile: Path = PROFILE):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (dispatches of the same timed frames), scaled by the
    copy-kernel calibration in the same file."""
    if not profile.exists():
        return None
    rows = {}
    for line in profile.read_text().splitlines():
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 4 and cells[0].startswith(("k_", "__")):
            try:
                rows[cells[0]] = (float(cells[2]), float(cells[3]))
            except ValueError:
                pass
    parts = kernel if isinstance(kernel, dict) else {kernel: 1}  # kernel -> launches per stage
    if "k_copy_calib" not in rows or any(k not in rows for k in parts):
        return None
    cal_f, cal_w = rows["k_copy_calib"]
    true_kib = (256 << 20) / 1024.0
    return sum(m * 1024.0 * (rows[k][0] * true_kib / cal_f + rows[k][1] * true_kib / cal_w) for k, m in parts.items())


def roofline_entry(stage, stages, n, nh, n_dyn, pairs, nc):
    us = stages.get(stage, (0.0, 0))[0]
    alg = algorithmic_bytes(stage, n, nh, n_dyn, pairs, nc)
    achieved = alg / (us * 1e-6) / 1e9 if us > 0 else 0.0
    kern = STAGE_KERNEL.get(stage, stage)
    return {"bound": "hbm", "kernel": f"{kern} ({stage})", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": profiled_traffic(STAGE_LAUNCHES.get(stage, kern)), "avg_us": round(us, 2), "alg_bytes": int(alg)}


def sweep_roofline(sweep_us: float, depth: int, items: int, handoff_us: float, nc: int) -> dict:
    """The sweep against its latency roofline: the serial order needs `depth`
    dependent steps (the longest chain of its dependency graph, sv_sweep_depth),
    and one step cannot cost less than one hand-off between waves through a
    device-coherent word (sv_debug_handoff, across XCDs).  achieved = levels
    per microsecond, peak = 1 / hand-off, frac = depth x hand-off / sweep."""
    lv = depth / sweep_us if sweep_us > 0 else 0.0
    peak = 1.0 / handoff_us if handoff_us > 0 else 0.0
    alg = 184.0 * items
    return {"bound": "latency", "kernel": f"{STAGE_KERNEL['phys/pgs_sweep']} (phys/pgs_sweep)",
            "achieved": round(lv, 4), "peak": round(peak, 4), "unit": "dependent levels/us",
            "frac": round(lv / peak, 4) if peak else None, "traffic": profiled_traffic(STAGE_KERNEL["phys/pgs_sweep"]),
            "avg_us": round(sweep_us, 2), "dag_depth": depth, "items": items, "handoff_us": round(handoff_us, 4),
            "us_per_level": round(sweep_us / depth, 3) if depth else None,
            "latency_floor_us": round(depth * handoff_us, 1), "alg_bytes": int(alg),
            "hbm_view_gbs": round(alg / (sweep_us * 1e-6) / 1e9, 1) if sweep_us > 0 else None}


def cpu_baseline(ctx, scene, n, substeps: int, mt: bool = True) -> dict:
    """Time the C oracle on `substeps` substeps from the device state: the
    OpenMP build (every host thread for the per-body, per-row and per-pair
    stages, the sweep on one core; the serial build's results) or, with
    mt=False, the scalar build on one core."""
    sys.path.insert(0, str(ROOT / "tests"))
    from harness import World, oracle
    w = World.from_scene(scene, max_pairs=ctx.caps.max_pairs, max_contacts=ctx.caps.max_contacts,
                         pgs_iterations=PGS_ITERS)
    st = w.download_state(ctx)
    for k, v in st.items():
        w.a[k][:] = v
    cores = oracle(mt=True).threads() if mt else 1
    t0 = time.perf_counter()
    w.oracle_step(substeps, DT * substeps / SUBSTEPS, GRAVITY, mt=mt)
    el = time.perf_counter() - t0
    how = f"C oracle, OpenMP build, {cores} threads (sweep serial)" if mt else "scalar C oracle, 1 core"
    return {"value": n * substeps / el, "unit": "body-substeps/s", "cores": cores, "kind": "port",
            "sample": f"{substeps} substeps of the same scene from the state at the start of the timed "
                      f"frames (N={n}), {how}, {el:.1f} s"}


TRI_STAGES = ("bvh/centroids", "bvh/scene_bounds", "bvh/morton_codes", "bvh/radix_sort", "bvh/build_tree",
              "bvh/leaf_aabbs", "bvh/clear_flags", "bvh/propagate")
TRI_KERNELS = {"bvh/centroids": {"k_centroids": 1}, "bvh/morton_codes": {"k_morton": 1},
               "bvh/radix_sort": {"k_os_hist<4>": 1, "k_os_pass<1024, 8>": 4}}


def triangle_bvh_section(ctx, counts: dict, reps: int = 5) -> dict:
    """The LBVH over the scene's render triangles (bvh.comp's build, the
    renderer's tree; 8.57M triangles of the 32k-body container): `reps` timed
    builds, then each stage timed alone with HIP events over 3 more builds
    (events around one stage only).  Rooflines for the bandwidth-bound
    stages: centroids read 3 indices per triangle (12 B) and each vertex once
    (12 B), and write a centroid and two indices (20 B); Morton codes read a
    centroid (12 B) and write a code (4 B); the radix sort of 32-bit keys
    reads the keys once for its histograms (4 B) and reads and writes key and
    payload in each of 4 passes (64 B)."""
    nt, nv = counts["num_triangles"], counts["num_vertices"]
    ctx.bvh_build(nt)
    w0, p0 = time.monotonic_ns(), time.perf_counter()
    for _ in range(reps):
        ctx.bvh_build(nt)
    el = time.perf_counter() - p0
    window("tri_bvh", w0, time.monotonic_ns())
    stages = {}
    for st in TRI_STAGES:
        ctx.enable_stage_timing(True, only=st)
        ctx.stage_times(reset=True)
        for _ in range(3):
            ctx.bvh_build(nt)
        stages[st] = ctx.stage_times(reset=True).get(st, (0.0, 0))[0]
    ctx.enable_stage_timing(False)
    alg = {"bvh/centroids": 32.0 * nt + 12.0 * nv, "bvh/morton_codes": 16.0 * nt, "bvh/radix_sort": 68.0 * nt}
    roof = {}
    for st, b in alg.items():
        us = stages[st]
        a = b / (us * 1e-6) / 1e9 if us > 0 else 0.0
        roof[st] = {"bound": "hbm", "kernel": " + ".join(f"{m} x {k}" if m > 1 else k for k, m in TRI_KERNELS[st].items()),
                    "achieved": round(a, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(a / HBM_PEAK_GBS, 4),
                    "traffic": profiled_traffic(TRI_KERNELS[st], PROFILE_TRI), "avg_us": round(us, 1),
                    "alg_bytes": int(b)}
    log(f"triangle BVH over {nt} triangles: {1000 * el / reps:.2f} ms per build; stages (us) " +
        ", ".join(f"{k.split('/')[1]} {v:.0f}" for k, v in stages.items()))
    return {"triangles": nt, "vertices": nv, "ms_per_build": round(1000 * el / reps, 3), "builds": reps,
            "stage_us": {k: round(v, 1) for k, v in stages.items()}, "roofline_hbm_stages": roof,
            "note": "stage times from HIP events around one stage per build; they include about 4 us of event "
                    "and launch overhead, which the kernel durations of the rocprof summary do not"}


def batched_section(frames: int, warmup: int, scenes: int = 4096, plain_sample: int = 32) -> dict:
    """The batched form on the 4,096-scene workload: env-steps/s over `frames`
    timed frames (one frame = 10 substeps of every scene, replayed as one
    hipGraph), against the plain path stepping a sample of the same scenes
    one by one (the per-scene frame time is averaged over the sample)."""
    from seevr_amd import Batch, Context, Scene, caps_for, default_caps, default_params
    from seevr_amd._lib import BatchCaps
    t0 = time.perf_counter()
    big = default_caps(max_bodies=scenes * 257 + 16, max_hull_verts=scenes * 257 * 290,
                       max_vertices=scenes * 257 * 290, max_triangles=scenes * 257 * 520,
                       max_support_entries=scenes * 257 * 193 + 1)
    sc = Scene(big)
    start, gs = sc.build_batch(0, scenes, 8, 256, 16, 1)
    cnt = sc.counts()
    log(f"batched: {scenes} scenes, {cnt['num_bodies']} bodies built in {time.perf_counter() - t0:.1f} s")
    bc = BatchCaps(max_scenes=scenes, max_bodies=cnt["num_bodies"], max_hull_verts=cnt["num_hull_verts"],
                   max_vertices=cnt["num_vertices"], max_triangles=max(cnt["num_triangles"], 1),
                   max_support_entries=cnt["num_support_entries"], max_scene_bodies=512, max_scene_pairs=4096,
                   max_scene_contacts=2048)
    b = Batch(bc, default_params(pgs_iterations=PGS_ITERS))
    sc.upload(b.ctx)
    grav = np.zeros((scenes, 3), np.float32)
    grav[:, 1] = np.float32(GRAVITY[1]) * gs
    b.set_scenes(start, grav)
    sc.close()
    nv = cnt["num_vertices"]
    for i in range(warmup):
        b.step(nv, SUBSTEPS, DT)
    threads, wgs, us = b.autotune(SUBSTEPS, DT)
    pair_tile = b.pair_tile
    log(f"batched autotune: {threads} threads x {wgs} workgroups/CU, broadphase pair tile {pair_tile} "
        f"lanes per row, {us:.0f} us/substep")
    b.step(nv, SUBSTEPS, DT)
    w0, t0 = time.monotonic_ns(), time.perf_counter()
    for i in range(frames):
        b.step(nv, SUBSTEPS, DT)
    el = time.perf_counter() - t0
    window("batched", w0, time.monotonic_ns())
    c = b.counters()
    env_steps = scenes * frames / el
    # where a settled substep goes: the kernel's phases by the device clock
    # (each workgroup's time per phase, summed), over 3 more frames
    b.set_phase_timing(True)
    t1 = time.perf_counter()
    for _ in range(3):
        b.step(nv, SUBSTEPS, DT)
    el_ph = (time.perf_counter() - t1) / 3
    ph, n_sc = b.phase_times()
    b.set_phase_timing(False)
    tot_ph = sum(ph.values())
    grid = min(scenes, wgs * 256)  # the persistent grid: workgroups per CU x the MI355X's 256 CUs
    phase = {"share_pct": {k: round(100 * v / tot_ph, 1) for k, v in ph.items()},
             "us_per_scene_substep": {k: round(v / max(n_sc, 1), 2) for k, v in ph.items()},
             "workgroup_busy_fraction": round(tot_ph * 1e-6 / (grid * 3 * el_ph), 3),
             "timed_ms_per_frame": round(1000 * el_ph, 2), "scene_substeps": n_sc,
             "note": "device wall clock (100 MHz) read by each workgroup's first thread after every phase; "
                     "busy fraction = summed workgroup time / (workgroups x wall time)"}
    log("batched phases (share of workgroup time): " +
        ", ".join(f"{k} {v}%" for k, v in phase["share_pct"].items()) +
        f"; workgroups busy {phase['workgroup_busy_fraction']:.2f} of the wall time")
    # plain path, one scene at a time, on an evenly spaced sample, each
    # started from that scene's current batched state
    ks = list(range(0, scenes, max(scenes // plain_sample, 1)))[:plain_sample]
    tot, bodies = 0.0, 0
    for k in ks:
        s1 = Scene(default_caps(max_bodies=300, max_hull_verts=300 * 290, max_vertices=300 * 290,
                                max_triangles=300 * 520, max_support_entries=300 * 193 + 1))
        s1.build_batch(k, 1, 8, 256, 16, 1)
        c1 = s1.counts()
        ctx = Context(caps_for(s1, 4096, 2048), default_params(pgs_iterations=PGS_ITERS))
        s1.upload(ctx)
        n1, v1, h1 = c1["num_bodies"], c1["num_vertices"], c1["num_hull_verts"]
        for key, w in (("BODY_POS", 3), ("BODY_QUAT", 4), ("BODY_VEL", 3), ("BODY_OMEGA", 3)):
            ctx.upload(key, b.ctx.download(key, w * n1, w * int(start[k])))
        g = tuple(float(x) for x in grav[k])
        ctx.step(n1, v1, h1, SUBSTEPS, DT, g)
        t1 = time.perf_counter()
        for _ in range(5):
            ctx.step(n1, v1, h1, SUBSTEPS, DT, g)
        tot += (time.perf_counter() - t1) / 5
        bodies += n1
        ctx.close()
        s1.close()
    b.close()
    plain_env = len(ks) / tot
    log(f"batched: {1000 * el / frames:.2f} ms/frame, {env_steps:.0f} env-steps/s; plain one by one "
        f"{1000 * tot / len(ks):.2f} ms per scene-frame -> {plain_env:.0f} env-steps/s")
    return {"metric": "env-steps/s @4096 uneven scenes", "value": round(env_steps, 1), "unit": "env-steps/s",
            "frames": frames, "frames_before_timing": warmup + 1,
            "data": "synthetic: 4096 seeded scenes, floor(exp(U[ln 8, ln 256])) bodies each plus a ground, "
                    "gravity x U[0.8, 1.2]", "ms_per_frame": round(1000 * el / frames, 3),
            "scenes": scenes, "bodies": cnt["num_bodies"],
            "body_substeps_per_s": round(cnt["num_bodies"] * SUBSTEPS * frames / el, 1),
            "contacts_last_substep": {"mean": round(float(c[:, 1].mean()), 2), "max": int(c[:, 1].max())},
            "launch_shape": {"threads": threads, "workgroups_per_cu": wgs, "pair_tile_lanes_per_row": pair_tile,
                             "autotune_us_per_substep": round(us, 1)},
            "plain_one_by_one": {"env_steps_per_s": round(plain_env, 1), "sample_scenes": len(ks),
                                 "sample_mean_bodies": round(bodies / len(ks), 1)},
            "speedup_vs_plain": round(env_steps / plain_env, 1), "phases": phase}


def differentiable_section(frames: int = 60, scenes: int = 256, bodies: int = 64, reps: int = 3) -> dict:
    """Cost of one backward pass relative to the forward rollout on the
    256-scene x 64-body workload: initial v, w ~ N(0, 1), per-scene,
    per-material friction U[0.2, 0.8] and restitution U[0, 0.5], 60 frames.
    Forward = sv_batch_diff_forward (rollout + frame checkpoints); backward =
    sv_batch_diff_backward (recompute each frame from its checkpoint with the
    tape, then the adjoint kernels).  The loss is the squared distance of the
    final positions to targets (final positions + U[-0.5, 0.5])."""
    from seevr_amd import Batch, Scene, default_caps, default_params
    from seevr_amd._lib import BatchCaps
    rng = np.random.default_rng(4)
    big = default_caps(max_bodies=scenes * (bodies + 1), max_hull_verts=scenes * (bodies + 1) * 290,
                       max_vertices=scenes * (bodies + 1) * 290, max_triangles=scenes * (bodies + 1) * 520,
                       max_support_entries=scenes * (bodies + 1) * 193 + 1)
    sc = Scene(big)
    start, gs = sc.build_batch(0, scenes, bodies, bodies, 16, 4, top=6.0)
    cnt = sc.counts()
    bc = BatchCaps(max_scenes=scenes, max_bodies=cnt["num_bodies"], max_hull_verts=cnt["num_hull_verts"],
                   max_vertices=cnt["num_vertices"], max_triangles=max(cnt["num_triangles"], 1),
                   max_support_entries=cnt["num_support_entries"], max_scene_bodies=512, max_scene_pairs=2048,
                   max_scene_contacts=512)
    b = Batch(bc, default_params(pgs_iterations=PGS_ITERS))
    sc.upload(b.ctx)
    sc.close()
    grav = np.zeros((scenes, 3), np.float32)
    grav[:, 1] = np.float32(GRAVITY[1]) * gs
    b.set_scenes(start, grav)
    n = b.num_bodies
    x0 = b.packed_state()
    inv_mass = b.ctx.download("BODY_INV_MASS", n)
    dyn = inv_mass > 0
    vel = x0[7 * n:10 * n].reshape(n, 3)
    omg = x0[10 * n:13 * n].reshape(n, 3)
    vel[dyn] = rng.normal(size=(dyn.sum(), 3))
    omg[dyn] = rng.normal(size=(dyn.sum(), 3))
    mu_m = rng.uniform(0.2, 0.8, (scenes, 2)).astype(np.float32)
    e_m = rng.uniform(0.0, 0.5, (scenes, 2)).astype(np.float32)
    scene_of = np.repeat(np.arange(scenes), np.diff(start))
    mat = np.where(dyn, 1, 0)
    fric, rest = mu_m[scene_of, mat], e_m[scene_of, mat]
    b.ctx.upload("GEOM_FRICTION", fric)
    b.ctx.upload("GEOM_RESTITUTION", rest)
    b.diff_reserve(frames, SUBSTEPS)
    b.set_packed_state(x0)
    b.diff_forward(frames, SUBSTEPS, DT)
    target = b.packed_state()[: 3 * n] + rng.uniform(-0.5, 0.5, 3 * n).astype(np.float32)
    fwd, bwd = [], []
    c = None
    w0 = time.monotonic_ns()
    for _ in range(reps):
        b.set_packed_state(x0)
        t0 = time.perf_counter()
        b.diff_forward(frames, SUBSTEPS, DT)
        fwd.append(time.perf_counter() - t0)
        c = b.counters()  # the forward rollout's last substep (backward recomputes earlier frames)
        xf = b.packed_state()
        g = np.zeros(13 * n, np.float32)
        g[: 3 * n] = 2.0 * (xf[: 3 * n] - target)
        t0 = time.perf_counter()
        gstate, gfric, grest = b.diff_backward(frames, g, SUBSTEPS, DT)
        bwd.append(time.perf_counter() - t0)
    window("diff", w0, time.monotonic_ns())
    b.close()
    tf, tb = min(fwd), min(bwd)
    gmu = np.zeros((scenes, 2))
    np.add.at(gmu, (scene_of, mat), gfric)
    log(f"differentiable: forward {1000 * tf:.1f} ms, backward {1000 * tb:.1f} ms, ratio {(tf + tb) / tf:.2f}")
    return {"metric": "fwd+bwd vs fwd cost", "value": round((tf + tb) / tf, 3), "unit": "x",
            "higher_is_better": False, "scenes": scenes, "bodies_per_scene": bodies + 1, "frames": frames,
            "forward_ms": round(1000 * tf, 2), "backward_ms": round(1000 * tb, 2),
            "contacts_last_forward_substep": {"mean": round(float(c[:, 1].mean()), 2), "max": int(c[:, 1].max())},
            "grad_norms": {"state": float(np.linalg.norm(gstate)), "friction_per_material": float(np.linalg.norm(gmu))},
            "data": "synthetic: 256 seeded scenes of 64 bodies over a ground, v and w ~ N(0, 1), per-scene "
                    "per-material friction U[0.2, 0.8] and restitution U[0, 0.5]; targets = final positions "
                    "+ U[-0.5, 0.5]"}


def sweep_variant_section(ctx, n, default_variant: str = "seq") -> dict:
    """The multi-workgroup sweep's hand-off variants from the same state (the
    pile where the timed frames left it): one substep's collision stages,
    then the sweep timed with HIP events, for each variant, with its time per
    level of the sweep's dependency graph; all must leave bit-identical
    velocities."""
    state = {k: ctx.download(k, w * n) for k, w in (("BODY_POS", 3), ("BODY_QUAT", 4), ("BODY_VEL", 3),
                                                      ("BODY_OMEGA", 3))}
    dt_sub = float(np.float32(DT) / np.float32(SUBSTEPS))
    out, ref = {}, None
    for variant in ("paired", "pipelined", "seq_basic", "coherent", "fenced"):
        for key, v in state.items():
            ctx.upload(key, v)
        ctx.set_sweep_variant(variant)
        ctx.dispatch(8, 1)
        ctx.dispatch(0, n, dt_sub, GRAVITY)
        ctx.dispatch(9, n)
        ctx.dispatch(2, n)
        ctx.dispatch(3, 0, num_bodies=n)
        ctx.dispatch(4, 0)
        ctx.enable_stage_timing(True)
        ctx.stage_times(reset=True)
        ctx.dispatch(5, n, dt_sub)
        st = ctx.stage_times(reset=True)
        ctx.enable_stage_timing(False)
        vel = np.concatenate([ctx.download("BODY_VEL", 3 * n), ctx.download("BODY_OMEGA", 3 * n)])
        ref = vel if ref is None else ref
        depth, _ = ctx.sweep_depth()
        us = st.get("phys/pgs_sweep", (0.0, 0))[0]
        out[variant] = {"sweep_us": round(us, 1), "us_per_level": round(us / depth, 3) if depth else None,
                        "bit_identical": bool(np.array_equal(vel.view(np.uint32), ref.view(np.uint32)))}
    ctx.set_sweep_variant(default_variant)
    for key, v in state.items():
        ctx.upload(key, v)
    log("sweep hand-off variants from the same state: " +
        ", ".join(f"{k} {v['sweep_us']:.0f} us" for k, v in out.items()))
    return out


def warm_start_section(ctx, n, nv, nh, factor: float = 1.0, iterations=(4, 10, 30)) -> dict:
    """Warm start beside the reference path, on the pile where the timed
    frames left it: one frame with warm start from that state fills the
    contact cache, then the next substep's collision stages and a sweep of
    k iterations run from the same state cold (impulses cleared, the
    reference) and warm (carried over); the normal-constraint residual and
    the sweep time are reported for each.  The frame is rerun before every
    probe so each starts from the same state and cache."""
    state = {k: ctx.download(k, w * n) for k, w in (("BODY_POS", 3), ("BODY_QUAT", 4), ("BODY_VEL", 3),
                                                      ("BODY_OMEGA", 3))}
    dt_sub = float(np.float32(DT) / np.float32(SUBSTEPS))
    out = {}
    for k in iterations:
        row = {}
        for mode in ("cold", "warm"):
            for key, v in state.items():
                ctx.upload(key, v)
            ctx.set_params(pgs_iterations=PGS_ITERS)
            ctx.set_warm_start(factor)
            ctx.step(n, nv, nh, SUBSTEPS, DT, GRAVITY)
            if mode == "cold":
                ctx.set_warm_start(0.0)
            ctx.set_params(pgs_iterations=k)
            ctx.dispatch(8, 1)
            ctx.dispatch(0, n, dt_sub, GRAVITY)
            ctx.dispatch(9, n)
            ctx.dispatch(2, n)
            ctx.dispatch(3, 0, num_bodies=n)
            ctx.dispatch(4, 0)
            ctx.enable_stage_timing(True)
            ctx.stage_times(reset=True)
            ctx.dispatch(5, n, dt_sub)
            st = ctx.stage_times(reset=True)
            ctx.enable_stage_timing(False)
            rms, mx, cnt = ctx.residual()
            row[mode] = {"residual_rms": rms, "residual_max": mx, "contacts": cnt,
                         "sweep_us": round(st.get("phys/pgs_sweep", (0.0, 0))[0], 1)}
        out[f"{k}_iterations"] = row
    ctx.set_warm_start(0.0)
    ctx.set_params(pgs_iterations=PGS_ITERS)
    log(f"warm start (factor {factor}) vs cold from the same state: " +
        ", ".join(f"{k}: rms {v['cold']['residual_rms']:.4g} -> {v['warm']['residual_rms']:.4g}" for k, v in out.items()))
    return {"factor": factor, "from_same_state": out}


def large_scene_section(frames: int = 120, timed: int = 20, compare: bool = True, solver_wgs: int = 0,
                        sweep_variant: str = "seq") -> dict:
    """The 1,048,576-body scene (70% spheres of 81 hull verts, 200x200x50 m
    over a 220x1x220 m ground; 62M hull verts, so hull ranges are read as
    uint32): `frames` frames with the LBVH broadphase, the last `timed` of
    them timed; then, from that state, the ordered sweep (reference order)
    against the graph-coloured one: time per frame over 3 frames each and
    the normal-constraint residual of one substep."""
    from seevr_amd import Context, Scene, default_caps, default_params
    n = 1048577
    t0 = time.perf_counter()
    sc = Scene(default_caps(max_bodies=n, max_hull_verts=n * 82, max_vertices=n * 82, max_triangles=n * 130,
                            max_support_entries=n * 193 + 1))
    sc.build_workload(5, 0, 0, 1)
    c = sc.counts()
    nb, nv, nh = c["num_bodies"], c["num_vertices"], c["num_hull_verts"]
    caps = default_caps(max_bodies=nb, max_hull_verts=nh, max_vertices=nv, max_triangles=c["num_triangles"],
                        max_support_entries=c["num_support_entries"], max_pairs=8 * nb + 1024,
                        max_contacts=6 * nb + 1024)
    ctx = Context(caps, default_params(pgs_iterations=PGS_ITERS, hull_range_u32=1))
    sc.upload(ctx)
    if solver_wgs:
        ctx.set_solver("ordered", solver_wgs)
    ctx.set_sweep_variant(sweep_variant)
    sc.close()
    log(f"large scene: {nb} bodies, {nh} hull verts, built and uploaded in {time.perf_counter() - t0:.1f} s")
    for i in range(frames - timed):
        ctx.step(nb, nv, nh, SUBSTEPS, DT, GRAVITY)
        if i % 10 == 9:
            log(f"large scene: frame {i + 1}/{frames}, counters {ctx.read_counters()}")
    w0, t0 = time.monotonic_ns(), time.perf_counter()
    for _ in range(timed):
        ctx.step(nb, nv, nh, SUBSTEPS, DT, GRAVITY)
    el = time.perf_counter() - t0
    window("large", w0, time.monotonic_ns())
    depth, items = ctx.sweep_depth()
    # stage times in two more frames (events around every stage slow a frame)
    ctx.enable_stage_timing(True)
    ctx.stage_times(reset=True)
    for _ in range(2):
        ctx.step(nb, nv, nh, SUBSTEPS, DT, GRAVITY)
    stages = ctx.stage_times(reset=True)
    ctx.enable_stage_timing(False)
    pairs, contacts = ctx.read_counters()
    state = {k: ctx.download(k, w * nb) for k, w in (("BODY_POS", 3), ("BODY_QUAT", 4), ("BODY_VEL", 3),
                                                       ("BODY_OMEGA", 3))}
    dt_sub = float(np.float32(DT) / np.float32(SUBSTEPS))
    cmp = {}
    for solver in (("auto", "colored") if compare else ()):
        for k, v in state.items():
            ctx.upload(k, v)
        ctx.set_solver(solver)
        t1 = time.perf_counter()
        for _ in range(3):
            ctx.step(nb, nv, nh, SUBSTEPS, DT, GRAVITY)
        ms = 1000 * (time.perf_counter() - t1) / 3
        for k, v in state.items():
            ctx.upload(k, v)
        ctx.dispatch(8, 1)
        ctx.dispatch(0, nb, dt_sub, GRAVITY)
        ctx.dispatch(9, nb)
        ctx.dispatch(2, nb)
        ctx.dispatch(3, 0, num_bodies=nb)
        ctx.dispatch(4, 0)
        ctx.enable_stage_timing(True)
        ctx.stage_times(reset=True)
        ctx.dispatch(5, nb, dt_sub)
        sweep = ctx.stage_times(reset=True).get("phys/pgs_sweep", (0.0, 0))[0]
        ctx.enable_stage_timing(False)
        rms, mx, cnt = ctx.residual()
        cmp["ordered" if solver == "auto" else "colored"] = {
            "ms_per_frame": round(ms, 2), "sweep_us_per_substep": round(sweep, 1), "residual_rms": rms,
            "residual_max": mx, "contacts": cnt}
    ctx.close()
    val = nb * SUBSTEPS * timed / el
    log(f"large scene: {1000 * el / timed:.1f} ms/frame over frames {frames - timed + 1}-{frames}, "
        f"{val:.3e} body-substeps/s; solvers {cmp}")
    top = {k: round(v[0], 1) for k, v in sorted(stages.items(), key=lambda kv: -kv[1][0] * kv[1][1])[:6]}
    return {"metric": "body-substeps/s @1M bodies", "value": round(val, 1), "unit": "body-substeps/s",
            "bodies": nb, "hull_verts": nh, "frames": frames, "timed_frames": timed,
            "ms_per_frame": round(1000 * el / timed, 2), "pairs_last_substep": pairs,
            "contacts_last_substep": contacts, "broadphase": "lbvh", "stage_us_per_launch": top,
            "sweep_dag_depth": depth, "sweep_items": items,
            "solvers_from_same_state": cmp,
            "data": "synthetic: 1,048,576 seeded bodies, 70% spheres r U[0.1, 0.3] (8 segments), 30% boxes "
                    "half-extents U[0.1, 0.4], spawned in 200x200x50 m"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=20, help="frames after the 240 settling frames, before timing")
    ap.add_argument("--bodies", type=int, default=32768)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-sample-substeps", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-engine section")
    ap.add_argument("--batched-frames", type=int, default=60)
    ap.add_argument("--batched-warmup", type=int, default=240)
    ap.add_argument("--no-diff", action="store_true", help="skip the differentiable-rollout section")
    ap.add_argument("--no-large", action="store_true", help="skip the 1M-body section")
    ap.add_argument("--large-frames", type=int, default=120)
    ap.add_argument("--large-timed", type=int, default=20)
    ap.add_argument("--large-solver-wgs", type=int, default=0, help="sweep workgroups of the 1M-body scene")
    ap.add_argument("--no-large-compare", action="store_true",
                    help="skip the ordered-vs-coloured comparison of the 1M-body section")
    ap.add_argument("--no-warm", action="store_true", help="skip the warm-start comparison")
    ap.add_argument("--solver", default="auto", choices=["auto", "ordered", "flow", "colored"],
                    help="schedule of the contact sweep (all give the serial result bit for bit)")
    ap.add_argument("--solver-wgs", type=int, default=0, help="workgroups of the sweep kernel (0 = default)")
    ap.add_argument("--sweep-variant", default="seq", choices=["seq", "pipelined", "paired", "seq_basic", "coherent", "fenced"],
                    help="hand-off variant of the multi-workgroup sweep")
    ap.add_argument("--broadphase", default="auto", choices=["auto", "all_pairs", "lbvh", "sweep"],
                    help="broadphase method (both give the all-pairs list bit for bit)")
    ap.add_argument("--no-tri", action="store_true", help="skip the triangle-BVH section")
    ap.add_argument("--window-file", type=Path, default=None,
                    help="write the timed windows (CLOCK_MONOTONIC ns) of each section here, for seevr_amd.profsum")
    ap.add_argument("--calib", action="store_true",
                    help="also run the 256 MiB copy kernel once (byte-counter calibration under rocprofv3)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(the scene is stepped on a single device); run with --gpus 1", file=sys.stderr)
        return 2

    from seevr_amd import Context, default_params
    t0 = time.perf_counter()
    scene = build_workload(args.bodies, args.seed)
    counts = scene.counts()
    n, nv, nh = counts["num_bodies"], counts["num_vertices"], counts["num_hull_verts"]
    n_dyn = args.bodies
    log(f"scene: {n} bodies, {nh} hull verts, built in {time.perf_counter() - t0:.1f} s")
    caps = scene.caps
    caps.max_pairs = 8 * n + 1024
    caps.max_contacts = 6 * n + 1024
    ctx = Context(caps, default_params(pgs_iterations=PGS_ITERS))
    scene.upload(ctx)
    ctx.set_solver(args.solver, args.solver_wgs)
    ctx.set_broadphase(args.broadphase)
    bp_method = ctx.broadphase_method(n)
    STAGE_KERNEL["phys/broad_phase"] = BROADPHASE_KERNEL[bp_method]
    ctx.set_sweep_variant(args.sweep_variant)
    name, arch = ctx.device_info()
    log(f"device: {name} {arch}")
    copy_gbs = ctx.copy_bandwidth(256 << 20, 1 if args.calib else 20)
    log(f"plain copy kernel (256 MiB): {copy_gbs:.0f} GB/s read+write")

    # The hand-off probe needs its two one-wave workgroups resident together;
    # when a wait passes its 2 s bound the probe reports a time-out (the
    # kernel has ended), recorded here instead of ending the bench.
    probe_notes = []

    def probe(same_xcd):
        try:
            return ctx.handoff_us(20000, same_xcd=same_xcd)
        except Exception as e:  # noqa: BLE001 - a bounded probe's time-out
            probe_notes.append(f"hand-off probe ({'same' if same_xcd else 'across'} XCD) failed: {e}")
            log(probe_notes[-1])
            return None, None

    handoff, hx = probe(False)
    handoff_same, hxs = probe(True)
    if handoff is None and handoff_same is not None:
        probe_notes.append("roofline peak from the same-XCD hand-off")
        handoff = handoff_same
    if handoff is not None:
        log(f"hand-off between two waves: {handoff:.3f} us one way across XCDs (XCC {hx}), "
            f"{handoff_same if handoff_same is not None else float('nan'):.3f} us within one (XCC {hxs})")

    pre = SETTLE_FRAMES + args.warmup
    for i in range(pre):
        ctx.step(n, nv, nh, SUBSTEPS, DT, GRAVITY)
        if i % 60 == 59:
            log(f"settle/warmup frame {i + 1}/{pre}: counters {ctx.read_counters()}")
    depth0, items0 = ctx.sweep_depth()

    # Timed region: K frames, device synchronized on both sides (every call is
    # synchronous).  A frame is the physics step and the LBVH over the body
    # boxes (the workload rebuilds it every frame).  HIP events on the
    # library stream bracket the sweep (the roofline kernel) in the same
    # frames; events around every stage cost 1.5 ms per frame, so the other
    # stages are timed in frames after these.
    ctx.enable_stage_timing(True, only="phys/pgs_sweep")
    ctx.stage_times(reset=True)
    el = el_bvh = 0.0
    w0 = time.monotonic_ns()
    for i in range(args.steps):
        t0 = time.perf_counter()
        ctx.step(n, nv, nh, SUBSTEPS, DT, GRAVITY)
        t1 = time.perf_counter()
        ctx.bvh_build_bodies(n)
        t2 = time.perf_counter()
        el_bvh += t2 - t1
        el += t1 - t0
    window("plain", w0, time.monotonic_ns())
    stages = ctx.stage_times(reset=True)
    pairs, contacts = ctx.read_counters()
    depth, items = ctx.sweep_depth()
    ms = 1000.0 * (el + el_bvh) / args.steps
    # every stage, in frames after the timed ones (and one LBVH build each)
    ctx.enable_stage_timing(True)
    for i in range(STAGE_FRAMES):
        ctx.step(n, nv, nh, SUBSTEPS, DT, GRAVITY)
        ctx.bvh_build_bodies(n)
    for k, v in ctx.stage_times(reset=True).items():
        if k != "phys/pgs_sweep":
            stages[k] = v
    ctx.enable_stage_timing(False)
    value = n * SUBSTEPS * args.steps / (el + el_bvh)
    value_phys = n * SUBSTEPS * args.steps / el
    log(f"timed {args.steps} frames (frames {pre + 1}-{pre + args.steps}) in {el + el_bvh:.2f} s: {ms:.2f} ms/frame "
        f"({1000 * el / args.steps:.2f} physics + {1000 * el_bvh / args.steps:.3f} LBVH), pairs {pairs}, "
        f"contacts {contacts}, sweep DAG depth {depth0} -> {depth}")
    if args.solver == "colored":
        log(f"contacts per colour (last class = one-lane remainder): {ctx.color_counts()}")

    nc = min(contacts, caps.max_contacts)
    sweep_us = stages.get("phys/pgs_sweep", (0.0, 0))[0]
    roof = sweep_roofline(sweep_us, depth, items, handoff if handoff is not None else 0.0, nc)
    roof["dag_depth_first_timed_frame"] = depth0
    roof["handoff_us_same_xcd"] = round(handoff_same, 4) if handoff_same is not None else None
    if probe_notes:
        roof["probe_notes"] = probe_notes
    hbm_stages = {k: roofline_entry(k, stages, n, nh, n_dyn, pairs, nc)
                  for k in ("phys/aabb", "phys/broad_phase", "phys/integrate", "phys/gravity", "bvh/radix_sort",
                            "bvh/morton_codes")}
    # the all-pairs test loop is bound by vector issue, not HBM: its tests/s
    # against the non-packed FP32 issue rate (256 CUs x 4 SIMDs x 16 lanes x
    # 2.4 GHz; packed FP32 is not used), at 10 vector instructions per test
    # (six subtractions, three maxima, one compare)
    bp_us = stages.get("phys/broad_phase", (0.0, 0))[0]
    tests = n * (n - 1) / 2.0
    hbm_stages["phys/broad_phase"]["method"] = bp_method
    if bp_us > 0 and bp_method == "all_pairs":
        lane_ops = 10.0 * tests / (bp_us * 1e-6)
        hbm_stages["phys/broad_phase"]["valu"] = {
            "tests_per_launch": int(tests), "tests_per_s": round(tests / (bp_us * 1e-6), 1),
            "vector_ops_per_test": 10, "lane_ops_per_s": round(lane_ops, 1), "peak_lane_ops_per_s": VALU_PEAK,
            "frac": round(lane_ops / VALU_PEAK, 4),
            "note": "whole stage time (count pass, scan, placement, emit pass) against the test loop's issue "
                    "rate at 10 vector instructions per test (an instruction count from the kernel's source, "
                    "not a counter)"}
    lbvh = {"ms_per_frame": round(1000.0 * el_bvh / args.steps, 3), "primitives": n,
            "stage_us": {k: round(v[0], 1) for k, v in stages.items() if k.startswith("bvh/")}}
    log(f"LBVH over body boxes: {lbvh['ms_per_frame']:.3f} ms per frame; {lbvh['stage_us']}")
    roof["copy_kernel_gbs"] = round(copy_gbs, 1)
    phys = {k: v for k, v in stages.items() if k.startswith("phys/") and k != "phys/pgs_sweep"}
    total_us = sum(v[0] * v[1] for v in phys.values())
    breakdown = {k: round(100.0 * v[0] * v[1] / total_us, 1) for k, v in phys.items()}

    variants = sweep_variant_section(ctx, n, args.sweep_variant)
    tri = None if args.no_tri else triangle_bvh_section(ctx, counts)

    cpu = cpu1 = None
    if not args.no_cpu:
        log("timing the CPU oracle baseline ...")
        cpu = cpu_baseline(ctx, scene, n, args.cpu_sample_substeps)
        cpu1 = cpu_baseline(ctx, scene, n, 2, mt=False)
        log(f"cpu baseline: {cpu['value']:.0f} body-substeps/s on {cpu['cores']} threads, "
            f"{cpu1['value']:.0f} on one core")
    warm = None
    if not args.no_warm:
        warm = warm_start_section(ctx, n, nv, nh)
    ctx.close()
    scene.close()

    batched = None
    if not args.no_batched:
        batched = batched_section(args.batched_frames, args.batched_warmup)
    diff = None
    if not args.no_diff:
        diff = differentiable_section()
    large = None
    if not args.no_large:
        large = large_scene_section(args.large_frames, args.large_timed, not args.no_large_compare,
                                    args.large_solver_wgs, args.sweep_variant)

    out = {
        "metric": "body-substeps/s @32k bodies", "value": round(value, 1), "unit": "body-substeps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded scene generator)",
        "config": {"workload": "32,768 bodies (+5 static) in a 30x30 m walled container, 10 substeps, "
                               "30 PGS iterations, dt 1/60, LBVH over the body boxes every frame",
                   "bodies": n, "hull_verts": nh, "substeps": SUBSTEPS, "pgs_iterations": PGS_ITERS,
                   "frames_before_timing": pre, "timed_frames": [pre + 1, pre + args.steps],
                   "pairs_last_substep": pairs, "contacts_last_substep": contacts},
        "physics_only": {"value": round(value_phys, 1), "ms_per_frame": round(1000.0 * el / args.steps, 3)},
        "roofline": roof, "roofline_hbm_stages": hbm_stages, "cpu_baseline": cpu, "cpu_baseline_one_core": cpu1,
        "stage_share_pct": breakdown,
        "lbvh_per_frame": lbvh, "sweep_variants": variants, "triangle_bvh": tri, "warm_start": warm,
        "batched": batched, "differentiable": diff, "large_scene": large, "timed_windows_ns": WINDOWS,
    }
    if batched is not None:
        batched["body_substeps_vs_plain_32k"] = round(batched["body_substeps_per_s"] / value, 3)
    if args.window_file:
        args.window_file.parent.mkdir(parents=True, exist_ok=True)
        args.window_file.write_text(json.dumps(WINDOWS))
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
