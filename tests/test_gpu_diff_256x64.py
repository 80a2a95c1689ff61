"""The differentiable workload at its own size (tests/c4_workload.py): 256
scenes x 64 bodies over a ground, 60 frames x 10 substeps, v and w ~ N(0, 1),
per-material friction and restitution, loss = MSE of the final positions to
seeded targets.  The device runs sv_batch_diff_forward / _backward over all
256 scenes (FP32 rollout, FP64 reverse pass).  Checks, per scene:
  - the device rollout's final positions equal the FP32 oracle's bit for bit;
  - the gradient blocks (initial velocities, angular velocities,
    per-material friction and restitution) are within 1e-3 relative of the
    exact derivative of the same FP32 computation (the oracle adjoint along
    the same FP32 trajectory with the FP32 sweep's own branch decisions and
    an FP64 reverse pass): on the 16 contact-stable scenes of
    tests/golden/c4_smooth_scenes.json and on 8 others;
  - on the 16 contact-stable scenes (chosen on the CPU without looking at a
    device gradient: smooth at the central-difference step, and a gradient
    that does not depend on the working precision), also within 1e-3 of the
    FP64 adjoint of the FP64 rollout and of FP64 central differences
    (eps 1e-4) along the scene's seeded directions.
"""
import json
from pathlib import Path

import numpy as np
import pytest

from c4_workload import BODIES, DT, FRAMES, SCENES, SEED, SEG, SUB, TOP, FixedContactMap, scene_inputs
from harness import assert_bitwise

pytestmark = pytest.mark.gpu

SEL = json.loads((Path(__file__).resolve().parent / "golden" / "c4_smooth_scenes.json").read_text())


def rel(got, want):
    return float(np.linalg.norm(np.asarray(got) - np.asarray(want)) / max(np.linalg.norm(want), 1e-30))


def test_diff_256x64_60_frames_against_fp64_adjoint_and_fd():
    from seevr_amd import Batch, Scene, default_caps, default_params
    from seevr_amd._lib import BatchCaps
    m = BODIES + 1
    sc = Scene(default_caps(max_bodies=SCENES * m, max_hull_verts=SCENES * m * 290, max_vertices=SCENES * m * 290,
                            max_triangles=SCENES * m * 520, max_support_entries=SCENES * m * 193 + 1))
    start, gs = sc.build_batch(0, SCENES, BODIES, BODIES, SEG, SEED, top=TOP)
    cnt = sc.counts()
    bc = BatchCaps(max_scenes=SCENES, max_bodies=cnt["num_bodies"], max_hull_verts=cnt["num_hull_verts"],
                   max_vertices=cnt["num_vertices"], max_triangles=max(cnt["num_triangles"], 1),
                   max_support_entries=cnt["num_support_entries"], max_scene_bodies=512, max_scene_pairs=2048,
                   max_scene_contacts=512)
    b = Batch(bc, default_params(pgs_iterations=30))
    sc.upload(b.ctx)
    sc.close()
    grav = np.zeros((SCENES, 3), np.float32)
    grav[:, 1] = np.float32(-9.81) * gs
    b.set_scenes(start, grav)
    n = b.num_bodies
    inv_mass = b.ctx.download("BODY_INV_MASS", n)
    x0 = b.packed_state()
    fric = b.ctx.download("GEOM_FRICTION", n)
    rest = b.ctx.download("GEOM_RESTITUTION", n)
    inputs = []
    for k in range(SCENES):
        b0, b1 = int(start[k]), int(start[k + 1])
        inp = scene_inputs(k, inv_mass[b0:b1])
        x0[7 * n + 3 * b0: 7 * n + 3 * b1] = inp["vel"].ravel()
        x0[10 * n + 3 * b0: 10 * n + 3 * b1] = inp["omega"].ravel()
        fric[b0:b1], rest[b0:b1] = inp["fric"], inp["rest"]
        inputs.append(inp)
    b.ctx.upload("GEOM_FRICTION", fric)
    b.ctx.upload("GEOM_RESTITUTION", rest)
    b.set_packed_state(x0)
    b.diff_reserve(FRAMES, SUB)
    b.diff_forward(FRAMES, SUB, DT)
    xf = b.packed_state()
    g = np.zeros(13 * n, np.float32)
    for k in range(SCENES):
        b0, b1 = int(start[k]), int(start[k + 1])
        target = xf[3 * b0: 3 * b1].astype(np.float64) + inputs[k]["offsets"]
        g[3 * b0: 3 * b1] = 2.0 * (xf[3 * b0: 3 * b1] - target) / (3 * (b1 - b0))
    g_state, g_fric, g_rest = b.diff_backward(FRAMES, g, SUB, DT)
    c = b.counters()
    print(f"contacts in the last substep: mean {c[:, 1].mean():.1f}, max {c[:, 1].max()}", flush=True)
    worst = {}
    others = [k for k in range(8, 64, 7) if k not in SEL["scenes"]][:8]
    for k in SEL["scenes"] + others:
        stable = k in SEL["scenes"]
        b0, b1 = int(start[k]), int(start[k + 1])
        cm = FixedContactMap(k)
        assert_bitwise(xf[3 * b0: 3 * b1], cm.final32, f"scene {k} final positions vs the FP32 oracle")
        mat = cm.inp["mat"]
        gv = g_state[7 * n + 3 * b0: 7 * n + 3 * b1].astype(np.float64)
        gw = g_state[10 * n + 3 * b0: 10 * n + 3 * b1].astype(np.float64)
        gfm = np.bincount(mat, g_fric[b0:b1].astype(np.float64), 2)
        gem = np.bincount(mat, g_rest[b0:b1].astype(np.float64), 2)
        dev = {"vel": gv, "omega": gw, "friction": gfm, "restitution": gem}
        same_point = cm.blocks(*cm.grad_at_fp32(forward32=True))
        errs = {f"{key} (same point)": rel(dev[key], same_point[key]) for key in dev}
        if stable:
            fp64 = cm.blocks(*cm.grad())
            errs.update({f"{key} (FP64 trajectory)": rel(dev[key], fp64[key]) for key in dev})
            # A directional difference is compared relative to the larger of
            # itself and the typical size of a random projection of the
            # gradient, |g| |d| / sqrt(dim): a seeded direction nearly
            # orthogonal to the gradient (cosine 5e-3 where 1/sqrt(dim) is
            # 5e-2) would otherwise turn an error of 3e-5 |g| |d| into 6e-3.
            gs_all, gm_all = np.concatenate([gv, gw]), np.concatenate([gfm, gem])
            for d in cm.inp["dirs"]:
                fs, fm = cm.fd(d)
                ds = np.concatenate([d["vel"], d["omega"]])
                dm = np.concatenate([d["fric_m"], d["rest_m"]])
                an_s, an_m = float(np.dot(gs_all, ds)), float(np.dot(gm_all, dm))
                sc_s = max(abs(fs), np.linalg.norm(gs_all) * np.linalg.norm(ds) / np.sqrt(np.count_nonzero(ds)))
                sc_m = max(abs(fm), np.linalg.norm(gm_all) * np.linalg.norm(dm) / np.sqrt(np.count_nonzero(dm)))
                errs["fd state"] = max(errs.get("fd state", 0.0), abs(an_s - fs) / sc_s)
                errs["fd material"] = max(errs.get("fd material", 0.0), abs(an_m - fm) / sc_m)
        print(f"scene {k}{' (stable)' if stable else ''}: " + ", ".join(f"{key} {v:.1e}" for key, v in errs.items()),
              flush=True)
        for key, v in errs.items():
            worst[key] = max(worst.get(key, 0.0), v)
            assert v < 1e-3, f"scene {k} {key}: relative error {v:.2e}"
    print("worst relative errors: " + ", ".join(f"{k} {v:.2e}" for k, v in worst.items()))
    b.close()
