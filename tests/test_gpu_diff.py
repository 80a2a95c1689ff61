"""Differentiable batched rollout: device gradients (FP32 rollout, FP64
reverse pass) against the CPU adjoint of the same fixed-contact map, in a
contact-stable regime (resting piles given small sideways kicks, so
contacts persist and some slide).  Bar: 1e-3 relative, per block (state
parts, per-geom friction, per-geom restitution) against the exact
derivative of the FP32 computation (the oracle adjoint along the FP32
rollout with the FP32 sweep's branch decisions, FP64 reverse), and for the
state blocks and the whole gradient against the FP64 adjoint of the FP64
rollout; a few inputs against FP64 central differences.  Also: the
torch.autograd.Function gives the same gradients as the C ABI, forward +
backward leaves the batch state at the end of the forward run, and a second
backward pass gives the same bits.
"""
import numpy as np
import pytest

from adjoint_harness import STATE, Problem, record_base, record_fp32
from harness import World, assert_bitwise

pytestmark = pytest.mark.gpu

SEG = 8
SUB = 10
DT = 1.0 / 60.0
FRAMES = 2
SEED = 7
WIDTH = {"pos": 3, "quat": 4, "vel": 3, "omega": 3}


def make_batch(n_scenes=4, min_n=6, max_n=10):
    from seevr_amd import Batch, Scene, default_caps, default_params
    from seevr_amd._lib import BatchCaps
    big = default_caps(max_bodies=n_scenes * 300, max_hull_verts=n_scenes * 300 * 82,
                       max_vertices=n_scenes * 300 * 82, max_triangles=n_scenes * 300 * 130,
                       max_support_entries=n_scenes * 300 * 193 + 1)
    sc = Scene(big)
    start, gs = sc.build_batch(0, n_scenes, min_n, max_n, SEG, SEED, top=2.5)
    cnt = sc.counts()
    bc = BatchCaps(max_scenes=n_scenes, max_bodies=cnt["num_bodies"], max_hull_verts=cnt["num_hull_verts"],
                   max_vertices=cnt["num_vertices"], max_triangles=max(cnt["num_triangles"], 1),
                   max_support_entries=cnt["num_support_entries"], max_scene_bodies=512, max_scene_pairs=1024,
                   max_scene_contacts=512)
    b = Batch(bc, default_params(pgs_iterations=30))
    sc.upload(b.ctx)
    grav = np.zeros((n_scenes, 3), np.float32)
    grav[:, 1] = np.float32(-9.81) * gs
    b.set_scenes(start, grav)
    n = b.num_bodies
    for _ in range(90):  # settle into resting piles
        b.step(cnt["num_vertices"], SUB, DT)
    rng = np.random.default_rng(SEED)
    x = b.packed_state()
    inv_mass = b.ctx.download("BODY_INV_MASS", n)
    vel = x[7 * n:10 * n].reshape(n, 3)
    dyn = inv_mass > 0
    vel[dyn, 0] += rng.normal(0, 0.3, dyn.sum()).astype(np.float32)
    vel[dyn, 2] += rng.normal(0, 0.3, dyn.sum()).astype(np.float32)
    b.set_packed_state(x)
    fric = rng.uniform(0.2, 0.8, n).astype(np.float32)  # distinct per geom: no max() ties
    rest = rng.uniform(0.0, 0.5, n).astype(np.float32)
    b.ctx.upload("GEOM_FRICTION", fric)
    b.ctx.upload("GEOM_RESTITUTION", rest)
    b.diff_reserve(FRAMES, SUB)
    return b, start, grav, fric, rest


def cpu_problem(k, start, grav, x0_packed, fric, rest, n_all):
    from seevr_amd import Scene, default_caps
    b0, b1 = int(start[k]), int(start[k + 1])
    nk = b1 - b0
    sc = Scene(default_caps(max_bodies=300, max_hull_verts=300 * 82, max_vertices=300 * 82, max_triangles=300 * 130,
                            max_support_entries=300 * 193 + 1))
    sc.build_batch(k, 1, 6, 10, SEG, SEED, top=2.5)
    w = World.from_scene(sc, max_pairs=4000, max_contacts=2000, pgs_iterations=30, double=True)
    assert w.n == nk
    off = 0
    for key, name in (("pos", "BODY_POS"), ("quat", "BODY_QUAT"), ("vel", "BODY_VEL"), ("omega", "BODY_OMEGA")):
        wd = WIDTH[key]
        w.a[name][: wd * nk] = x0_packed[off + wd * b0: off + wd * b1]
        off += wd * n_all
    w.a["GEOM_FRICTION"][:nk] = fric[b0:b1]
    w.a["GEOM_RESTITUTION"][:nk] = rest[b0:b1]
    dt_sub = float(np.float32(DT) / np.float32(SUB))
    g = tuple(float(v) for v in grav[k])
    # the FP32 rollout (the device's, bit for bit) and its branch decisions
    w32 = World.from_scene(sc, max_pairs=4000, max_contacts=2000, pgs_iterations=30)
    for name in ("BODY_POS", "BODY_QUAT", "BODY_VEL", "BODY_OMEGA", "GEOM_FRICTION", "GEOM_RESTITUTION"):
        w32.a[name][:] = w.a[name].astype(np.float32)
    _, c32, s32, starts32 = record_fp32(w32, FRAMES * SUB, dt_sub, g)
    same = (Problem(s32, c32, dt_sub, g, iters=30), starts32, s32)
    x0, contacts, statics = record_base(w, FRAMES * SUB, dt_sub, g)
    return Problem(statics, contacts, dt_sub, g, iters=30), x0, statics, same


def split(packed, n):
    out, off = {}, 0
    for key in STATE:
        out[key] = packed[off: off + WIDTH[key] * n]
        off += WIDTH[key] * n
    return out


def rel_err(got, want):
    return float(np.linalg.norm(got - want) / max(np.linalg.norm(want), 1e-30))


def test_gpu_gradients_match_fp64_adjoint_and_finite_differences():
    b, start, grav, fric, rest = make_batch()
    n = b.num_bodies
    x0 = b.packed_state()
    rng = np.random.default_rng(11)
    wts = rng.normal(size=13 * n).astype(np.float32)
    b.diff_forward(FRAMES, SUB, DT)
    xf = b.packed_state()
    g_state, g_fric, g_rest = b.diff_backward(FRAMES, wts, SUB, DT)
    assert np.array_equal(b.packed_state(), xf), "backward must leave the forward's final state"
    gs = split(g_state, n)
    ws = split(wts.astype(np.float64), n)
    worst = 0.0
    for k in range(b.num_scenes):
        prob, x0k, statics, (prob32, starts32, s32) = cpu_problem(k, start, grav, x0, fric, rest, n)
        assert sum(c["nc"] for c in prob.contacts) > 0
        b0, b1 = int(start[k]), int(start[k + 1])
        wk = {key: ws[key][WIDTH[key] * b0: WIDTH[key] * b1] for key in STATE}
        adj, af, ae = prob.grad(x0k, statics["fric"], statics["rest"], wk)
        adj32, af32, ae32 = prob32.grad_along(starts32, s32["fric"], s32["rest"], wk, forward32=True)
        got_all, want_all = [], []
        for key in STATE:
            got = gs[key][WIDTH[key] * b0: WIDTH[key] * b1].astype(np.float64)
            e32 = rel_err(got, adj32[key])
            e = rel_err(got, adj[key])
            worst = max(worst, e)
            print(f"scene {k} d/d{key}: {e32:.1e} vs the FP32 computation's derivative, {e:.1e} vs FP64")
            assert e32 < 1e-3, f"scene {k} d/d{key}: relative error {e32:.2e} (same FP32 computation)"
            assert e < 1e-3, f"scene {k} d/d{key}: relative error {e:.2e} (FP64 rollout)"
            got_all.append(got)
            want_all.append(adj[key])
        # Per-geom material gradients, summed per geom in FP64 in incidence
        # order (deterministic).  Friction gradients here are ~1e-4 of the
        # others and a few sweep items that FP32 and FP64 evaluate on
        # different sides of a stick/slip or tangent threshold move them by
        # ~1e-3 of their own norm, so the per-block bar is against the same
        # FP32 computation's derivative.
        for name, got, want32, want in (("friction", g_fric[b0:b1], af32, af),
                                        ("restitution", g_rest[b0:b1], ae32, ae)):
            e32 = rel_err(got.astype(np.float64), want32)
            e = rel_err(got.astype(np.float64), want)
            print(f"scene {k} d/d{name}: {e32:.1e} vs the FP32 computation's derivative, {e:.1e} vs FP64")
            assert e32 < 1e-3, f"scene {k} d/d{name}: relative error {e32:.2e} (same FP32 computation)"
            got_all.append(got.astype(np.float64))
            want_all.append(want)
        e_all = rel_err(np.concatenate(got_all), np.concatenate(want_all))
        worst = max(worst, e_all)
        assert e_all < 1e-3, f"scene {k}: whole-gradient relative error {e_all:.2e}"
        if k == 0:  # FP64 central differences of the same map, a few inputs
            def loss(x, fr, re):
                xf_, *_ = prob.forward(x, fr, re)
                return sum(float(np.dot(wk[kk], xf_[kk])) for kk in STATE)
            h = 1e-6
            dyn = np.nonzero(statics["inv_mass"] > 0)[0][:2]
            for body in dyn:
                for key in ("vel", "omega", "pos"):
                    i = WIDTH[key] * body
                    xp = {kk: v.copy() for kk, v in x0k.items()}
                    xm = {kk: v.copy() for kk, v in x0k.items()}
                    xp[key][i] += h
                    xm[key][i] -= h
                    fd = (loss(xp, statics["fric"], statics["rest"]) - loss(xm, statics["fric"], statics["rest"])) / (2 * h)
                    got = float(gs[key][WIDTH[key] * (b0 + body)])
                    scale = max(abs(fd), 1e-2 * np.abs(adj[key]).max())
                    assert abs(got - fd) < 1e-3 * scale, f"d/d{key}[{i}]: gpu {got} fd {fd}"
    print(f"worst relative gradient error vs FP64 adjoint: {worst:.2e}")
    # deterministic: a second backward pass gives the same bits
    b.set_packed_state(x0)
    b.diff_forward(FRAMES, SUB, DT)
    again = b.diff_backward(FRAMES, wts, SUB, DT)
    for got, first, name in zip(again, (g_state, g_fric, g_rest), ("state", "friction", "restitution")):
        assert_bitwise(got, first, f"repeated backward, d/d{name}")
    b.close()


def test_torch_function_matches_c_abi():
    torch = pytest.importorskip("torch")
    from seevr_amd.diff import rollout
    b, start, grav, fric, rest = make_batch(n_scenes=2)
    n = b.num_bodies
    x0 = b.packed_state()
    rng = np.random.default_rng(5)
    wts = rng.normal(size=13 * n).astype(np.float32)
    s0 = torch.tensor(x0, requires_grad=True)
    fr = torch.tensor(fric, requires_grad=True)
    re = torch.tensor(rest, requires_grad=True)
    out = rollout(b, s0, fr, re, FRAMES, SUB, DT)
    loss = (out * torch.from_numpy(wts)).sum()
    loss.backward()
    b.set_packed_state(x0)
    b.ctx.upload("GEOM_FRICTION", fric)
    b.ctx.upload("GEOM_RESTITUTION", rest)
    b.diff_forward(FRAMES, SUB, DT)
    assert np.array_equal(b.packed_state(), out.detach().numpy())
    g_state, g_fric, g_rest = b.diff_backward(FRAMES, wts, SUB, DT)
    assert np.array_equal(s0.grad.numpy(), g_state)
    # per-geom material sums are FP64 in incidence order: the same bits
    assert np.array_equal(fr.grad.numpy(), g_fric)
    assert np.array_equal(re.grad.numpy(), g_rest)
    b.close()
