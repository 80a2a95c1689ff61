"""The FP64 CPU adjoint of the fixed-contact substep map (oracle/sv_adjoint.c)
against FP64 central finite differences of the same map, and the map itself
against the double-precision oracle's own substeps.  These pin the adjoint
derivation that the library's differentiable step implements on the GPU.
"""
import numpy as np
import pytest

from adjoint_harness import STATE, Problem, record_base
from harness import World, make_scene

SUB = 10
DT = 1.0 / 60.0
G = (0.0, -9.81, 0.0)


def base_problem(n_bodies=6, seed=4, settle=30, substeps=SUB, kick=True):
    sc = make_scene(4, n_bodies, seed=seed, segments=4)
    w = World.from_scene(sc, max_pairs=4000, max_contacts=2000, pgs_iterations=30, double=True)
    for _ in range(settle):
        w.oracle_step(SUB, DT, G)
    rng = np.random.default_rng(seed)
    n = w.n
    # distinct per-geom materials: e = max(e_a, e_b) has a kink at ties
    w.a["GEOM_FRICTION"][:n] = rng.uniform(0.2, 0.8, n)
    w.a["GEOM_RESTITUTION"][:n] = rng.uniform(0.0, 0.5, n)
    if kick:  # sideways velocities so that some contacts slide
        v = w.a["BODY_VEL"][: 3 * n].reshape(n, 3)
        dyn = w.a["BODY_INV_MASS"][:n] > 0
        v[dyn, 0] += rng.normal(0, 0.5, dyn.sum())
        v[dyn, 2] += rng.normal(0, 0.5, dyn.sum())
    ref = w.copy()
    dt_sub = float(np.float32(DT) / np.float32(SUB))
    x0, contacts, statics = record_base(w, substeps, dt_sub, G)
    prob = Problem(statics, contacts, dt_sub, G, iters=30)
    return prob, x0, statics, w, ref


def loss(prob, x0, fric, rest, wts):
    xf, *_ = prob.forward(x0, fric, rest)
    return sum(float(np.dot(wts[k], xf[k])) for k in STATE)


def test_fixed_contact_map_equals_oracle_substeps():
    prob, x0, statics, w_after, _ = base_problem()
    assert sum(c["nc"] for c in prob.contacts) > 10, "the scene should be in contact"
    xf, *_ = prob.forward(x0, statics["fric"], statics["rest"])
    n = prob.n
    for k, name, width in (("pos", "BODY_POS", 3), ("quat", "BODY_QUAT", 4), ("vel", "BODY_VEL", 3),
                           ("omega", "BODY_OMEGA", 3)):
        np.testing.assert_allclose(xf[k], w_after.a[name][: width * n], rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("seed", [4, 9])
def test_adjoint_matches_finite_differences(seed):
    prob, x0, statics, _, _ = base_problem(seed=seed)
    n = prob.n
    rng = np.random.default_rng(100 + seed)
    wts = {k: rng.normal(size=x0[k].size) for k in STATE}
    fric, rest = statics["fric"].copy(), statics["rest"].copy()
    adj, af, ae = prob.grad(x0, fric, rest, wts)
    dyn = np.nonzero(statics["inv_mass"] > 0)[0]
    h = 1e-6
    checks = []
    for b in dyn[:4]:
        for k, width in (("vel", 3), ("omega", 3), ("pos", 3), ("quat", 4)):
            for comp in range(width):
                i = width * b + comp
                xp = {kk: v.copy() for kk, v in x0.items()}
                xm = {kk: v.copy() for kk, v in x0.items()}
                xp[k][i] += h
                xm[k][i] -= h
                fd = (loss(prob, xp, fric, rest, wts) - loss(prob, xm, fric, rest, wts)) / (2 * h)
                checks.append((f"{k}[{i}]", fd, adj[k][i]))
    for name, arr, g in (("fric", fric, af), ("rest", rest, ae)):
        for i in range(n):
            p, m = arr.copy(), arr.copy()
            p[i] += h
            m[i] -= h
            args_p = (p, rest) if name == "fric" else (fric, p)
            args_m = (m, rest) if name == "fric" else (fric, m)
            fd = (loss(prob, x0, *args_p, wts) - loss(prob, x0, *args_m, wts)) / (2 * h)
            checks.append((f"{name}[{i}]", fd, g[i]))
    fd = np.array([c[1] for c in checks])
    ad = np.array([c[2] for c in checks])
    scale = np.abs(fd).max()
    assert scale > 1e-3
    err = np.abs(ad - fd)
    bad = [(c[0], c[1], c[2]) for c, e in zip(checks, err) if e > 1e-5 * max(abs(c[1]), scale * 1e-2)]
    assert not bad, f"adjoint vs finite differences: {bad[:6]}"
    assert np.count_nonzero(np.abs(np.array([c[1] for c in checks if c[0].startswith('fric')])) > 1e-9) > 0, \
        "some contact should slide (friction gradient non-zero)"
