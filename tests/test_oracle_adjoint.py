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


def one_contact_problem(mu: float, kick: float):
    """One dynamic body settled on the ground (one contact per substep),
    per-geom friction mu, a sideways kick; one substep of the fixed-contact
    map."""
    sc = make_scene(4, 1, seed=3, segments=4)
    w = World.from_scene(sc, max_pairs=4000, max_contacts=2000, pgs_iterations=30, double=True)
    for _ in range(60):
        w.oracle_step(SUB, DT, G)
    n = w.n
    w.a["GEOM_FRICTION"][:n] = mu
    w.a["GEOM_RESTITUTION"][:n] = np.linspace(0.1, 0.3, n)
    dyn = w.a["BODY_INV_MASS"][:n] > 0
    w.a["BODY_VEL"][: 3 * n].reshape(n, 3)[dyn, 0] += kick
    dt_sub = float(np.float32(DT) / np.float32(SUB))
    x0, contacts, statics = record_base(w, 1, dt_sub, G)
    assert [c["nc"] for c in contacts] == [1]
    return Problem(statics, contacts, dt_sub, G, iters=30), x0, statics


def central_difference(prob, x0, fric, rest, wts, key, i, h=1e-6):
    if key in STATE:
        xp = {k: v.copy() for k, v in x0.items()}
        xm = {k: v.copy() for k, v in x0.items()}
        xp[key][i] += h
        xm[key][i] -= h
        return (loss(prob, xp, fric, rest, wts) - loss(prob, xm, fric, rest, wts)) / (2 * h)
    fp, fm = fric.copy(), fric.copy()
    fp[i] += h
    fm[i] -= h
    return (loss(prob, x0, fp, rest, wts) - loss(prob, x0, fm, rest, wts)) / (2 * h)


def test_linearizing_at_the_converged_sweep_loses_the_applied_impulse():
    """Implicit differentiation at the Gauss-Seidel fixed point (every
    reversed pass linearized at the last pass's inputs, adj_set_implicit)
    against the exact reverse and FP64 central differences, for one
    frictionless contact that the first pass already solves.  The velocity
    adjoints agree: the item map is affine in the velocities and the impulse
    while the projection and the friction branch keep their sides.  The
    position and rotation adjoints do not: at the fixed point the normal
    update is zero, so the derivative of the impulse that pass 0 applied
    through the lever arm, the rotation and k_n is gone.  The unrolled
    reverse, the library's adjoint, matches the differences."""
    prob, x0, st = one_contact_problem(mu=0.0, kick=0.0)
    fric, rest = st["fric"].copy(), st["rest"].copy()
    rng = np.random.default_rng(1)
    wts = {k: rng.normal(size=x0[k].size) for k in STATE}
    gu, _, _ = prob.grad(x0, fric, rest, wts)
    gi, _, _ = prob.grad(x0, fric, rest, wts, implicit=True)
    for k in ("vel", "omega"):
        np.testing.assert_allclose(gi[k], gu[k], rtol=1e-12, atol=1e-12)
    body = int(np.nonzero(st["inv_mass"] > 0)[0][0])
    worst_implicit = 0.0
    for key, width in (("pos", 3), ("quat", 4)):
        for comp in range(width):
            i = width * body + comp
            fd = central_difference(prob, x0, fric, rest, wts, key, i)
            scale = max(abs(fd), 1e-3)
            assert abs(gu[key][i] - fd) < 1e-6 * scale + 1e-9, f"unrolled d/d{key}[{i}] {gu[key][i]} vs fd {fd}"
            worst_implicit = max(worst_implicit, abs(gi[key][i] - fd) / scale)
    assert worst_implicit > 1e-2, "the fixed-point linearization should miss the applied impulse's derivative"


def test_linearizing_at_the_converged_sweep_drops_the_friction():
    """A sliding contact whose friction (recomputed each pass, never
    accumulated, physics.comp:560-580) acts in the early passes: the exact
    reverse's friction gradient matches FP64 central differences, and the
    fixed-point linearization's is off by more than half of it (the friction
    branch of the last pass is not what acted in the first)."""
    prob, x0, st = one_contact_problem(mu=0.5, kick=0.6)
    fric, rest = st["fric"].copy(), st["rest"].copy()
    rng = np.random.default_rng(2)
    wts = {k: rng.normal(size=x0[k].size) for k in STATE}
    _, af_u, _ = prob.grad(x0, fric, rest, wts)
    _, af_i, _ = prob.grad(x0, fric, rest, wts, implicit=True)
    fd = np.array([central_difference(prob, x0, fric, rest, wts, "fric", i) for i in range(prob.n)])
    assert np.abs(fd).max() > 1e-5, "the contact should slide under friction"
    np.testing.assert_allclose(af_u, fd, rtol=1e-4)  # central differences of a ~1e-4 gradient: ~1e-5
    assert np.linalg.norm(af_i - fd) > 0.5 * np.linalg.norm(fd)
