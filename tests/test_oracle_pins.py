"""Known-answer pins of two oracle functions that the other pins leave open:
the Gauss-Seidel visiting order of the contact sweep (physics.comp:523-582)
and the bin mapping of the lookup support (physics.comp:132-161).

Sweep order.  Bodies stacked on a vertical line with contacts on that line
and normals +y: every lever arm is parallel to the normal, so r x n = 0, no
angular velocity moves and the tangential velocity is exactly zero (no
friction).  The reference's update then reduces to the scalar equations
    dl = (-(v_b - v_a) + bias) / (1/m_a + 1/m_b),  l' = max(l + dl, 0),
    v_a -= (l' - l) / m_a,  v_b += (l' - l) / m_b,
visited contact by contact in index order, iteration after iteration.
With unit inverse masses and dyadic velocities every value is exact in
float32, so the known answers below are hand-derived and bit-exact; a
reversed, Jacobi or per-iteration-shuffled sweep gives other numbers.

Lookup binning.  A one-body world whose lookup table sends every (axis, bin)
entry to a vertex far behind the query direction except one entry, which
sends it to the vertex along it.  The support returns the good vertex iff
the lookup consulted that entry, so directions built at known angles on each
axis circle (the angle of axis 0 is atan(d.z, d.y), of axis 1 atan(d.z, d.x),
of axis 2 atan(d.y, d.x), each in [0, 2 pi), 64 bins; the candidates are
bins b0 and b0 + 1 mod 64) pin which entries are read.
"""
import ctypes as C
import math
from fractions import Fraction

import numpy as np
import pytest

from harness import World, oracle
from test_oracle_physics import build

SUB_DT = float(np.float32(np.float32(1 / 60) / np.float32(10)))


# ---------------------------------------------------------------- sweep order
def stack_world(vels, contacts, iters, inv_mass=None, double=False):
    """Ground (static, body 0) and len(vels) unit cubes stacked above it at
    x = z = 0; contact k joins bodies contacts[k] = (a, b), a below b, at a
    point on the vertical line through the centres, normal +y."""
    nb = len(vels)
    bodies = [("box", (0, -0.5, 0), (5, 0.5, 5), 0.0, 0.5, 0.0)]
    bodies += [("box", (0, 0.5 + i, 0), (0.5, 0.5, 0.5), 1000.0, 0.5, 0.0) for i in range(nb)]
    sc, w = build(bodies, iters=iters)
    if double:
        w = World.from_scene(sc, w.max_pairs, w.max_contacts, iters, double=True)
    w.a["BODY_INV_MASS"][1:] = 1.0 if inv_mass is None else inv_mass
    w.a["BODY_VEL"][:] = 0.0
    for i, v in enumerate(vels):
        w.a["BODY_VEL"][3 * (i + 1) + 1] = v
    for k, (a, b) in enumerate(contacts):
        w.a["CONTACT_POS"][3 * k:3 * k + 3] = [0.0, float(b) - 0.5, 0.0]  # top face of body a
        w.a["CONTACT_NORMAL"][3 * k:3 * k + 3] = [0.0, 1.0, 0.0]
        w.a["CONTACT_PENETRATION"][k] = 0.0
        w.a["CONTACT_GEOM_A"][k] = a
        w.a["CONTACT_GEOM_B"][k] = b
    w.a["ATOMIC_COUNTERS"][:] = [len(contacts), len(contacts), 0, 0]
    w._scene = sc  # capacities and static arrays for a device context
    return w


def solve(w):
    w.oracle_stage(4)
    w.oracle_stage(5, dt=SUB_DT)
    nb = w.n
    v = w.a["BODY_VEL"].reshape(nb, 3)
    return v[:, 1].copy(), w.a["CONTACT_LAMBDA_N"][: w.ncontacts()].copy(), v, w.a["BODY_OMEGA"].copy()


def scalar_gs(vels, contacts, iters, inv_mass=None):
    """Exact rational restatement of the reduced equations (module docstring)."""
    v = [Fraction(0)] + [Fraction(x) for x in vels]
    im = [Fraction(0)] + [Fraction(x) for x in (inv_mass if inv_mass is not None else [1] * len(vels))]
    lam = [Fraction(0)] * len(contacts)
    for _ in range(iters):
        for k, (a, b) in enumerate(contacts):
            dl = -(v[b] - v[a]) / (im[a] + im[b])
            nl = max(lam[k] + dl, Fraction(0))
            dl = nl - lam[k]
            lam[k] = nl
            v[a] -= im[a] * dl
            v[b] += im[b] * dl
    return v, lam


CHAIN = [(0, 1), (1, 2), (2, 3)]


def test_sweep_order_known_answer_one_iteration():
    # v = (-1, -2, -3): c0 stops body 1 (l = 1), c1 then closes 0 - (-2) with
    # k = 2 (l = 1, v1 = -1, v2 = -1), c2 closes -3 - (-1) (l = 1).
    vy, lam, v, om = solve(stack_world([-1.0, -2.0, -3.0], CHAIN, 1))
    assert vy.tolist() == [0.0, -1.0, -2.0, -2.0]
    assert lam.tolist() == [1.0, 1.0, 1.0]
    assert not om.any() and not v[:, [0, 2]].any()
    # the reversed sweep would give v = (0, -1.75, -2.5), lambda = (1.75, 0.75, 0.5);
    # a Jacobi sweep lambda = (1, 0.5, 0.5)


def test_sweep_order_known_answer_two_iterations():
    vy, lam, _, _ = solve(stack_world([-1.0, -2.0, -3.0], CHAIN, 2))
    assert vy.tolist() == [0.0, -1.0, -1.5, -1.5]
    assert lam.tolist() == [2.0, 2.0, 1.5]


def test_sweep_follows_contact_index_not_geometry():
    # the same chain listed top contact first: c0 = (2, 3) is visited first
    vy, lam, _, _ = solve(stack_world([-1.0, -2.0, -3.0], [(2, 3), (1, 2), (0, 1)], 1))
    assert vy.tolist() == [0.0, 0.0, -1.75, -2.5]
    assert lam.tolist() == [0.5, 0.75, 1.75]


def test_sweep_clamps_separating_contacts():
    # body 2 moves up away from body 1: c1 would pull (dl = -1.5) and is clamped
    vy, lam, _, _ = solve(stack_world([-1.0, 3.0, 0.0], CHAIN, 1))
    assert vy.tolist() == [0.0, 0.0, 1.5, 1.5]
    assert lam.tolist() == [1.0, 0.0, 1.5]


@pytest.mark.parametrize("seed", range(6))
def test_sweep_matches_exact_scalar_gauss_seidel(seed):
    """Longer stacks, shuffled contact lists, several iterations, and (seeds
    3-5) inverse masses 1 and 1/2 mixed: bit-exact against exact rationals
    while every k is 1 or 2 (all values dyadic), else within 1e-5."""
    rng = np.random.default_rng(seed)
    nb = int(rng.integers(4, 8))
    contacts = [(i, i + 1) for i in range(nb)]
    rng.shuffle(contacts)
    contacts = [tuple(int(x) for x in c) for c in contacts]
    vels = (rng.integers(-16, 17, nb) / 4.0).tolist()
    iters = int(rng.integers(1, 5))
    im = [1.0] * nb if seed < 3 else rng.choice([1.0, 0.5], nb).tolist()
    vy, lam, _, _ = solve(stack_world(vels, contacts, iters, inv_mass=im))
    v_ex, l_ex = scalar_gs(vels, contacts, iters, inv_mass=im)
    if seed < 3:  # k in {1, 2}: every value dyadic
        assert vy.tolist() == [float(x) for x in v_ex]
        assert lam.tolist() == [float(x) for x in l_ex]
    else:
        np.testing.assert_allclose(vy, [float(x) for x in v_ex], rtol=0, atol=1e-5)
        np.testing.assert_allclose(lam, [float(x) for x in l_ex], rtol=0, atol=1e-5)


def test_sweep_bias_baumgarte_restitution_fp64():
    """Penetration and restitution bias (physics.comp:505-521), computed from
    the pre-sweep velocities, in the FP64 oracle: the reduced equations with
    bias = 0.2 max(pen - 0.005, 0) / dt + max(e_a, e_b) max(-vn, 0)."""
    rng = np.random.default_rng(11)
    nb = 5
    contacts = [(0, 1), (2, 3), (1, 2), (4, 5), (3, 4)]
    vels = rng.uniform(-3, 1, nb).tolist()
    im = rng.uniform(0.2, 2.0, nb).tolist()
    e = rng.uniform(0.0, 0.6, nb + 1)
    pen = rng.uniform(0.0, 0.05, len(contacts))
    w = stack_world(vels, contacts, 3, inv_mass=im, double=True)
    w.a["GEOM_RESTITUTION"][:] = e
    w.a["CONTACT_PENETRATION"][: len(contacts)] = pen
    vy, lam, _, om = solve(w)
    v = [0.0] + list(vels)
    imv = [0.0] + list(im)
    bias = [0.2 * max(p - 0.005, 0.0) / SUB_DT + max(e[a], e[b]) * max(-(v[b] - v[a]), 0.0)
            for p, (a, b) in zip(pen, contacts)]
    lamr = [0.0] * len(contacts)
    for _ in range(3):
        for k, (a, b) in enumerate(contacts):
            dl = (-(v[b] - v[a]) + bias[k]) / (imv[a] + imv[b])
            nl = max(lamr[k] + dl, 0.0)
            dl = nl - lamr[k]
            lamr[k] = nl
            v[a] -= imv[a] * dl
            v[b] += imv[b] * dl
    np.testing.assert_allclose(vy, v, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(lam, lamr, rtol=1e-12, atol=1e-12)
    assert not om.any()


# -------------------------------------------------------------- lookup binning
def lookup_world():
    sc, w = build([("sphere", (0, 0, 0), 0.5, 16, 1000.0, 0.5, 0.0)])
    assert int(w.a["SUPPORT_LOOKUP_OFFSET"][0]) == 1  # the lookup path
    return w


def support_of(w, d):
    out = (C.c_float * 3)()
    oracle().support(C.byref(w.bufs()), 0, (C.c_float * 3)(*d), out)
    return np.array(out[:])


def direction(axis, theta, off=0.3):
    c, s = math.cos(theta), math.sin(theta)
    # the two coordinates axis `axis` measures its angle from, (cos, sin) order
    d = {0: (off, c, s), 1: (c, off, s), 2: (c, s, off)}[axis]
    return np.array(d) / np.linalg.norm(d)


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("b", [0, 5, 17, 31, 32, 50, 63])
@pytest.mark.parametrize("frac", [0.02, 0.5, 0.98])
def test_lookup_reads_bins_b0_and_next(axis, b, frac):
    w = lookup_world()
    vs = int(w.a["GEOM_DATA"][0])
    theta = (b + frac) * 2 * math.pi / 64
    if theta > math.pi:
        theta -= 2 * math.pi  # atan returns the negative angle; the lookup adds 2 pi
    d = direction(axis, theta)
    hull = w.a["HULL_VERTS"]
    hull[3 * vs:3 * vs + 3] = (d * 0.5).astype(np.float32)         # vertex 0: along d
    hull[3 * vs + 3:3 * vs + 6] = (-d * 0.5).astype(np.float32)    # vertex 1: behind
    table = w.a["SUPPORT_LOOKUP"]
    for g in (b - 1, b, b + 1, b + 2):
        table[1:193] = 1
        table[1 + axis * 64 + g % 64] = 0
        got = support_of(w, d)
        want_good = g % 64 in (b % 64, (b + 1) % 64)
        assert np.allclose(got, 0.5 * d if want_good else -0.5 * d, atol=1e-6), (axis, b, frac, g, got)
    # the same bin on another axis circle is not a candidate (that axis reads
    # the bins around its own angle; pick the other axis' bin opposite it)
    other = (axis + 1) % 3
    u, v = {0: (d[1], d[2]), 1: (d[0], d[2]), 2: (d[0], d[1])}[other]
    own = int((math.atan2(v, u) % (2 * math.pi)) / (2 * math.pi) * 64) % 64
    table[1:193] = 1
    table[1 + other * 64 + (own + 32) % 64] = 0
    assert np.allclose(support_of(w, d), -0.5 * d, atol=1e-6)
    table[1:193] = 1
    table[1 + other * 64 + own] = 0  # ... while its own bin is
    assert np.allclose(support_of(w, d), 0.5 * d, atol=1e-6)
