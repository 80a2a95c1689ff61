"""Warm start in the oracle (an extension beside the reference path): the
lookup's key semantics on a hand-made cache, factor 0 reproducing the
reference step bit for bit, and the property warm start exists for: on a
resting stack, carried-over impulses leave a smaller normal-velocity
residual after a short sweep than cold starts do."""
import ctypes as C

import numpy as np
import pytest

from harness import World, make_scene, oracle, ptr


def test_lookup_keys_and_factor():
    from seevr_amd import Scene, default_caps
    sc = Scene(default_caps(max_bodies=8, max_hull_verts=100, max_vertices=100, max_triangles=100))
    m = sc.add_material(1000.0)
    for i in range(4):
        sc.add_box((2.0 * i, 0, 0), (0.5, 0.5, 0.5), material=m)
    w = World.from_scene(sc, max_pairs=16, max_contacts=6)
    # contacts (0,1) (0,2) (1,3) (2,3); cache holds (0,2) (1,2) (2,3)
    ga = np.array([0, 0, 1, 2], np.uint32)
    gb = np.array([1, 2, 3, 3], np.uint32)
    w.a["CONTACT_GEOM_A"][:4] = ga
    w.a["CONTACT_GEOM_B"][:4] = gb
    w.a["ATOMIC_COUNTERS"][1] = 4
    w.a["CONTACT_LAMBDA_N"][:] = 7.0
    keys = np.array([(0 << 32) | 2, (1 << 32) | 2, (2 << 32) | 3], np.uint64)
    lam = np.array([0.5, 0.25, 2.0], np.float32)
    b = w.bufs()
    oracle().warm_lookup(C.byref(b), ptr(keys), ptr(lam), 3, C.c_float(0.5))
    assert w.a["CONTACT_LAMBDA_N"][:6].tolist() == [0.0, 0.25, 0.0, 1.0, 0.0, 0.0]
    out_k, out_l = np.zeros(6, np.uint64), np.zeros(6, np.float32)
    assert oracle().warm_save(C.byref(b), ptr(out_k), ptr(out_l)) == 4
    assert out_k[:4].tolist() == [(int(a) << 32) | int(c) for a, c in zip(ga, gb)]
    assert out_l[:4].tolist() == [0.0, 0.25, 0.0, 1.0]


def test_factor_zero_is_the_reference_step():
    sc = make_scene(1, 60, seed=3)
    w = World.from_scene(sc, max_pairs=4000, max_contacts=2000)
    for _ in range(40):
        w.oracle_step()
    cold, warm = w.copy(), w.copy()
    for _ in range(5):
        cold.oracle_step()
        warm.oracle_step_warm(0.0)
    for k in ("BODY_POS", "BODY_QUAT", "BODY_VEL", "BODY_OMEGA", "CONTACT_LAMBDA_N"):
        assert np.array_equal(cold.a[k].view(np.uint32), warm.a[k].view(np.uint32)), k


def chain_world(boxes=6):
    """A static ground under a column of boxes at rest, face on face, with a
    hand-made contact list: one contact at the centre of each touching face,
    normal +y, no penetration, no restitution and no friction.  The normal
    impulses of a sweep from the state after one gravity step then converge
    to the weight above each contact times dt, with no torques."""
    from seevr_amd import Scene, default_caps
    sc = Scene(default_caps(max_bodies=16, max_hull_verts=200, max_vertices=200, max_triangles=400))
    g = sc.add_material(0.0)
    d = sc.add_material(500.0)
    sc.add_box((0, -0.5, 0), (10, 0.5, 10), material=g, friction=0.0, restitution=0.0)
    for k in range(boxes):
        sc.add_box((0.0, 0.2 + 0.4 * k, 0.0), (0.6, 0.2, 0.6), material=d, friction=0.0, restitution=0.0)
    w = World.from_scene(sc, max_pairs=64, max_contacts=64, pgs_iterations=30)
    a = w.a
    a["CONTACT_GEOM_A"][:boxes] = np.arange(boxes)
    a["CONTACT_GEOM_B"][:boxes] = np.arange(1, boxes + 1)
    a["CONTACT_POS"][: 3 * boxes] = np.stack([np.zeros(boxes), 0.4 * np.arange(boxes), np.zeros(boxes)], 1).ravel()
    a["CONTACT_NORMAL"][: 3 * boxes] = np.tile([0.0, 1.0, 0.0], boxes)
    a["CONTACT_PENETRATION"][:boxes] = 0.0
    a["ATOMIC_COUNTERS"][1] = boxes
    return w


def chain_sweep(w, iters, factor, cache, dt):
    x = w.copy()
    b = x.bufs()
    b.pgs_iterations = iters
    keys, lam = cache
    if factor:
        x.lib().warm_lookup(C.byref(b), ptr(keys), ptr(lam), len(keys), C.c_float(factor))
        x.lib().step_solve_warm(C.byref(b), C.c_float(dt))
    else:
        x.lib().step_clear(C.byref(b))
        x.lib().step_solve(C.byref(b), C.c_float(dt))
    nc = x.ncontacts()
    vy = x.a["BODY_VEL"][1::3].astype(np.float64)
    vn = vy[1: nc + 1] - vy[:nc]  # normals are +y and the contact points lie on the axis
    return vn, x.a["CONTACT_LAMBDA_N"][:nc].copy()


@pytest.mark.parametrize("iters", [1, 2, 4])
def test_warm_start_starts_at_the_fixed_point(iters):
    """From the converged impulses of the same substep (the cache a resting
    column carries over), a warm-started sweep of `iters` iterations leaves
    the column at rest, while a cold one still lets it sink; half the
    impulses (factor 0.5) land in between.  A warm start applied at each
    contact's own first update instead of before the sweep would equal the
    cold sweep and fail this."""
    dt = float(np.float32(1 / 60) / np.float32(10))
    w = chain_world()
    w.oracle_stage(0, dt=dt, g=(0.0, -9.81, 0.0))
    _, lam_star = chain_sweep(w, 400, 0.0, (np.zeros(0, np.uint64), np.zeros(0, np.float32)), dt)
    m = 500.0 * 1.2 * 0.4 * 1.2
    assert np.allclose(lam_star, m * 9.81 * dt * np.arange(6, 0, -1), rtol=1e-4)
    # cache in pair order, with an unrelated pair between: lookup by key
    keys = np.array([(0 << 32) | 1, (0 << 32) | 5] + [((k << 32) | (k + 1)) for k in range(1, 6)], np.uint64)
    lam = np.concatenate([lam_star[:1], [123.0], lam_star[1:]]).astype(np.float32)
    speed = {}
    for f in (0.0, 0.5, 1.0):
        vn, _ = chain_sweep(w, iters, f, (keys, lam), dt)
        speed[f] = float(np.sqrt(np.mean(np.minimum(vn, 0.0) ** 2)))
    print(f"iterations {iters}: approach speed {speed}")
    assert speed[1.0] < 1e-3 * speed[0.0]
    assert speed[1.0] < speed[0.5] < speed[0.0]
