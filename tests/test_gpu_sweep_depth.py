"""sv_sweep_depth: the longest dependency chain of the last substep's
contact sweep (items in the serial order of physics.comp:500-583, each
after the previous item of each of its dynamic bodies), read from the
device's prepared contacts, against the same chain restated here over the
oracle's contact list of the same substep (the device state equals the
oracle's bit for bit, so the lists are the same).  sv_debug_handoff: the
hand-off probe returns a finite latency."""
import numpy as np
import pytest

from harness import World, make_scene

pytestmark = pytest.mark.gpu


def chain_depth(a, b, dyn, iters):
    level = np.zeros(dyn.size, np.int64)
    depth = items = 0
    for _ in range(iters):
        for x, y in zip(a, b):
            if not (dyn[x] or dyn[y]):
                continue  # both bodies static: not an item
            lx = level[x] if dyn[x] else 0
            ly = level[y] if dyn[y] else 0
            lv = max(lx, ly) + 1
            if dyn[x]:
                level[x] = lv
            if dyn[y]:
                level[y] = lv
            depth = max(depth, lv)
            items += 1
    return depth, items


def test_sweep_depth_matches_the_oracles_contacts():
    from seevr_amd import Context, caps_for, default_params
    sc = make_scene(1, 300, seed=3)
    w = World.from_scene(sc, max_pairs=40000, max_contacts=20000, pgs_iterations=30)
    for _ in range(60):  # a pile forms on the oracle
        w.oracle_step()
    ctx = Context(caps_for(sc, w.max_pairs, w.max_contacts), default_params(pgs_iterations=30))
    sc.upload(ctx)
    w.upload(ctx)
    ctx.step(w.n, w.nv, w.nh, 10, 1 / 60, (0.0, -9.81, 0.0))
    w.oracle_step()
    nc = w.ncontacts()
    assert nc > 20
    gb = w.a["GEOM_BODY_IDX"]
    a = gb[w.a["CONTACT_GEOM_A"][:nc].astype(np.int64)]
    b = gb[w.a["CONTACT_GEOM_B"][:nc].astype(np.int64)]
    dyn = w.a["BODY_INV_MASS"][: w.n] > 0
    want = chain_depth(a, b, dyn, 30)
    got = ctx.sweep_depth()
    assert got == want
    assert want[0] >= 30 and want[1] >= want[0]
    ctx.close()


def test_handoff_probe():
    from seevr_amd import Context, default_caps
    ctx = Context(default_caps(max_bodies=64, max_hull_verts=64, max_vertices=64, max_triangles=64))
    for same in (False, True):
        us, xcc = ctx.handoff_us(2000, same_xcd=same)
        assert 0.01 < us < 50.0, (same, us, xcc)
    ctx.close()
