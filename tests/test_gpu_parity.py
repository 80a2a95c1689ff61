"""Parity of every physics stage of libseevr_hip.so with the CPU oracle,
from identical inputs, through the C ABI.

Bar: bit-exact for integer outputs (pair lists, contact body ids, counters)
and, for float state, bit-exact as well: both sides evaluate the same
operation sequence without contraction, with correctly rounded division and
square root and the same atan2f.  The north-star tolerance (1e-5 relative
per substep) is asserted in addition, so a failure says which bar broke.
"""
import numpy as np
import pytest

from harness import World, assert_bitwise, device_contacts, device_pairs, make_scene, oracle, rel_close

pytestmark = pytest.mark.gpu

RTOL = 1e-5  # single-substep float tolerance


def new_ctx(scene, pairs, contacts, iters=30):
    from seevr_amd import Context, caps_for, default_params
    ctx = Context(caps_for(scene, pairs, contacts), default_params(pgs_iterations=iters))
    scene.upload(ctx)
    return ctx


@pytest.fixture(scope="module")
def settled():
    """A 300-body column run 120 frames on the oracle: a live pile."""
    sc = make_scene(1, 300, seed=7)
    w = World.from_scene(sc, max_pairs=40000, max_contacts=20000, pgs_iterations=30)
    for _ in range(120):
        w.oracle_step()
    return sc, w


def test_atan2_bitwise():
    from seevr_amd import Context, default_caps
    rng = np.random.default_rng(0)
    y = rng.standard_normal(20000).astype(np.float32) * 10 ** rng.uniform(-8, 8, 20000).astype(np.float32)
    x = rng.standard_normal(20000).astype(np.float32) * 10 ** rng.uniform(-8, 8, 20000).astype(np.float32)
    sp = np.array([0.0, -0.0, 1.0, -1.0, np.inf, -np.inf, 1e-38, -1e-38], np.float32)
    y = np.concatenate([y, np.repeat(sp, sp.size)])
    x = np.concatenate([x, np.tile(sp, sp.size)])
    with Context(default_caps(max_bodies=4, max_hull_verts=8, max_vertices=8, max_triangles=4)) as ctx:
        got = ctx.debug_atan2(y, x)
    lib = oracle()
    want = np.array([lib.atan2f(float(a), float(b)) for a, b in zip(y, x)], np.float32)
    assert_bitwise(got, want, "atan2f")


def _stage_world(settled):
    sc, w = settled
    return sc, w.copy()


def test_gravity_aabb_bitwise(settled):
    sc, w = _stage_world(settled)
    ctx = new_ctx(sc, w.max_pairs, w.max_contacts)
    w.upload(ctx)
    dt, g = np.float32(1 / 60) / np.float32(10), (0.0, -9.81, 0.0)
    ctx.dispatch(0, w.n, float(dt), g)
    w.oracle_stage(0, dt=float(dt), g=g)
    st = w.download_state(ctx)
    assert_bitwise(st["BODY_VEL"], w.a["BODY_VEL"], "vel after gravity")
    assert_bitwise(st["BODY_OMEGA"], w.a["BODY_OMEGA"], "omega after gravity")
    ctx.dispatch(9, w.n)
    w.oracle_stage(1)
    assert_bitwise(ctx.download("BODY_AABB_MIN", 3 * w.n), w.a["BODY_AABB_MIN"], "aabb min")
    assert_bitwise(ctx.download("BODY_AABB_MAX", 3 * w.n), w.a["BODY_AABB_MAX"], "aabb max")
    ctx.close()


def test_broadphase_narrowphase_bitwise(settled):
    sc, w = _stage_world(settled)
    w.oracle_stage(1)
    ctx = new_ctx(sc, w.max_pairs, w.max_contacts)
    w.upload(ctx)
    w.oracle_stage(8)
    w.oracle_stage(2)
    ctx.dispatch(8, 1)
    ctx.dispatch(2, w.n)
    got = device_pairs(ctx, w.max_pairs)
    assert int(ctx.download("ATOMIC_COUNTERS", 1)[0]) == int(w.a["ATOMIC_COUNTERS"][0])
    assert len(w.pairs()) > 50, "scene should have overlapping boxes"
    assert_bitwise(got, w.pairs(), "pairs")
    # narrowphase from the same pairs
    w.oracle_stage(3)
    ctx.dispatch(3, 0, num_bodies=w.n)
    assert int(ctx.download("ATOMIC_COUNTERS", 2)[1]) == int(w.a["ATOMIC_COUNTERS"][1])
    dc, oc = device_contacts(ctx, w.max_contacts), w.contacts()
    assert len(oc["pen"]) > 20
    for k in ("ga", "gb", "pos", "normal", "pen"):
        assert_bitwise(dc[k], oc[k], f"contact {k}")
    ctx.close()


# (schedule, workgroups, hand-off variant of the multi-workgroup forms)
SOLVER_FORMS = [("auto", 0, "seq"), ("ordered", 1, "seq"), ("ordered", 8, "seq"), ("ordered", 64, "seq"),
                ("ordered", 8, "coherent"), ("ordered", 8, "fenced"), ("ordered", 8, "pipelined"),
                ("ordered", 64, "pipelined"), ("ordered", 8, "seq_basic"), ("ordered", 64, "seq_basic"),
                ("ordered", 8, "paired"), ("ordered", 64, "paired"), ("flow", 1, "seq"), ("flow", 64, "coherent"), ("flow", 64, "fenced")]


@pytest.mark.parametrize("schedule,wgs,variant", SOLVER_FORMS)
def test_solve_bitwise(settled, schedule, wgs, variant):
    """Every sweep schedule reproduces the serial Gauss-Seidel loop bit for bit."""
    sc, w = _stage_world(settled)
    for st in (1, 8, 2, 3):
        w.oracle_stage(st)
    ctx = new_ctx(sc, w.max_pairs, w.max_contacts)
    ctx.set_solver(schedule, wgs)
    ctx.set_sweep_variant(variant)
    w.upload(ctx)  # includes the oracle's contact list and counters
    dt = float(np.float32(1 / 60) / np.float32(10))
    w.oracle_stage(4)
    w.oracle_stage(5, dt=dt)
    ctx.dispatch(4, 0)
    ctx.dispatch(5, w.n, dt)
    st = w.download_state(ctx)
    nc = w.ncontacts()
    assert nc > 20
    ok, err = rel_close(st["BODY_VEL"], w.a["BODY_VEL"], RTOL, 1e-6)
    assert ok, f"vel after PGS max rel err {err}"
    assert_bitwise(st["BODY_VEL"], w.a["BODY_VEL"], "vel after PGS")
    assert_bitwise(st["BODY_OMEGA"], w.a["BODY_OMEGA"], "omega after PGS")
    assert_bitwise(ctx.download("CONTACT_LAMBDA_N", nc), w.a["CONTACT_LAMBDA_N"][:nc], "lambda_n")
    ctx.close()


def test_integrate_verts_bitwise(settled):
    sc, w = _stage_world(settled)
    ctx = new_ctx(sc, w.max_pairs, w.max_contacts)
    w.upload(ctx)
    dt = float(np.float32(1 / 60) / np.float32(10))
    ctx.dispatch(6, w.n, dt)
    w.oracle_stage(6, dt=dt)
    st = w.download_state(ctx)
    assert_bitwise(st["BODY_POS"], w.a["BODY_POS"], "pos after integrate")
    assert_bitwise(st["BODY_QUAT"], w.a["BODY_QUAT"], "quat after integrate")
    ctx.dispatch(7, w.nv, num_bodies=w.n)
    w.oracle_stage(7, count=w.nv)
    assert_bitwise(ctx.download("VERTICES", 3 * w.nv), w.a["VERTICES"][: 3 * w.nv], "render verts")
    ctx.close()


@pytest.mark.parametrize("frames,schedule,wgs,broadphase,variant",
                         [(1, "auto", 0, "auto", "seq"), (8, "auto", 0, "auto", "seq"), (8, "flow", 0, "auto", "seq"),
                          (8, "auto", 0, "lbvh", "seq"), (8, "ordered", 64, "auto", "seq"),
                          (8, "ordered", 64, "auto", "pipelined"), (8, "ordered", 64, "auto", "seq_basic"),
                          (8, "ordered", 64, "auto", "paired")])
def test_full_step_matches_oracle(settled, frames, schedule, wgs, broadphase, variant):
    sc, w = _stage_world(settled)
    ctx = new_ctx(sc, w.max_pairs, w.max_contacts)
    ctx.set_solver(schedule, wgs)
    ctx.set_broadphase(broadphase)
    ctx.set_sweep_variant(variant)
    w.upload(ctx)
    for f in range(frames):
        ctx.step(w.n, w.nv, w.nh, 10, 1 / 60, (0, -9.81, 0))
        w.oracle_step(10, 1 / 60, (0, -9.81, 0))
        st = w.download_state(ctx)
        for k in ("BODY_POS", "BODY_QUAT", "BODY_VEL", "BODY_OMEGA"):
            assert_bitwise(st[k], w.a[k], f"{k} after frame {f}")
    assert ctx.read_counters() == tuple(int(v) for v in w.a["ATOMIC_COUNTERS"][:2])
    assert_bitwise(ctx.download("VERTICES", 3 * w.nv), w.a["VERTICES"][: 3 * w.nv], "render verts")
    ctx.close()


def test_full_step_with_external_geometry_pointers(settled):
    """Once a caller holds device pointers to the hulls or support lookups
    (and may write them itself), every step derives its lookup vertices
    again; the result is unchanged."""
    sc, w = _stage_world(settled)
    ctx = new_ctx(sc, w.max_pairs, w.max_contacts)
    assert ctx.buffer_ptr("SUPPORT_LOOKUP") and ctx.buffer_ptr("HULL_VERTS")
    w.upload(ctx)
    for f in range(2):
        ctx.step(w.n, w.nv, w.nh, 10, 1 / 60, (0, -9.81, 0))
        w.oracle_step(10, 1 / 60, (0, -9.81, 0))
        st = w.download_state(ctx)
        for k in ("BODY_POS", "BODY_QUAT", "BODY_VEL", "BODY_OMEGA"):
            assert_bitwise(st[k], w.a[k], f"{k} after frame {f}")
    ctx.close()


def test_overflow_truncates_in_canonical_order(settled):
    sc, w = _stage_world(settled)
    w.max_pairs, w.max_contacts = 40, 15
    w.a["COLLISION_PAIRS"][:] = 0
    ctx = new_ctx(sc, 40, 15)
    w.upload(ctx)
    ctx.step(w.n, w.nv, w.nh, 1, 1 / 600, (0, -9.81, 0))
    w.oracle_step(1, 1 / 600, (0, -9.81, 0))
    cnt = ctx.download("ATOMIC_COUNTERS", 3)
    assert int(cnt[0]) == int(w.a["ATOMIC_COUNTERS"][0]) > 40
    assert int(cnt[1]) == int(w.a["ATOMIC_COUNTERS"][1])
    assert int(cnt[2]) & 1, "pair overflow flag"
    assert_bitwise(device_pairs(ctx, 40), w.pairs(), "truncated pairs")
    st = w.download_state(ctx)
    assert_bitwise(st["BODY_VEL"], w.a["BODY_VEL"], "vel with truncated lists")
    ctx.close()


def test_repeat_runs_identical(settled):
    sc, w = _stage_world(settled)
    outs = []
    for _ in range(2):
        ctx = new_ctx(sc, w.max_pairs, w.max_contacts)
        w.upload(ctx)
        for _ in range(3):
            ctx.step(w.n, w.nv, w.nh)
        outs.append(w.download_state(ctx))
        ctx.close()
    for k in outs[0]:
        assert_bitwise(outs[1][k], outs[0][k], f"repeat {k}")


@pytest.mark.parametrize("layout", ["empty_bodies", "gap_before_first", "count_below_last_range"])
def test_render_vertex_owners_edge_cases(settled, layout):
    """Step 7's owner of a vertex is the last body whose body_vert_start is
    <= it (the reference's binary search, physics.comp:611-619); the device
    walks each body's range instead.  Starts with bodies that own no vertex
    (repeated starts), vertices below the first body's start, and a vertex
    count that ends inside the last body's range must give the oracle's
    values bit for bit."""
    sc, w = _stage_world(settled)
    ctx = new_ctx(sc, w.max_pairs, w.max_contacts)
    w.upload(ctx)
    n, nv = w.n, w.nv
    start = w.a["BODY_VERT_START"][:n].copy()
    if layout == "empty_bodies":  # every third body owns nothing (its start repeats the next one's)
        for b in range(1, n - 1, 3):
            start[b] = start[b + 1]
    elif layout == "gap_before_first":
        start[0] = start[1] // 2
    count = nv if layout != "count_below_last_range" else int(start[n - 1]) + 3
    w.a["BODY_VERT_START"][:n] = start
    ctx.upload("BODY_VERT_START", start)
    ctx.dispatch(7, count, num_bodies=n)
    w.oracle_stage(7, count=count)
    assert_bitwise(ctx.download("VERTICES", 3 * count), w.a["VERTICES"][: 3 * count], f"render verts ({layout})")
    ctx.close()
