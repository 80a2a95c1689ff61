"""Warm start (an extension beside the reference path) on the device against
the oracle's warm-start restatement, bit for bit: full steps with the
ordered and flow sweeps (cache lookup, pre-sweep impulses, sweep, cache
update every substep), the coloured sweep against the serial sweep in
colour order after the same pre-pass, and factor 0 returning the
reference's path exactly."""
import numpy as np
import pytest

from harness import World, assert_bitwise, make_scene
from test_gpu_colored import color_order

pytestmark = pytest.mark.gpu

FACTOR = 0.8


@pytest.fixture(scope="module")
def settled():
    sc = make_scene(1, 300, seed=7)
    w = World.from_scene(sc, max_pairs=40000, max_contacts=20000, pgs_iterations=30)
    for _ in range(120):
        w.oracle_step()
    return sc, w


def new_ctx(sc, w, schedule="auto"):
    from seevr_amd import Context, caps_for, default_params
    ctx = Context(caps_for(sc, w.max_pairs, w.max_contacts), default_params(pgs_iterations=30))
    sc.upload(ctx)
    ctx.set_solver(schedule)
    w.upload(ctx)
    return ctx


@pytest.mark.parametrize("schedule", ["ordered", "flow"])
def test_warm_full_steps_match_oracle(settled, schedule):
    sc, w0 = settled
    w = w0.copy()
    w.warm_reset()
    ctx = new_ctx(sc, w, schedule)
    ctx.set_warm_start(FACTOR)
    for f in range(4):
        ctx.step(w.n, w.nv, w.nh, 10, 1 / 60, (0, -9.81, 0))
        w.oracle_step_warm(FACTOR, 10, 1 / 60, (0, -9.81, 0))
        st = w.download_state(ctx)
        for k in ("BODY_POS", "BODY_QUAT", "BODY_VEL", "BODY_OMEGA"):
            assert_bitwise(st[k], w.a[k], f"{k} after warm frame {f}")
    nc = w.ncontacts()
    assert w.warm_n == nc > 50
    assert_bitwise(ctx.download("CONTACT_LAMBDA_N", nc), w.a["CONTACT_LAMBDA_N"][:nc], "impulses")
    assert ctx.read_counters() == tuple(int(v) for v in w.a["ATOMIC_COUNTERS"][:2])
    ctx.close()


def test_warm_colored_sweep_equals_serial_in_colour_order(settled):
    sc, w0 = settled
    w = w0.copy()
    w.warm_reset()
    ctx = new_ctx(sc, w, "ordered")
    ctx.set_warm_start(FACTOR)
    # two frames in lockstep fill both caches
    for _ in range(2):
        ctx.step(w.n, w.nv, w.nh, 10, 1 / 60, (0, -9.81, 0))
        w.oracle_step_warm(FACTOR, 10, 1 / 60, (0, -9.81, 0))
    ctx.set_solver("colored")
    dt = float(np.float32(1 / 60) / np.float32(10))
    g = (0.0, -9.81, 0.0)
    for st in (8,):
        ctx.dispatch(st, 1)
    ctx.dispatch(0, w.n, dt, g)
    ctx.dispatch(9, w.n)
    ctx.dispatch(2, w.n)
    ctx.dispatch(3, 0, num_bodies=w.n)
    ctx.dispatch(4, 0)
    ctx.dispatch(5, w.n, dt)
    w.oracle_stage(8)
    w.oracle_stage(0, dt=dt, g=g)
    w.oracle_stage(1)
    w.oracle_stage(2)
    w.oracle_stage(3)
    import ctypes as C
    from harness import ptr
    b = w.bufs()
    w.lib().warm_lookup(C.byref(b), ptr(w.warm_keys), ptr(w.warm_lam), w.warm_n, C.c_float(FACTOR))
    nc = w.ncontacts()
    assert (w.a["CONTACT_LAMBDA_N"][:nc] > 0).sum() > 50
    order, color = color_order(w)
    w.lib().step_solve_in_order(C.byref(b), C.c_float(dt), 1, ptr(np.ascontiguousarray(order, np.uint32)))
    st = w.download_state(ctx)
    assert_bitwise(st["BODY_VEL"], w.a["BODY_VEL"], "vel after warm coloured sweep")
    assert_bitwise(st["BODY_OMEGA"], w.a["BODY_OMEGA"], "omega after warm coloured sweep")
    assert_bitwise(ctx.download("CONTACT_LAMBDA_N", nc), w.a["CONTACT_LAMBDA_N"][:nc], "impulses")
    ctx.close()


def test_factor_zero_is_the_reference_path(settled):
    sc, w0 = settled
    outs = []
    for warm in (False, True):
        w = w0.copy()
        ctx = new_ctx(sc, w)
        if warm:
            ctx.set_warm_start(0.5)
            ctx.step(w.n, w.nv, w.nh, 10, 1 / 60, (0, -9.81, 0))
            w = w0.copy()
            w.upload(ctx)
            ctx.set_warm_start(0.0)
        for _ in range(2):
            ctx.step(w.n, w.nv, w.nh, 10, 1 / 60, (0, -9.81, 0))
        outs.append(w.download_state(ctx))
        ctx.close()
    for k in outs[0]:
        assert_bitwise(outs[1][k], outs[0][k], k)


def test_set_warm_start_rejects_bad_factors(settled):
    from seevr_amd import SeevrError
    sc, w = settled
    ctx = new_ctx(sc, w.copy())
    for bad in (-0.1, 1.5, float("nan")):
        with pytest.raises(SeevrError):
            ctx.set_warm_start(bad)
    ctx.close()
