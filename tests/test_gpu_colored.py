"""Graph-coloured sweep: a different (deterministic) order than the
reference's, so it is pinned by running the oracle's serial sweep over the
contacts sorted the same way (colouring restated below), and judged by the
normal-constraint residual against the ordered sweep from the same state.
"""
import numpy as np
import pytest

from harness import World, assert_bitwise, make_scene

pytestmark = pytest.mark.gpu

ROUNDS = 32


def fmix32(h):
    h = h.astype(np.uint64) & 0xFFFFFFFF
    h ^= h >> 16
    h = (h * 0x85EBCA6B) & 0xFFFFFFFF
    h ^= h >> 13
    h = (h * 0xC2B2AE35) & 0xFFFFFFFF
    h ^= h >> 16
    return h


def color_order(w):
    """Colour classes of the library's colouring (hashed-priority rounds),
    returned as the contact order (colour, index); skipped contacts last."""
    nc = w.ncontacts()
    a = w.a["GEOM_BODY_IDX"][w.a["CONTACT_GEOM_A"][:nc]].astype(np.int64)
    b = w.a["GEOM_BODY_IDX"][w.a["CONTACT_GEOM_B"][:nc]].astype(np.int64)
    im = w.a["BODY_INV_MASS"]
    ii = w.a["BODY_INV_INERTIA"].reshape(-1, 3)
    dyn = ~((im == 0) & np.all(ii == 0, axis=1))
    skip = ((im[a] == 0) & (im[b] == 0)) | (a == b)
    da, db = dyn[a], dyn[b]
    c = np.arange(nc, dtype=np.uint64)
    h = fmix32((c * 0x9E3779B9) & 0xFFFFFFFF)
    color = np.full(nc, -1, np.int64)
    color[skip] = ROUNDS + 1
    claim = np.zeros(w.n, np.uint64)
    for r in range(ROUNDS):
        key = (np.uint64(r) << np.uint64(56)) | ((h >> np.uint64(8)) << np.uint64(24)) | (c & np.uint64(0xFFFFFF))
        un = color < 0
        np.maximum.at(claim, a[un & da], key[un & da])
        np.maximum.at(claim, b[un & db], key[un & db])
        win = un & (~da | (claim[a] == key)) & (~db | (claim[b] == key))
        color[win] = r
    color[color < 0] = ROUNDS
    return np.argsort(color, kind="stable"), color


@pytest.fixture(scope="module")
def settled():
    sc = make_scene(1, 300, seed=7)
    w = World.from_scene(sc, max_pairs=40000, max_contacts=20000, pgs_iterations=30)
    for _ in range(120):
        w.oracle_step()
    for st in (1, 8, 2, 3):
        w.oracle_stage(st)
    return sc, w


def new_ctx(sc, w, schedule):
    from seevr_amd import Context, caps_for, default_params
    ctx = Context(caps_for(sc, w.max_pairs, w.max_contacts), default_params(pgs_iterations=30))
    sc.upload(ctx)
    ctx.set_solver(schedule)
    w.upload(ctx)
    return ctx


def test_colored_sweep_equals_serial_sweep_in_colour_order(settled):
    sc, w0 = settled
    w = w0.copy()
    nc = w.ncontacts()
    order, color = color_order(w)
    assert color[color <= ROUNDS].max() > 1 and nc > 50
    dt = float(np.float32(1 / 60) / np.float32(10))
    ctx = new_ctx(sc, w, "colored")
    ctx.dispatch(4, 0)
    ctx.dispatch(5, w.n, dt)
    # oracle: the same serial sweep over the permuted contact list
    wp = w.copy()
    for key, width in (("CONTACT_GEOM_A", 1), ("CONTACT_GEOM_B", 1), ("CONTACT_POS", 3), ("CONTACT_NORMAL", 3),
                       ("CONTACT_PENETRATION", 1)):
        arr = wp.a[key][: width * nc].reshape(nc, width)
        wp.a[key][: width * nc] = arr[order].ravel()
    wp.oracle_stage(4)
    wp.oracle_stage(5, dt=dt)
    st = w.download_state(ctx)
    assert_bitwise(st["BODY_VEL"], wp.a["BODY_VEL"], "vel after coloured sweep")
    assert_bitwise(st["BODY_OMEGA"], wp.a["BODY_OMEGA"], "omega after coloured sweep")
    lam = ctx.download("CONTACT_LAMBDA_N", nc)
    assert_bitwise(lam[order], wp.a["CONTACT_LAMBDA_N"][:nc], "lambda in colour order")
    ctx.close()


def test_colored_residual_close_to_ordered(settled):
    sc, w0 = settled
    dt = float(np.float32(1 / 60) / np.float32(10))
    res = {}
    for schedule in ("ordered", "colored"):
        w = w0.copy()
        ctx = new_ctx(sc, w, schedule)
        ctx.dispatch(4, 0)
        ctx.dispatch(5, w.n, dt)
        res[schedule] = ctx.residual()
        ctx.close()
    (ro, mo, no), (rc, mc, ncol) = res["ordered"], res["colored"]
    assert no == ncol > 50
    print(f"residual rms ordered {ro:.3e} coloured {rc:.3e}; max {mo:.3e} {mc:.3e}")
    assert rc < 2.0 * ro + 1e-4 and mc < 2.0 * mo + 1e-3


def test_colored_full_steps_deterministic(settled):
    sc, w0 = settled
    outs = []
    for _ in range(2):
        w = w0.copy()
        ctx = new_ctx(sc, w, "colored")
        for _ in range(3):
            ctx.step(w.n, w.nv, w.nh)
        outs.append(w.download_state(ctx))
        ctx.close()
    for k in outs[0]:
        assert_bitwise(outs[1][k], outs[0][k], f"coloured repeat {k}")
