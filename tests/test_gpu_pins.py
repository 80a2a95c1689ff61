"""The hand-derived Gauss-Seidel known answers of test_oracle_pins run on the
device, through the C ABI, with every sweep schedule: the stacked-box chains
whose reduced scalar equations give exact float32 values (physics.comp:
523-582 visited in contact index order)."""
import numpy as np
import pytest

from harness import assert_bitwise
from test_gpu_parity import SOLVER_FORMS, new_ctx
from test_oracle_pins import SUB_DT, scalar_gs, stack_world

pytestmark = pytest.mark.gpu

CASES = [
    # (velocities, contacts, iterations, expected velocities incl. ground, expected impulses)
    ([-1.0, -2.0, -3.0], [(0, 1), (1, 2), (2, 3)], 1, [0.0, -1.0, -2.0, -2.0], [1.0, 1.0, 1.0]),
    ([-1.0, -2.0, -3.0], [(0, 1), (1, 2), (2, 3)], 2, [0.0, -1.0, -1.5, -1.5], [2.0, 2.0, 1.5]),
    ([-1.0, -2.0, -3.0], [(2, 3), (1, 2), (0, 1)], 1, [0.0, 0.0, -1.75, -2.5], [0.5, 0.75, 1.75]),
    ([-1.0, 3.0, 0.0], [(0, 1), (1, 2), (2, 3)], 1, [0.0, 0.0, 1.5, 1.5], [1.0, 0.0, 1.5]),
]


def run_device(w, schedule, wgs, variant):
    ctx = new_ctx(w._scene, w.max_pairs, w.max_contacts, iters=w.pgs_iterations)
    try:
        ctx.set_solver(schedule, wgs)
        ctx.set_sweep_variant(variant)
        w.upload(ctx)
        ctx.dispatch(4, 0)
        ctx.dispatch(5, w.n, SUB_DT)
        nc = w.ncontacts()
        return (ctx.download("BODY_VEL", 3 * w.n).reshape(-1, 3), ctx.download("CONTACT_LAMBDA_N", nc),
                ctx.download("BODY_OMEGA", 3 * w.n))
    finally:
        ctx.close()


@pytest.mark.parametrize("schedule,wgs,variant", SOLVER_FORMS)
def test_sweep_order_known_answers_on_device(schedule, wgs, variant):
    for vels, contacts, iters, want_v, want_l in CASES:
        w = stack_world(vels, contacts, iters)
        v, lam, om = run_device(w, schedule, wgs, variant)
        assert v[:, 1].tolist() == want_v, (vels, contacts, iters)
        assert lam.tolist() == want_l
        assert not om.any() and not v[:, [0, 2]].any()


@pytest.mark.parametrize("seed", range(3))
def test_sweep_exact_scalar_chains_on_device(seed):
    rng = np.random.default_rng(100 + seed)
    nb = 7
    contacts = [(i, i + 1) for i in range(nb)]
    rng.shuffle(contacts)
    contacts = [tuple(int(x) for x in c) for c in contacts]
    vels = (rng.integers(-16, 17, nb) / 4.0).tolist()
    w = stack_world(vels, contacts, 4)
    v_ex, l_ex = scalar_gs(vels, contacts, 4)
    for schedule, wgs, variant in SOLVER_FORMS[:4]:
        v, lam, _ = run_device(w, schedule, wgs, variant)
        assert_bitwise(v[:, 1], np.array([float(x) for x in v_ex], np.float32), f"{schedule}/{wgs} velocities")
        assert_bitwise(lam, np.array([float(x) for x in l_ex], np.float32), f"{schedule}/{wgs} impulses")
