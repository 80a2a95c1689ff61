"""300 frames of the 1,000-body column on the device equal the oracle's
trajectory bit for bit: SHA-256 of the body state and the raw counters at
every 25th frame, against the committed fixture."""
import pytest

from harness import make_scene
from pile_trajectory import digest, load

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("solver,wgs", [("auto", 0), ("flow", 0), ("ordered", 32)])
def test_pile_trajectory_matches_fixture(solver, wgs):
    from seevr_amd import Context, caps_for, default_params
    fx = load()
    sc = make_scene(1, 1000, seed=1)
    c = sc.counts()
    ctx = Context(caps_for(sc, fx["max_pairs"], fx["max_contacts"]),
                  default_params(pgs_iterations=fx["pgs_iterations"]))
    sc.upload(ctx)
    ctx.set_solver(solver, wgs)
    n = c["num_bodies"]
    cps = {cp["frame"]: cp for cp in fx["checkpoints"]}
    for f in range(1, max(cps) + 1):
        ctx.step(n, c["num_vertices"], c["num_hull_verts"], fx["substeps"], fx["dt"], (0.0, -9.81, 0.0))
        if f in cps:
            st = [ctx.download("BODY_POS", 3 * n), ctx.download("BODY_QUAT", 4 * n),
                  ctx.download("BODY_VEL", 3 * n), ctx.download("BODY_OMEGA", 3 * n)]
            assert list(ctx.read_counters()) == cps[f]["counters"], f"counters at frame {f}"
            assert digest(st) == cps[f]["state_sha256"], f"state at frame {f}"
    ctx.close()
