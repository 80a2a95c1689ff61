"""Stage timing limited to one stage: with sv_stage_timing_only only that
stage records events (ten per ten-substep frame), and with the limit
cleared every stage does."""
import pytest

from harness import make_scene

pytestmark = pytest.mark.gpu


def test_timing_only_one_stage():
    from seevr_amd import Context, caps_for, default_params
    sc = make_scene(1, 300, seed=2)
    c = sc.counts()
    ctx = Context(caps_for(sc, 8192, 4096), default_params(pgs_iterations=10))
    sc.upload(ctx)
    n, nv, nh = c["num_bodies"], c["num_vertices"], c["num_hull_verts"]
    for _ in range(30):
        ctx.step(n, nv, nh)
    ctx.enable_stage_timing(True, only="phys/pgs_solve")
    ctx.stage_times(reset=True)
    ctx.step(n, nv, nh)
    got = {k: v for k, v in ctx.stage_times(reset=True).items() if v[1] > 0}
    assert list(got) == ["phys/pgs_solve"] and got["phys/pgs_solve"][1] == 10
    assert got["phys/pgs_solve"][0] > 0.0
    ctx.enable_stage_timing(True)
    ctx.step(n, nv, nh)
    got = {k: v for k, v in ctx.stage_times(reset=True).items() if v[1] > 0}
    for k in ("phys/gravity", "phys/aabb", "phys/broad_phase", "phys/narrow_phase", "phys/pgs_solve",
              "phys/integrate"):
        assert got[k][1] == 10, k
    assert got["phys/render_verts"][1] == 1
    ctx.enable_stage_timing(False)
    ctx.close()
    sc.close()
