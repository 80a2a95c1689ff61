"""Parity where the automatic choices pick the multi-workgroup paths: a
4,096-body container scene is stepped 240 frames on the device, its state is
handed to the oracle, and two more frames must agree bit for bit (the
sequence-form sweep, chosen from 2,048 bodies, in its paired pipelined
kernel, and the all-pairs broadphase; then the LBVH broadphase, the basic
sequence kernel and the one-lane pipelined kernel)."""
import numpy as np
import pytest

from harness import World, assert_bitwise

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("broadphase,variant", [("auto", "seq"), ("lbvh", "seq"), ("auto", "seq_basic"), ("auto", "pipelined")])
def test_multi_workgroup_paths_match_oracle(broadphase, variant):
    from seevr_amd import Context, Scene, caps_for, default_caps, default_params
    n = 4096 + 5
    sc = Scene(default_caps(max_bodies=n, max_hull_verts=n * 289, max_vertices=n * 289, max_triangles=n * 512,
                            max_support_entries=n * 192 + 1))
    sc.build_workload(2, 4096, 16, 3)
    c = sc.counts()
    pairs, contacts = 8 * n + 1024, 6 * n + 1024
    ctx = Context(caps_for(sc, pairs, contacts), default_params(pgs_iterations=30))
    sc.upload(ctx)
    ctx.set_broadphase(broadphase)
    ctx.set_sweep_variant(variant)
    for _ in range(240):
        ctx.step(c["num_bodies"], c["num_vertices"], c["num_hull_verts"])
    w = World.from_scene(sc, max_pairs=pairs, max_contacts=contacts, pgs_iterations=30)
    for k, v in w.download_state(ctx).items():
        w.a[k][: v.size] = v
    w.a["VERTICES"][: 3 * w.nv] = ctx.download("VERTICES", 3 * w.nv)
    for f in range(2):
        ctx.step(w.n, w.nv, w.nh)
        w.oracle_step()
        st = w.download_state(ctx)
        for k in ("BODY_POS", "BODY_QUAT", "BODY_VEL", "BODY_OMEGA"):
            assert_bitwise(st[k], w.a[k], f"{k} after frame {f}")
    got = ctx.read_counters()
    assert got == tuple(int(v) for v in w.a["ATOMIC_COUNTERS"][:2])
    assert got[1] > 500, "the pile should be in contact"
    ctx.close()
