"""The 32,768-body walled-container workload at its own size: 32,773 bodies
(32,768 seeded boxes and 16-segment spheres plus ground and 4 walls,
4.86M hull vertices, 8.57M render triangles) settled 240 frames on the
device, then handed to the oracle.  The next 10 substeps run one at a time
on both: after each, body state, body boxes, the pair list, the contact
list with impulses and the counters must agree bit for bit.  Then the LBVH
over the body boxes and the triangle LBVH over the render mesh (the
reference's BVH.build) are compared with the oracle's, bit for bit."""
import numpy as np
import pytest

from harness import World, assert_bitwise, device_contacts, device_pairs
from test_gpu_bvh import device_bvh, oracle_bvh

pytestmark = pytest.mark.gpu

SUB_DT = float(np.float32(np.float32(1 / 60) / np.float32(10)))
G = (0.0, -9.81, 0.0)


def compare_bvh(got, want, n, what):
    for k in ("morton", "sort_idx", "node_left", "node_count", "node_parent", "prim_idx", "node_min", "node_max",
              "centroids", "bounds"):
        assert_bitwise(got[k], want[k], f"{what} {k}")
    assert_bitwise(got["node_right"][: n - 1], want["node_right"][: n - 1], f"{what} node_right")


def test_container_32k_per_substep_and_lbvh():
    from seevr_amd import Context, Scene, caps_for, default_caps, default_params
    n = 32768 + 5
    sc = Scene(default_caps(max_bodies=n, max_hull_verts=n * 289, max_vertices=n * 289, max_triangles=n * 512,
                            max_support_entries=n * 192 + 1))
    sc.build_workload(2, 32768, 16, 1)
    c = sc.counts()
    nb, nv, nh, nt = c["num_bodies"], c["num_vertices"], c["num_hull_verts"], c["num_triangles"]
    assert nb == n
    pairs, contacts = 8 * n + 1024, 6 * n + 1024
    ctx = Context(caps_for(sc, pairs, contacts), default_params(pgs_iterations=30))
    sc.upload(ctx)
    for f in range(240):
        ctx.step(nb, nv, nh, 10, 1 / 60, G)
    print(f"settled 240 frames: counters {ctx.read_counters()}", flush=True)
    w = World.from_scene(sc, max_pairs=pairs, max_contacts=contacts, pgs_iterations=30)
    for k, v in w.download_state(ctx).items():
        w.a[k][: v.size] = v
    w.a["VERTICES"][: 3 * nv] = ctx.download("VERTICES", 3 * nv)
    for s in range(10):
        ctx.step(nb, nv, nh, 1, SUB_DT, G)
        w.oracle_step(1, SUB_DT, G, mt=True)
        st = w.download_state(ctx)
        for k in ("BODY_POS", "BODY_QUAT", "BODY_VEL", "BODY_OMEGA"):
            assert_bitwise(st[k], w.a[k], f"{k} after substep {s}")
        for k in ("BODY_AABB_MIN", "BODY_AABB_MAX"):
            assert_bitwise(ctx.download(k, 3 * nb), w.a[k][: 3 * nb], f"{k} of substep {s}")
        got_cnt = ctx.read_counters()
        assert got_cnt == tuple(int(v) for v in w.a["ATOMIC_COUNTERS"][:2]), f"counters of substep {s}"
        assert_bitwise(device_pairs(ctx, pairs), w.pairs(), f"pair list of substep {s}")
        dc, oc = device_contacts(ctx, contacts), w.contacts()
        for k in ("ga", "gb", "pos", "normal", "pen", "lam"):
            assert_bitwise(dc[k], oc[k], f"contact {k} of substep {s}")
        print(f"substep {s}: {got_cnt[0]} pairs, {got_cnt[1]} contacts bit-exact", flush=True)
    assert got_cnt[1] > 30000, "the settled container should hold a deep pile"
    assert_bitwise(ctx.download("VERTICES", 3 * nv), w.a["VERTICES"][: 3 * nv], "render vertices")

    # LBVH over the body boxes of the last substep (rebuilt every frame in this workload)
    ctx.bvh_build_bodies(nb)
    amin, amax = ctx.download("BODY_AABB_MIN", 3 * nb), ctx.download("BODY_AABB_MAX", 3 * nb)
    compare_bvh(device_bvh(ctx, nb), oracle_bvh(None, None, nb, bodies=(amin, amax)), nb, "body LBVH")

    # the reference's BVH.build over the render triangles of the current vertices
    ctx.bvh_build(nt)
    verts = ctx.download("VERTICES", 3 * nv)
    idx = w.a["INDICES"][: 3 * nt].copy()
    compare_bvh(device_bvh(ctx, nt), oracle_bvh(verts, idx, nt), nt, "triangle LBVH")
    print(f"body LBVH ({nb} leaves) and triangle LBVH ({nt} leaves) bit-exact", flush=True)
    ctx.close()
