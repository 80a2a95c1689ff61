"""BVH layout interop with the reference path tracer and the render-vertex
hand-off: after physics frames (step 7 writes VERTICES in place), the
triangle LBVH is built on the device from that same buffer, and the
restated closest-hit query of raytrace.comp:86-133, run on the CPU over the
downloaded node arrays, returns for every ray the brute-force closest hit and
exactly what it returns over the oracle's BVH of the same vertices."""
import numpy as np
import pytest

from harness import make_scene
from test_oracle_trace import oracle_bvh
from trace_oracle import check_hits, rays_at, trace

pytestmark = pytest.mark.gpu


def device_nodes(ctx, n):
    m = 2 * n - 1
    return dict(node_min=ctx.download("BVH_AABB_MIN", 3 * m), node_max=ctx.download("BVH_AABB_MAX", 3 * m),
                node_left=ctx.download("BVH_LEFT", m), node_right=ctx.download("BVH_RIGHT", m),
                node_count=ctx.download("BVH_COUNT", m), prim_idx=ctx.download("BVH_PRIM_INDICES", n))


@pytest.mark.parametrize("bodies,frames", [(40, 0), (1000, 30)])
def test_tracer_hits_after_physics_frames(bodies, frames):
    from seevr_amd import Context, caps_for, default_params
    sc = make_scene(1, bodies, seed=9)
    cn = sc.counts()
    ctx = Context(caps_for(sc, 200000, 100000), default_params(pgs_iterations=30))
    sc.upload(ctx)
    for _ in range(frames):
        ctx.step(cn["num_bodies"], cn["num_vertices"], cn["num_hull_verts"])
    n = cn["num_triangles"]
    ctx.bvh_build(n)
    verts = ctx.download("VERTICES", 3 * cn["num_vertices"])
    idx = ctx.download("INDICES", 3 * n)
    if frames:
        assert not np.array_equal(verts, sc.array("VERTICES")[: verts.size]), "vertices did not move"
    got = device_nodes(ctx, n)
    ctx.close()
    ro, rd = rays_at(verts, idx, 2000, seed=bodies)
    hits, dropped = check_hits(verts, idx, got, ro, rd, n)
    print(f"{bodies} bodies, {n} triangles: {hits} of 2000 rays hit, {dropped} dropped pushes")
    assert hits > 1200
    want = oracle_bvh(verts, idx, n)
    a, b = trace(verts, idx, got, ro, rd), trace(verts, idx, want, ro, rd)
    for x, y in zip(a[:3], b[:3]):
        assert np.array_equal(np.asarray(x).view(np.uint8), np.asarray(y).view(np.uint8))
    assert a[3] == b[3]
