"""The 1,048,576-body workload at its own size: 1,048,577 bodies (70%
8-segment spheres, 30% boxes, and the ground), 62M hull vertices, so the
hull offsets pass 2^24 and are read exactly from SV_GEOM_HULL_RANGE
(sv_params.hull_range_u32; seeVR's float offsets, data.zig:468-473, are
exact only below 2^24).  After 110 device frames (a forming pile), one
substep runs stage by stage on the device and on the oracle (with the same
exact hull ranges): gravity, body boxes, narrowphase, the ordered sweep and
integration bit for bit.  The oracle's O(N^2) broadphase is out of reach at
this size (5.5e11 tests), so the broadphase is checked three ways: the LBVH
list equals the all-pairs kernel's list, sampled rows equal a brute-force
float64 scan of every other body, and the oracle's narrowphase and sweep
then run on that list."""
import numpy as np
import pytest

from harness import World, assert_bitwise, device_contacts, device_pairs

pytestmark = pytest.mark.gpu

SUB_DT = float(np.float32(np.float32(1 / 60) / np.float32(10)))
G = (0.0, -9.81, 0.0)


def test_million_bodies_u32_hull_ranges_substep():
    from seevr_amd import Context, Scene, caps_for, default_caps, default_params
    n = 1048576 + 1
    sc = Scene(default_caps(max_bodies=n, max_hull_verts=n * 82, max_vertices=n * 82, max_triangles=n * 130,
                            max_support_entries=n * 193 + 1))
    sc.build_workload(5, 0, 8, 1)
    c = sc.counts()
    nb, nv, nh = c["num_bodies"], c["num_vertices"], c["num_hull_verts"]
    assert nb == n and nh > (1 << 24), "the workload must need exact hull offsets"
    pairs, contacts = 8 * n + 1024, 6 * n + 1024
    ctx = Context(caps_for(sc, pairs, contacts), default_params(pgs_iterations=30, hull_range_u32=1))
    sc.upload(ctx)
    for f in range(110):
        ctx.step(nb, nv, nh, 10, 1 / 60, G)
    print(f"110 frames: counters {ctx.read_counters()}", flush=True)

    w = World.from_scene(sc, max_pairs=pairs, max_contacts=contacts, pgs_iterations=30)
    sc.close()
    w.use_hull_range = True
    for k, v in w.download_state(ctx).items():
        w.a[k][: v.size] = v
    # gravity and body boxes
    ctx.dispatch(8, 1)
    ctx.dispatch(0, nb, SUB_DT, G)
    ctx.dispatch(9, nb)
    w.oracle_stage(8)
    w.oracle_stage(0, dt=SUB_DT, g=G, mt=True)
    w.oracle_stage(1, mt=True)
    st = w.download_state(ctx)
    assert_bitwise(st["BODY_VEL"], w.a["BODY_VEL"], "vel after gravity")
    assert_bitwise(st["BODY_OMEGA"], w.a["BODY_OMEGA"], "omega after gravity")
    amin, amax = ctx.download("BODY_AABB_MIN", 3 * nb), ctx.download("BODY_AABB_MAX", 3 * nb)
    assert_bitwise(amin, w.a["BODY_AABB_MIN"][: 3 * nb], "body box min")
    assert_bitwise(amax, w.a["BODY_AABB_MAX"][: 3 * nb], "body box max")

    # broadphase: LBVH (the automatic choice at this size) == all pairs
    ctx.dispatch(2, nb)
    cnt = ctx.download("ATOMIC_COUNTERS", 2)
    lbvh = device_pairs(ctx, pairs)
    ctx.set_broadphase("all_pairs")
    ctx.dispatch(8, 1)
    ctx.dispatch(2, nb)
    assert_bitwise(ctx.download("ATOMIC_COUNTERS", 2), cnt, "pair counters, all pairs vs LBVH")
    assert_bitwise(device_pairs(ctx, pairs), lbvh, "pair list, all pairs vs LBVH")
    assert len(lbvh) > 500000
    # sampled rows against a brute-force scan (the overlap test of
    # physics.comp:459-466: touching counts, static-static pairs skipped)
    lo, hi = amin.reshape(-1, 3).astype(np.float64), amax.reshape(-1, 3).astype(np.float64)
    static = w.a["BODY_INV_MASS"][:nb] <= 0
    rng = np.random.default_rng(3)
    rows = np.concatenate([[0], rng.choice(nb - 1, 300, replace=False)])
    starts = np.searchsorted(lbvh[:, 0], rows)
    ends = np.searchsorted(lbvh[:, 0], rows, side="right")
    for i, s0, s1 in zip(rows, starts, ends):
        j = np.arange(i + 1, nb)
        ov = ~((hi[i] < lo[j]).any(1) | (hi[j] < lo[i]).any(1))
        ov &= ~(static[i] & static[j])
        assert np.array_equal(lbvh[s0:s1, 1], j[ov]), f"row {i}"
    assert int(cnt[0]) <= pairs, "no truncation at this capacity"
    w.a["COLLISION_PAIRS"][: 2 * len(lbvh)] = lbvh.ravel()
    w.a["ATOMIC_COUNTERS"][0] = cnt[0]

    # narrowphase, sweep and integration from the same pair list
    ctx.dispatch(3, 0, num_bodies=nb)
    w.oracle_stage(3, mt=True)
    dc, oc = device_contacts(ctx, contacts), w.contacts()
    assert len(oc["pen"]) > 100000
    for k in ("ga", "gb", "pos", "normal", "pen"):
        assert_bitwise(dc[k], oc[k], f"contact {k}")
    ctx.dispatch(4, 0)
    ctx.dispatch(5, nb, SUB_DT)
    w.oracle_stage(4)
    w.oracle_stage(5, dt=SUB_DT)
    st = w.download_state(ctx)
    assert_bitwise(st["BODY_VEL"], w.a["BODY_VEL"], "vel after the sweep")
    assert_bitwise(st["BODY_OMEGA"], w.a["BODY_OMEGA"], "omega after the sweep")
    nc = w.ncontacts()
    assert_bitwise(ctx.download("CONTACT_LAMBDA_N", nc), w.a["CONTACT_LAMBDA_N"][:nc], "impulses")
    ctx.dispatch(6, nb, SUB_DT)
    w.oracle_stage(6, dt=SUB_DT, mt=True)
    st = w.download_state(ctx)
    assert_bitwise(st["BODY_POS"], w.a["BODY_POS"], "pos after integrate")
    assert_bitwise(st["BODY_QUAT"], w.a["BODY_QUAT"], "quat after integrate")
    print(f"1M-body substep bit-exact: {len(lbvh)} pairs, {nc} contacts", flush=True)
    ctx.close()
