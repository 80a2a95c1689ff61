"""Caller-owned buffers (sv_bind_buffer): the physics and BVH binding tables
of the reference (src/physics.zig:37-68, src/bvh.zig:32-51) bound to device
memory the caller allocated (hipMalloc through ctypes), then full steps and a
body LBVH build checked bit for bit against the oracle, with the results
read back through the caller's memory rather than through the library."""
import numpy as np
import pytest

from harness import HipMem, World, assert_bitwise, make_scene
from test_gpu_bvh import oracle_bvh

pytestmark = pytest.mark.gpu

PHYSICS = ["BODY_POS", "BODY_QUAT", "BODY_VEL", "BODY_OMEGA", "BODY_INV_MASS", "BODY_INV_INERTIA",
           "BODY_VERT_START", "BODY_VERT_COUNT", "GEOM_BODY_IDX", "GEOM_DATA", "GEOM_FRICTION",
           "GEOM_RESTITUTION", "HULL_VERTS", "VERTICES", "ORIGINAL_VERTICES", "CONTACT_POS", "CONTACT_NORMAL",
           "CONTACT_PENETRATION", "CONTACT_GEOM_A", "CONTACT_GEOM_B", "CONTACT_LAMBDA_N", "ATOMIC_COUNTERS",
           "BODY_AABB_MIN", "BODY_AABB_MAX", "COLLISION_PAIRS", "SUPPORT_LOOKUP_OFFSET", "SUPPORT_LOOKUP",
           "GEOM_HULL_RANGE"]
BVH = ["INDICES", "BVH_CENTROIDS", "BVH_MORTON", "BVH_MORTON_TEMP", "BVH_SORT_INDICES", "BVH_SORT_TEMP",
       "BVH_SCENE_BOUNDS", "BVH_AABB_MIN", "BVH_AABB_MAX", "BVH_LEFT", "BVH_RIGHT", "BVH_COUNT", "BVH_PARENT",
       "BVH_PRIM_INDICES", "BVH_FLAGS", "BVH_TEMP"]


def test_bound_buffers_step_and_lbvh_bitwise():
    from seevr_amd import Context, caps_for, default_params
    hm = HipMem()

    def as_np(p, count, dtype):
        return hm.download(p, count, dtype)

    sc = make_scene(1, 300, seed=3)
    w = World.from_scene(sc, max_pairs=40000, max_contacts=20000, pgs_iterations=30)
    for _ in range(60):  # a pile forms on the oracle
        w.oracle_step()
    ctx = Context(caps_for(sc, w.max_pairs, w.max_contacts), default_params(pgs_iterations=30))
    mem = {}
    for name in PHYSICS + BVH:
        nb = ctx.buffer_bytes(name)
        if nb == 0:
            continue
        p = hm.alloc(nb + 32, fill=-1)  # caller memory, not zeroed
        ctx.bind_buffer(name, p, nb + 32)
        assert ctx.buffer_ptr(name) == p
        mem[name] = p
    with pytest.raises(Exception):  # smaller than the capacities need
        ctx.bind_buffer("BODY_POS", hm.alloc(16), 16)
    sc.upload(ctx)
    w.upload(ctx)
    for _ in range(4):
        ctx.step(w.n, w.nv, w.nh, 10, 1 / 60, (0.0, -9.81, 0.0))
        w.oracle_step()
    n, nv = w.n, w.nv
    for k in ("BODY_POS", "BODY_QUAT", "BODY_VEL", "BODY_OMEGA"):
        cnt = (4 if k == "BODY_QUAT" else 3) * n
        assert_bitwise(as_np(mem[k], cnt, np.float32), w.a[k], f"{k} in caller memory")
    assert ctx.read_counters() == tuple(int(v) for v in w.a["ATOMIC_COUNTERS"][:2])
    # render vertices written in place for a renderer that binds the same memory
    assert_bitwise(as_np(mem["VERTICES"], 3 * nv, np.float32), w.a["VERTICES"][: 3 * nv], "render vertices")
    nc = w.ncontacts()
    assert nc > 20
    assert_bitwise(as_np(mem["CONTACT_LAMBDA_N"], nc, np.float32), w.a["CONTACT_LAMBDA_N"][:nc], "impulses")

    # the body LBVH into the bound BVH arrays, from the step's last body boxes
    ctx.bvh_build_bodies(n)
    dev_amin = as_np(mem["BODY_AABB_MIN"], 3 * n, np.float32)
    dev_amax = as_np(mem["BODY_AABB_MAX"], 3 * n, np.float32)
    want = oracle_bvh(None, None, n, bodies=(dev_amin, dev_amax))
    nn = 2 * n - 1
    got = dict(morton=as_np(mem["BVH_MORTON"], n, np.uint32), sort_idx=as_np(mem["BVH_SORT_INDICES"], n, np.uint32),
               node_left=as_np(mem["BVH_LEFT"], nn, np.uint32), node_parent=as_np(mem["BVH_PARENT"], nn, np.uint32),
               node_count=as_np(mem["BVH_COUNT"], nn, np.uint32),
               node_min=as_np(mem["BVH_AABB_MIN"], 3 * nn, np.float32),
               node_max=as_np(mem["BVH_AABB_MAX"], 3 * nn, np.float32),
               prim_idx=as_np(mem["BVH_PRIM_INDICES"], n, np.uint32))
    for k, v in got.items():
        assert_bitwise(v, want[k], f"LBVH {k}")
    assert_bitwise(as_np(mem["BVH_RIGHT"], nn, np.uint32)[: n - 1], want["node_right"][: n - 1], "LBVH right")
    pos = as_np(mem["BODY_POS"], 3 * n, np.float32)
    ctx.close()
    # the caller's memory outlives the context (the library did not free it)
    assert_bitwise(as_np(mem["BODY_POS"], 3 * n, np.float32), pos, "caller memory after sv_destroy")
    hm.close()
