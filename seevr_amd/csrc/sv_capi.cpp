// sv_capi.cpp - the extern "C" boundary of libseevr_hip.so: context and
// buffer ownership (the role of data.zig's Data), Physics.step sequencing
// (physics.zig:80-165), BVH.build pass order (bvh.zig:63-172), stage
// dispatch by the reference's step ids, and HIP-event stage timing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>

#include <vector>

#include "sv_internal.h"

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIP_OK(expr)                                                                               \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess) return fail(SV_E_DEVICE, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

// Byte size of every buffer for the given capacities.
void buffer_sizes(const sv_caps &c, size_t *bytes) {
    size_t N = c.max_bodies, H = c.max_hull_verts, V = c.max_vertices, T = c.max_triangles;
    size_t P = c.max_pairs, C = c.max_contacts, L = c.max_support_entries;
    size_t NP = std::max(T, N), NN = NP ? 2 * NP - 1 : 0;
    size_t f = sizeof(float), u = sizeof(uint32_t);
    size_t s[SV_BUF_COUNT] = {};
    s[SV_BODY_POS] = 3 * N * f;
    s[SV_BODY_QUAT] = 4 * N * f;
    s[SV_BODY_VEL] = 3 * N * f;
    s[SV_BODY_OMEGA] = 3 * N * f;
    s[SV_BODY_INV_MASS] = N * f;
    s[SV_BODY_INV_INERTIA] = 3 * N * f;
    s[SV_BODY_VERT_START] = N * u;
    s[SV_BODY_VERT_COUNT] = N * u;
    s[SV_GEOM_BODY_IDX] = N * u;
    s[SV_GEOM_DATA] = 8 * N * f;
    s[SV_GEOM_FRICTION] = N * f;
    s[SV_GEOM_RESTITUTION] = N * f;
    s[SV_HULL_VERTS] = 3 * H * f;
    s[SV_VERTICES] = 3 * V * f;
    s[SV_ORIGINAL_VERTICES] = 3 * V * f;
    s[SV_CONTACT_POS] = 3 * C * f;
    s[SV_CONTACT_NORMAL] = 3 * C * f;
    s[SV_CONTACT_PENETRATION] = C * f;
    s[SV_CONTACT_GEOM_A] = C * u;
    s[SV_CONTACT_GEOM_B] = C * u;
    s[SV_CONTACT_LAMBDA_N] = C * f;
    s[SV_ATOMIC_COUNTERS] = 4 * u;
    s[SV_BODY_AABB_MIN] = 3 * N * f;
    s[SV_BODY_AABB_MAX] = 3 * N * f;
    s[SV_COLLISION_PAIRS] = 2 * P * u;
    s[SV_AABB_TEMP_MIN] = 0;
    s[SV_AABB_TEMP_MAX] = 0;
    s[SV_AABB_TEMP_BODY] = 0;
    s[SV_SUPPORT_LOOKUP_OFFSET] = N * u;
    s[SV_SUPPORT_LOOKUP] = L * u;
    s[SV_INDICES] = 3 * T * u;
    s[SV_BVH_CENTROIDS] = 3 * NP * f;
    s[SV_BVH_MORTON] = NP * u;
    s[SV_BVH_MORTON_TEMP] = NP * u;
    s[SV_BVH_SORT_INDICES] = NP * u;
    s[SV_BVH_SORT_TEMP] = NP * u;
    s[SV_BVH_SCENE_BOUNDS] = 6 * f;
    s[SV_BVH_AABB_MIN] = 3 * NN * f;
    s[SV_BVH_AABB_MAX] = 3 * NN * f;
    s[SV_BVH_LEFT] = NN * u;
    s[SV_BVH_RIGHT] = NN * u;
    s[SV_BVH_COUNT] = NN * u;
    s[SV_BVH_PARENT] = NN * u;
    s[SV_BVH_PRIM_INDICES] = NP * u;
    s[SV_BVH_FLAGS] = NP * u;
    s[SV_BVH_TEMP] = std::max<size_t>(6 * ((NP + 2047) / 2048 + 1) * f, 256);
    s[SV_GEOM_HULL_RANGE] = 2 * N * u;
    memcpy(bytes, s, sizeof(s));
}

template <typename T>
hipError_t dalloc(T *&p, size_t count) {
    p = nullptr;
    if (count == 0) return hipSuccess;
    return hipMalloc(reinterpret_cast<void **>(&p), count * sizeof(T));
}

void free_ctx(sv_ctx *c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (int i = 0; i < SV_BUF_COUNT; i++)
        if (c->buf[i] && c->owned[i]) (void)hipFree(c->buf[i]);
    sv::Scratch &s = c->s;
    void *ptrs[] = {s.hrange, s.rot, s.bpbox, s.row_count, s.row_off, s.bp_stage, s.bp_ovf, s.np_flag, s.np_idx, s.np_res,
                    s.cprep, s.inc_key[0], s.inc_key[1], s.inc_val[0], s.inc_val[1], s.deg, s.deg_off,
                    s.rank, s.ticket, s.solve_ctl, s.flow_count, s.flow_queue, s.scan_part, s.sort_hist, s.misc};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (void *p : s.lazy)
        if (p) (void)hipFree(p);
    for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int sync_and_finish(sv_ctx *c) {
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return fail(SV_E_DEVICE, "device work failed: %s", hipGetErrorString(e));
    e = hipGetLastError();  // a kernel launch that was rejected (resources, bad configuration)
    if (e == hipSuccess && c->launch_err != hipSuccess) e = c->launch_err;
    c->launch_err = hipSuccess;
    if (e != hipSuccess) return fail(SV_E_DEVICE, "kernel launch failed: %s", hipGetErrorString(e));
    if (c->timing) {
        for (auto &p : c->pending) {
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
                c->stage_ms[p.stage] += ms;
                c->stage_calls[p.stage] += 1;
            }
        }
    }
    c->pending.clear();
    c->ev_next = 0;
    return SV_OK;
}

int check_status(sv_ctx *c) {
    uint32_t status = 0;
    HIP_OK(hipMemcpy(&status, reinterpret_cast<uint32_t *>(c->buf[SV_ATOMIC_COUNTERS]) + 2, sizeof(status),
                     hipMemcpyDeviceToHost));
    if (status & SV_STATUS_SOLVER_TIMEOUT)
        return fail(SV_E_TIMEOUT, "ordered PGS made no progress for 2 s and was aborted");
    if (status & SV_STATUS_SORT_TIMEOUT) {
        (void)sv::sort_take_abort(c);
        return fail(SV_E_TIMEOUT, "radix sort look-back made no progress for 2 s and was aborted");
    }
    if (status & SV_STATUS_SCAN_TIMEOUT)
        return fail(SV_E_TIMEOUT, "scan look-back made no progress for 2 s and was aborted");
    return SV_OK;
}

bool valid_buf(int id) { return id >= 0 && id < SV_BUF_COUNT; }

}  // namespace

namespace sv {

int set_error(int code, const char *msg) { return fail(code, "%s", msg); }

int timing_begin(sv_ctx *c, const char *name) {
    if (!c->timing) return -1;
    if (!c->timing_only.empty() && c->timing_only != name) return -1;
    int stage = -1;
    for (size_t i = 0; i < c->stage_names.size(); i++)
        if (c->stage_names[i] == name) stage = (int)i;
    if (stage < 0) {
        c->stage_names.emplace_back(name);
        c->stage_ms.push_back(0.0);
        c->stage_calls.push_back(0);
        stage = (int)c->stage_names.size() - 1;
    }
    while (c->ev_pool.size() < c->ev_next + 2) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return -1;
        c->ev_pool.push_back(e);
    }
    hipEvent_t a = c->ev_pool[c->ev_next], b = c->ev_pool[c->ev_next + 1];
    c->ev_next += 2;
    (void)hipEventRecord(a, c->stream);
    c->pending.push_back({stage, a, b});
    return (int)c->pending.size() - 1;
}

void timing_end(sv_ctx *c, int token) {
    if (token < 0 || !c->timing) return;
    (void)hipEventRecord(c->pending[token].b, c->stream);
}

}  // namespace sv

extern "C" {

void sv_default_caps(sv_caps *c) {
    if (!c) return;
    c->max_bodies = 256;            // data.zig:13
    c->max_hull_verts = 10000000;   // data.zig:17
    c->max_vertices = 10000000;     // data.zig:11
    c->max_triangles = 10000000;    // data.zig:12
    c->max_pairs = 10000;           // data.zig:16
    c->max_contacts = 4000;         // data.zig:15
    c->max_support_entries = 256u * 192u + 1u;  // data.zig:208 plus the reserved slot 0
}

void sv_default_params(sv_params *p) {
    if (!p) return;
    p->pgs_iterations = 4;      // physics.comp:68
    p->baumgarte = 0.2f;        // physics.comp:66
    p->slop = 0.005f;           // physics.comp:67
    p->linear_damping = 0.999f; // physics.comp:422
    p->angular_damping = 0.998f;// physics.comp:425
    p->hull_range_u32 = 0;
}

const char *sv_last_error(void) { return g_err.c_str(); }

int sv_create(const sv_caps *caps, const sv_params *params, sv_ctx **out) {
    if (!caps || !out) return fail(SV_E_INVALID, "sv_create: null argument");
    *out = nullptr;
    if (caps->max_bodies == 0) return fail(SV_E_INVALID, "sv_create: max_bodies must be > 0");
    if (caps->max_bodies >= 0x7FFFFFFFu || caps->max_pairs >= 0x7FFFFFFFu || caps->max_contacts >= 0x3FFFFFFFu ||
        caps->max_triangles >= 0x7FFFFFFFu)
        return fail(SV_E_INVALID, "sv_create: capacity too large for 32-bit indices");
    sv_ctx *c = new (std::nothrow) sv_ctx();
    if (!c) return fail(SV_E_NOMEM, "sv_create: out of host memory");
    c->caps = *caps;
    if (params) c->params = *params; else sv_default_params(&c->params);
    if (const char *wgs = getenv("SEEVR_SOLVER_WGS")) {
        int v = atoi(wgs);
        c->solver_workgroups = v > 0 ? (uint32_t)v : 0u;
    }
    if (const char *f = getenv("SEEVR_SOLVER")) c->solver_flow = strcmp(f, "flow") == 0;
    if ((uint64_t)c->params.pgs_iterations * caps->max_contacts >= 0xFFFFFFFFull) {
        delete c;
        return fail(SV_E_INVALID, "sv_create: pgs_iterations * max_contacts must stay below 2^32");
    }
    hipError_t e = hipGetDevice(&c->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        free_ctx(c);
        return fail(SV_E_DEVICE, "sv_create: no usable HIP device: %s", hipGetErrorString(e));
    }
    c->solver_lds_bodies = sv::solver_lds_bodies(c);
    if (const char *e = getenv("SEEVR_SOLVER_LDS_BODIES")) c->solver_lds_bodies = (uint32_t)atoi(e);
    buffer_sizes(c->caps, c->bytes);
    for (int i = 0; i < SV_BUF_COUNT && e == hipSuccess; i++) {
        if (c->bytes[i] == 0) continue;
        e = hipMalloc(&c->buf[i], c->bytes[i]);
        if (e == hipSuccess) c->owned[i] = true;
        if (e == hipSuccess) e = hipMemset(c->buf[i], 0, c->bytes[i]);
    }
    const sv_caps &k = c->caps;
    size_t N = k.max_bodies, P = k.max_pairs, C = k.max_contacts;
    sv::Scratch &s = c->s;
    if (e == hipSuccess) e = dalloc(s.hrange, 2 * N);
    if (e == hipSuccess) e = dalloc(s.rot, 9 * N);
    if (e == hipSuccess) e = dalloc(s.bpbox, 2 * N);
    // broadphase cells for every n <= N (bp_nseg is not monotonic in n)
    const size_t CELLS = sv::bp_cells_cap((uint32_t)N);
    if (e == hipSuccess) e = dalloc(s.row_count, CELLS);
    if (e == hipSuccess) e = dalloc(s.row_off, CELLS);
    if (e == hipSuccess) e = dalloc(s.bp_stage, CELLS * sv::BP_STAGE);
    if (e == hipSuccess) e = dalloc(s.bp_ovf, sv::bp_ovf_words_cap((uint32_t)N));
    if (e == hipSuccess) e = dalloc(s.np_flag, P + 1);
    if (e == hipSuccess) e = dalloc(s.np_idx, P + 1);
    if (e == hipSuccess) e = dalloc(s.np_res, 8 * P + 8);
    if (e == hipSuccess) e = dalloc(s.cprep, (size_t)sv::CPREP_WORDS * C + sv::CPREP_WORDS);
    for (int q = 0; q < 2 && e == hipSuccess; q++) {
        e = dalloc(s.inc_key[q], 2 * C + 2);
        if (e == hipSuccess) e = dalloc(s.inc_val[q], 2 * C + 2);
    }
    if (e == hipSuccess) e = dalloc(s.deg, N + 1);
    if (e == hipSuccess) e = dalloc(s.deg_off, N + 1);
    if (e == hipSuccess) e = dalloc(s.rank, 2 * C + 2);
    if (e == hipSuccess) e = dalloc(s.ticket, N + 1);
    if (e == hipSuccess) e = dalloc(s.solve_ctl, sv::SOLVE_CTL_WORDS);
    if (e == hipSuccess) e = dalloc(s.misc, 64);
    if (e == hipSuccess) e = hipMemset(s.misc, 0, 64 * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemset(s.solve_ctl, 0, sv::SOLVE_CTL_WORDS * sizeof(uint32_t));
    if (e != hipSuccess) {
        free_ctx(c);
        return fail(SV_E_NOMEM, "sv_create: device allocation failed: %s", hipGetErrorString(e));
    }
    size_t NP = std::max<size_t>(k.max_triangles, N);
    sv::prims_reserve(c, (uint32_t)std::max<size_t>(CELLS, P), (uint32_t)std::max<size_t>(2 * C, NP));
    if (!s.scan_part || !s.sort_hist) {
        free_ctx(c);
        return fail(SV_E_NOMEM, "sv_create: scratch allocation failed");
    }
    // the zeroing above ran on the null stream; the context's stream does not wait for it
    e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        free_ctx(c);
        return fail(SV_E_DEVICE, "sv_create: %s", hipGetErrorString(e));
    }
    *out = c;
    return SV_OK;
}

int sv_destroy(sv_ctx *c) {
    free_ctx(c);
    return SV_OK;
}

int sv_get_params(sv_ctx *c, sv_params *out) {
    if (!c || !out) return fail(SV_E_INVALID, "sv_get_params: null argument");
    *out = c->params;
    return SV_OK;
}

int sv_set_params(sv_ctx *c, const sv_params *p) {
    if (!c || !p) return fail(SV_E_INVALID, "sv_set_params: null argument");
    if ((uint64_t)p->pgs_iterations * c->caps.max_contacts >= 0xFFFFFFFFull)
        return fail(SV_E_INVALID, "sv_set_params: pgs_iterations * max_contacts must stay below 2^32");
    if (p->hull_range_u32 != c->params.hull_range_u32) c->sup_dirty = true;
    c->params = *p;
    return SV_OK;
}

// buffers the expanded support lookup (Scratch::sup_xyz) is derived from
static bool touches_support(sv_buf id) {
    return id == SV_SUPPORT_LOOKUP || id == SV_SUPPORT_LOOKUP_OFFSET || id == SV_HULL_VERTS || id == SV_GEOM_DATA ||
           id == SV_GEOM_HULL_RANGE;
}

int sv_buffer_ptr(sv_ctx *c, sv_buf id, void **dev, size_t *bytes) {
    if (!c || !valid_buf(id)) return fail(SV_E_INVALID, "sv_buffer_ptr: bad context or buffer id %d", (int)id);
    if (dev) *dev = c->buf[id];
    if (bytes) *bytes = c->bytes[id];
    if (dev && touches_support(id)) c->sup_external = true;
    return SV_OK;
}

int sv_bind_buffer(sv_ctx *c, sv_buf id, void *dev, size_t nbytes) {
    if (!c || !valid_buf(id) || !dev) return fail(SV_E_INVALID, "sv_bind_buffer: bad context, buffer id or pointer");
    c->s.bvh_rng_n = 0;
    size_t need[SV_BUF_COUNT];
    buffer_sizes(c->caps, need);
    if (nbytes < need[id])
        return fail(SV_E_CAPACITY, "sv_bind_buffer: buffer %d needs %zu bytes for the capacities, got %zu", (int)id,
                    need[id], nbytes);
    if (reinterpret_cast<uintptr_t>(dev) % 16u)
        return fail(SV_E_INVALID, "sv_bind_buffer: device address must be 16-byte aligned");
    HIP_OK(hipStreamSynchronize(c->stream));  // no queued work still reads the old buffer
    if (c->buf[id] && c->owned[id]) HIP_OK(hipFree(c->buf[id]));
    c->buf[id] = dev;
    c->bytes[id] = nbytes;
    c->owned[id] = false;
    c->bind_epoch++;
    if (touches_support(id)) c->sup_dirty = c->sup_external = true;  // the caller may write it at any time
    return SV_OK;
}

int sv_upload(sv_ctx *c, sv_buf id, const void *src, size_t off, size_t n) {
    if (!c || !valid_buf(id) || (!src && n)) return fail(SV_E_INVALID, "sv_upload: bad argument");
    c->s.bvh_rng_n = 0;  // host data may replace the tree the leaf ranges describe
    if (off > c->bytes[id] || n > c->bytes[id] - off)
        return fail(SV_E_RANGE, "sv_upload: [%zu, %zu) outside buffer %d of %zu bytes", off, off + n, (int)id,
                    c->bytes[id]);
    if (!n) return SV_OK;
    if (touches_support(id)) c->sup_dirty = true;
    HIP_OK(hipMemcpyAsync(static_cast<char *>(c->buf[id]) + off, src, n, hipMemcpyDefault, c->stream));
    return sync_and_finish(c);
}

int sv_download(sv_ctx *c, sv_buf id, void *dst, size_t off, size_t n) {
    if (!c || !valid_buf(id) || (!dst && n)) return fail(SV_E_INVALID, "sv_download: bad argument");
    if (off > c->bytes[id] || n > c->bytes[id] - off)
        return fail(SV_E_RANGE, "sv_download: [%zu, %zu) outside buffer %d of %zu bytes", off, off + n, (int)id,
                    c->bytes[id]);
    if (!n) return SV_OK;
    HIP_OK(hipMemcpyAsync(dst, static_cast<const char *>(c->buf[id]) + off, n, hipMemcpyDefault, c->stream));
    return sync_and_finish(c);
}

int sv_physics_step(sv_ctx *c, uint32_t nb, uint32_t nv, uint32_t nh, uint32_t substeps, float dt,
                    const float gravity[3]) {
    if (!c || !gravity) return fail(SV_E_INVALID, "sv_physics_step: null argument");
    if (nb == 0 || substeps == 0) return SV_OK;  // physics.zig:81
    if (nb > c->caps.max_bodies || nv > c->caps.max_vertices || nh > c->caps.max_hull_verts)
        return fail(SV_E_CAPACITY, "sv_physics_step: counts (%u bodies, %u verts, %u hull verts) exceed capacity",
                    nb, nv, nh);
    const float sub_dt = dt / (float)substeps;  // physics.zig:90
    uint32_t *counters = reinterpret_cast<uint32_t *>(c->buf[SV_ATOMIC_COUNTERS]);
    sv::launch_hull_ranges(c, nb);
    sv::ensure_support_vertices(c, nb);
    sv::fill_u32(c, counters + 2, 0u, 1);
    // unless those stages are being timed, gravity rides on the AABB launch
    // and the integration on the sweep's write-back (same arithmetic, fewer
    // launches); timed, every stage keeps its own launch and timer
    auto timed = [c](const char *stage) { return c->timing && (c->timing_only.empty() || c->timing_only == stage); };
    const bool fuse_gravity = !timed("phys/gravity") && !timed("phys/aabb"), fuse_integrate = !timed("phys/integrate");
    for (uint32_t s = 0; s < substeps; s++) {
        if (fuse_gravity) {
            sv::launch_gravity_aabb(c, nb, sub_dt, gravity);
        } else {
            { sv::StageTimer t(c, "phys/gravity"); sv::launch_gravity(c, nb, sub_dt, gravity, true); }
            { sv::StageTimer t(c, "phys/aabb"); sv::launch_aabb(c, nb); }
        }
        { sv::StageTimer t(c, "phys/broad_phase"); sv::launch_broadphase(c, nb); }
        // without warm start the impulse clear rides on the contact compaction
        const bool warm = c->warm_factor > 0.0f;
        { sv::StageTimer t(c, "phys/narrow_phase"); sv::launch_narrowphase(c, true, !warm, nb); }
        if (warm) { sv::StageTimer t(c, "phys/clear_impulse"); sv::launch_clear_lambda(c); }
        c->rot_fresh = true;
        bool integrated;
        { sv::StageTimer t(c, "phys/pgs_solve"); integrated = sv::launch_solve(c, nb, sub_dt, fuse_integrate); }
        c->rot_fresh = false;
        if (!integrated) { sv::StageTimer t(c, "phys/integrate"); sv::launch_integrate(c, nb, sub_dt); }
    }
    { sv::StageTimer t(c, "phys/render_verts"); sv::launch_update_verts(c, nv, nb); }
    int rc = sync_and_finish(c);
    if (rc) return rc;
    return check_status(c);
}

int sv_read_counters(sv_ctx *c, uint32_t out[2]) {
    if (!c || !out) return fail(SV_E_INVALID, "sv_read_counters: null argument");
    return sv_download(c, SV_ATOMIC_COUNTERS, out, 0, 2 * sizeof(uint32_t));
}

int sv_physics_dispatch(sv_ctx *c, uint32_t step, uint32_t count, float dt, const float gravity[3],
                        uint32_t num_bodies) {
    if (!c) return fail(SV_E_INVALID, "sv_physics_dispatch: null context");
    const sv_caps &k = c->caps;
    switch (step) {
    case 0:
        if (!gravity) return fail(SV_E_INVALID, "sv_physics_dispatch: gravity needed for step 0");
        if (count > k.max_bodies) return fail(SV_E_CAPACITY, "step 0: %u bodies > capacity", count);
        sv::launch_gravity(c, count, dt, gravity);
        break;
    case 1:
    case 9:
        if (count > k.max_bodies) return fail(SV_E_CAPACITY, "step %u: %u bodies > capacity", step, count);
        sv::launch_hull_ranges(c, count);
        sv::launch_aabb(c, count);
        break;
    case 2:
        if (count > k.max_bodies) return fail(SV_E_CAPACITY, "step 2: %u bodies > capacity", count);
        sv::launch_pack_boxes(c, count);
        sv::launch_broadphase(c, count);
        break;
    case 3:
        if (num_bodies > k.max_bodies) return fail(SV_E_CAPACITY, "step 3: %u bodies > capacity", num_bodies);
        sv::launch_hull_ranges(c, num_bodies);
        sv::ensure_support_vertices(c, num_bodies);
        sv::launch_rot(c, num_bodies);
        sv::launch_narrowphase(c);
        break;
    case 4:
        sv::launch_clear_lambda(c);
        break;
    case 5:
        if (count > k.max_bodies) return fail(SV_E_CAPACITY, "step 5: %u bodies > capacity", count);
        sv::fill_u32(c, reinterpret_cast<uint32_t *>(c->buf[SV_ATOMIC_COUNTERS]) + 2, 0u, 1);
        sv::launch_solve(c, count, dt);
        break;
    case 6:
        if (count > k.max_bodies) return fail(SV_E_CAPACITY, "step 6: %u bodies > capacity", count);
        sv::launch_integrate(c, count, dt);
        break;
    case 7:
        if (count > k.max_vertices || num_bodies > k.max_bodies)
            return fail(SV_E_CAPACITY, "step 7: counts exceed capacity");
        sv::launch_update_verts(c, count, num_bodies);
        break;
    case 8:
        sv::launch_clear_counters(c);
        break;
    case 10:
    case 11:
        break;
    default:
        return fail(SV_E_INVALID, "sv_physics_dispatch: unknown step %u", step);
    }
    int rc = sync_and_finish(c);
    if (rc) return rc;
    return step == 5 ? check_status(c) : SV_OK;
}

}  // extern "C"

namespace sv {
int bvh_build_run(sv_ctx *c, uint32_t n, bool bodies) {
    static const char *names[8] = {"bvh/centroids", "bvh/scene_bounds", "bvh/morton_codes", "bvh/radix_sort",
                                   "bvh/build_tree", "bvh/leaf_aabbs", "bvh/clear_flags", "bvh/propagate"};
    for (uint32_t st = 0; st < 8; st++) {
        sv::StageTimer t(c, names[st]);
        sv::bvh_stage(c, st, n, bodies);
    }
    int rc = sync_and_finish(c);
    if (rc == SV_OK && sv::sort_take_abort(c))
        rc = fail(SV_E_TIMEOUT, "bvh build: radix sort look-back made no progress for 2 s and was aborted");
    return rc;
}
}  // namespace sv

extern "C" {

int sv_bvh_build(sv_ctx *c, uint32_t n) {
    if (!c) return fail(SV_E_INVALID, "sv_bvh_build: null context");
    if (n == 0) return SV_OK;  // bvh.zig:64
    if (n > c->caps.max_triangles) return fail(SV_E_CAPACITY, "sv_bvh_build: %u triangles > capacity", n);
    return sv::bvh_build_run(c, n, false);
}

int sv_bvh_build_bodies(sv_ctx *c, uint32_t n) {
    if (!c) return fail(SV_E_INVALID, "sv_bvh_build_bodies: null context");
    if (n == 0) return SV_OK;
    if (n > c->caps.max_bodies) return fail(SV_E_CAPACITY, "sv_bvh_build_bodies: %u bodies > capacity", n);
    return sv::bvh_build_run(c, n, true);
}

int sv_bvh_dispatch(sv_ctx *c, uint32_t step, uint32_t n) {
    if (!c) return fail(SV_E_INVALID, "sv_bvh_dispatch: null context");
    if (step > 11) return fail(SV_E_INVALID, "sv_bvh_dispatch: unknown step %u", step);
    if (n > std::max(c->caps.max_triangles, c->caps.max_bodies))
        return fail(SV_E_CAPACITY, "sv_bvh_dispatch: %u primitives > capacity", n);
    sv::bvh_stage(c, step, n, false);
    return sync_and_finish(c);
}

int sv_device_info(sv_ctx *c, char *name, size_t name_len, char *arch, size_t arch_len) {
    if (!c) return fail(SV_E_INVALID, "sv_device_info: null context");
    hipDeviceProp_t p;
    HIP_OK(hipGetDeviceProperties(&p, c->device));
    if (name && name_len) snprintf(name, name_len, "%s", p.name);
    if (arch && arch_len) snprintf(arch, arch_len, "%s", p.gcnArchName);
    return SV_OK;
}

int sv_enable_stage_timing(sv_ctx *c, int enable) {
    if (!c) return fail(SV_E_INVALID, "sv_enable_stage_timing: null context");
    c->timing = enable != 0;
    return SV_OK;
}

int sv_stage_timing_only(sv_ctx *c, const char *stage) {
    if (!c) return fail(SV_E_INVALID, "sv_stage_timing_only: null context");
    c->timing_only = stage ? stage : "";
    return SV_OK;
}

int sv_set_solver(sv_ctx *c, int schedule, uint32_t workgroups) {
    if (!c) return fail(SV_E_INVALID, "sv_set_solver: null context");
    if (schedule < SV_SOLVER_AUTO || schedule > SV_SOLVER_COLORED)
        return fail(SV_E_INVALID, "sv_set_solver: unknown schedule %d", schedule);
    if (workgroups > 4096) return fail(SV_E_INVALID, "sv_set_solver: at most 4096 workgroups");
    if (schedule == SV_SOLVER_COLORED) {
        int coop = 0;
        if (hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, c->device) != hipSuccess || !coop)
            return fail(SV_E_INVALID, "sv_set_solver: the device has no cooperative launch");
        if (c->caps.max_contacts >= (1u << 24))
            return fail(SV_E_CAPACITY, "sv_set_solver: the coloured sweep takes fewer than 2^24 contacts");
    }
    c->solver_flow = schedule == SV_SOLVER_FLOW;
    c->solver_colored = schedule == SV_SOLVER_COLORED;
    c->solver_workgroups = schedule == SV_SOLVER_AUTO
                               ? 0u
                               : (workgroups ? workgroups : (c->solver_flow || c->solver_colored ? 0u : 1u));
    return SV_OK;
}

int sv_solver_residual(sv_ctx *c, double *rms, double *max_abs, uint32_t *count) {
    if (!c) return fail(SV_E_INVALID, "sv_solver_residual: null context");
    int rc = sv::solver_residual(c, rms, max_abs, count);
    return rc ? fail(rc, "sv_solver_residual: device error") : SV_OK;
}

int sv_debug_color_counts(sv_ctx *c, uint32_t *out, uint32_t cap) {
    if (!c || !out) return fail(SV_E_INVALID, "sv_debug_color_counts: null argument");
    int n = sv::color_counts(c, out, cap);
    return n < 0 ? fail(n, "sv_debug_color_counts: no coloured sweep has run") : n;
}

int sv_set_warm_start(sv_ctx *c, float factor) {
    if (!c) return fail(SV_E_INVALID, "sv_set_warm_start: null context");
    if (!(factor >= 0.0f && factor <= 1.0f)) return fail(SV_E_INVALID, "sv_set_warm_start: factor must be in [0, 1]");
    if (factor > 0.0f && !sv::warm_reserve(c)) return fail(SV_E_NOMEM, "sv_set_warm_start: cache allocation failed");
    c->warm_factor = factor;
    return SV_OK;
}

int sv_debug_set_sweep_variant(sv_ctx *c, int variant) {
    if (!c) return fail(SV_E_INVALID, "sv_debug_set_sweep_variant: null context");
    if (variant < sv::SWEEP_SEQ || variant > sv::SWEEP_PAIR)
        return fail(SV_E_INVALID, "sv_debug_set_sweep_variant: unknown variant %d", variant);
    c->sweep_variant = variant;
    return SV_OK;
}

int sv_set_broadphase(sv_ctx *c, int method) {
    if (!c) return fail(SV_E_INVALID, "sv_set_broadphase: null context");
    if (method < SV_BROADPHASE_AUTO || method > SV_BROADPHASE_SWEEP)
        return fail(SV_E_INVALID, "sv_set_broadphase: unknown method %d", method);
    c->broadphase_mode = (uint32_t)method;
    return SV_OK;
}

int sv_broadphase_method(sv_ctx *c, uint32_t n) {
    if (!c) return fail(SV_E_INVALID, "sv_broadphase_method: null context");
    return sv::broadphase_method(c, n);
}

int sv_stage_times(sv_ctx *c, int reset, const char **names, double *avg_us, uint64_t *calls, int cap) {
    if (!c) return fail(SV_E_INVALID, "sv_stage_times: null context");
    int n = (int)c->stage_names.size();
    for (int i = 0; i < n && i < cap; i++) {
        if (names) names[i] = c->stage_names[i].c_str();
        if (avg_us) avg_us[i] = c->stage_calls[i] ? 1000.0 * c->stage_ms[i] / (double)c->stage_calls[i] : 0.0;
        if (calls) calls[i] = c->stage_calls[i];
    }
    if (reset) {
        std::fill(c->stage_ms.begin(), c->stage_ms.end(), 0.0);
        std::fill(c->stage_calls.begin(), c->stage_calls.end(), 0);
    }
    return n;
}

// Host-only: the all-pairs broadphase's need at n bodies against what a
// context of capacity max_bodies allocates.
uint32_t sv_debug_seq_streams(uint32_t wgs) { return sv::seq_streams(wgs); }

int sv_sweep_depth(sv_ctx *c, uint32_t *depth, uint64_t *items) {
    if (!c || !depth || !items) return fail(SV_E_INVALID, "sv_sweep_depth: bad argument");
    return sv::sweep_depth(c, depth, items);
}

int sv_debug_handoff(sv_ctx *c, uint32_t rounds, int same_xcd, double *us_one_way, uint32_t xcc[2]) {
    if (!c || !us_one_way || !xcc || rounds == 0 || rounds > (1u << 24))
        return fail(SV_E_INVALID, "sv_debug_handoff: bad argument");
    uint32_t *d = nullptr;
    HIP_OK(hipMalloc(&d, 64 * sizeof(uint32_t)));
    uint64_t *out = reinterpret_cast<uint64_t *>(d + 16);
    hipError_t e = sv::zero_device(d, 64 * sizeof(uint32_t));  // before the launch on the context stream
    uint64_t h[5] = {};
    if (e == hipSuccess) {
        sv::launch_handoff(c, d, rounds, same_xcd ? 8u : 1u, out);
        e = hipStreamSynchronize(c->stream);
    }
    if (e == hipSuccess) e = hipMemcpy(h, out, sizeof(h), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(SV_E_DEVICE, "sv_debug_handoff: %s", hipGetErrorString(e));
    if (h[4]) return fail(SV_E_TIMEOUT, "sv_debug_handoff: a wait passed 2 s");
    *us_one_way = (double)(h[1] - h[0]) * 0.01 / (2.0 * rounds - 1.0);
    xcc[0] = (uint32_t)h[2];
    xcc[1] = (uint32_t)h[3];
    return SV_OK;
}

int sv_debug_broadphase_scratch(uint32_t max_bodies, uint32_t n, uint64_t out[4]) {
    if (!out || n > max_bodies) return fail(SV_E_INVALID, "sv_debug_broadphase_scratch: bad argument");
    const bool all_pairs = n <= sv::BP_ALL_PAIRS_MAX_BODIES;
    const uint64_t nseg = sv::bp_nseg(n);
    out[0] = all_pairs ? (uint64_t)n * nseg + 1 : 0;                             // cells (+ scan total)
    out[1] = sv::bp_cells_cap(max_bodies);
    out[2] = all_pairs ? 1 + (uint64_t)sv::div_up(n, sv::BP_ROWS) * nseg : 0;  // overflow list
    out[3] = sv::bp_ovf_words_cap(max_bodies);
    return SV_OK;
}

// Evaluate the library's atan2f on the device (host arrays in and out).
int sv_debug_atan2(sv_ctx *c, const float *y, const float *x, float *out, uint32_t n) {
    if (!c || (n && (!y || !x || !out))) return fail(SV_E_INVALID, "sv_debug_atan2: bad argument");
    if (!n) return SV_OK;
    float *d = nullptr;
    HIP_OK(hipMalloc(&d, (size_t)3 * n * sizeof(float)));
    hipError_t e = hipMemcpy(d, y, n * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d + n, x, n * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        sv::launch_debug_atan2(c, d, d + n, d + 2 * n, n);
        e = hipStreamSynchronize(c->stream);
    }
    if (e == hipSuccess) e = hipMemcpy(out, d + 2 * n, n * sizeof(float), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(SV_E_DEVICE, "sv_debug_atan2: %s", hipGetErrorString(e));
    return SV_OK;
}

int sv_debug_radix_sort(sv_ctx *c, const uint32_t *keys, const uint32_t *vals, uint32_t n, uint32_t n_max,
                        uint32_t key_bits, uint32_t *keys_out, uint32_t *vals_out) {
    if (!c || n > n_max || n_max > sv::SORT_MAX_KEYS || key_bits < 1 || key_bits > 32 ||
        (n && (!keys || !vals || !keys_out || !vals_out)))
        return fail(SV_E_INVALID, "sv_debug_radix_sort: bad argument");
    if (!n_max) return SV_OK;
    uint32_t *d = nullptr;
    HIP_OK(hipMalloc(&d, (4 * (size_t)n_max + 1) * sizeof(uint32_t)));
    uint32_t *k0 = d, *v0 = d + n_max, *k1 = d + 2 * (size_t)n_max, *v1 = d + 3 * (size_t)n_max, *dn = d + 4 * (size_t)n_max;
    hipError_t e = hipMemset(d, 0xFF, 4 * (size_t)n_max * sizeof(uint32_t));  // past n: never read as keys
    if (e == hipSuccess && n) e = hipMemcpy(k0, keys, n * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess && n) e = hipMemcpy(v0, vals, n * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dn, &n, sizeof(n), hipMemcpyHostToDevice);
    int rc = SV_OK;
    if (e == hipSuccess) {
        const int where = sv::radix_sort(c, k0, v0, k1, v1, dn, n_max, (int)key_bits);
        rc = sync_and_finish(c);
        if (rc == SV_OK && sv::sort_take_abort(c))
            rc = fail(SV_E_TIMEOUT, "sv_debug_radix_sort: look-back made no progress for 2 s and was aborted");
        if (rc == SV_OK && n) {
            e = hipMemcpy(keys_out, where ? k1 : k0, n * sizeof(uint32_t), hipMemcpyDeviceToHost);
            if (e == hipSuccess) e = hipMemcpy(vals_out, where ? v1 : v0, n * sizeof(uint32_t), hipMemcpyDeviceToHost);
        }
    }
    (void)hipFree(d);
    if (rc != SV_OK) return rc;
    if (e != hipSuccess) return fail(SV_E_DEVICE, "sv_debug_radix_sort: %s", hipGetErrorString(e));
    return SV_OK;
}

int sv_debug_scan(sv_ctx *c, const uint32_t *in, uint32_t n, uint32_t n_max, uint32_t *out, uint32_t *total) {
    if (!c || !total || n > n_max || n_max > (1u << 30) || (n && (!in || !out)))
        return fail(SV_E_INVALID, "sv_debug_scan: bad argument");
    *total = 0;
    if (!n_max) return SV_OK;
    uint32_t *d = nullptr;
    HIP_OK(hipMalloc(&d, (2 * (size_t)n_max + 2) * sizeof(uint32_t)));
    uint32_t *din = d, *dout = d + n_max, *dn = d + 2 * (size_t)n_max, *dtot = dn + 1;
    hipError_t e = hipMemset(d, 0xFF, 2 * (size_t)n_max * sizeof(uint32_t));  // past n: never read
    if (e == hipSuccess && n) e = hipMemcpy(din, in, n * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dn, &n, sizeof(n), hipMemcpyHostToDevice);
    int rc = SV_OK;
    if (e == hipSuccess) {
        sv::scan_exclusive(c, din, dout, dn, n_max, dtot);
        rc = sync_and_finish(c);
        if (rc == SV_OK) rc = check_status(c);
        if (rc == SV_OK && n) e = hipMemcpy(out, dout, n * sizeof(uint32_t), hipMemcpyDeviceToHost);
        if (rc == SV_OK && e == hipSuccess) e = hipMemcpy(total, dtot, sizeof(uint32_t), hipMemcpyDeviceToHost);
    }
    (void)hipFree(d);
    if (rc != SV_OK) return rc;
    if (e != hipSuccess) return fail(SV_E_DEVICE, "sv_debug_scan: %s", hipGetErrorString(e));
    return SV_OK;
}

int sv_debug_sort_rows(sv_ctx *c, const uint32_t *cnt, const uint32_t *entries, uint32_t n, uint32_t max_pairs,
                       uint32_t window_words, uint32_t *pairs_out) {
    if (!c || n == 0 || !cnt || !pairs_out || (window_words && (window_words % 256 || window_words > (1u << 15))))
        return fail(SV_E_INVALID, "sv_debug_sort_rows: bad argument");
    std::vector<uint32_t> off(n);
    uint64_t total = 0;
    for (uint32_t i = 0; i < n; i++) {
        if (cnt[i] > n - 1 - i) return fail(SV_E_INVALID, "sv_debug_sort_rows: row %u holds more partners than j > i", i);
        off[i] = (uint32_t)total;
        total += cnt[i];
    }
    if (total > (1u << 30) || (total && !entries)) return fail(SV_E_INVALID, "sv_debug_sort_rows: bad entries");
    const size_t words = 3 * (size_t)n + 2 + total + 2 * (size_t)max_pairs + 1;
    uint32_t *d = nullptr;
    HIP_OK(hipMalloc(&d, words * sizeof(uint32_t)));
    uint32_t *dcnt = d, *doff = d + n, *dlist = d + 2 * (size_t)n, *dpk = dlist + n + 2, *dpairs = dpk + total + 1;
    hipError_t e = hipMemset(d, 0, words * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemcpy(dcnt, cnt, n * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(doff, off.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess && total) e = hipMemcpy(dpk, entries, total * sizeof(uint32_t), hipMemcpyHostToDevice);
    int rc = SV_OK;
    if (e == hipSuccess) {
        sv::debug_sort_rows(c, n, dcnt, doff, dpk, dpairs, max_pairs, dlist, window_words);
        rc = sync_and_finish(c);
        const uint64_t kept = total < max_pairs ? total : max_pairs;
        if (rc == SV_OK && kept)
            e = hipMemcpy(pairs_out, dpairs, 2 * kept * sizeof(uint32_t), hipMemcpyDeviceToHost);
    }
    (void)hipFree(d);
    if (rc != SV_OK) return rc;
    if (e != hipSuccess) return fail(SV_E_DEVICE, "sv_debug_sort_rows: %s", hipGetErrorString(e));
    return SV_OK;
}

// Copy `nbytes` device-to-device `iters` times with k_copy_calib and report
// the average bandwidth (read + write bytes per second).
int sv_debug_xcc_ids(sv_ctx *c, uint32_t nblocks, uint32_t *out) {
    if (!c || !out || nblocks == 0 || nblocks > 65536) return fail(SV_E_INVALID, "sv_debug_xcc_ids: bad argument");
    uint32_t *d = nullptr;
    hipError_t e = hipMalloc(&d, nblocks * sizeof(uint32_t));
    if (e == hipSuccess) {
        sv::launch_xcc_ids(c, d, nblocks);
        e = hipMemcpyAsync(out, d, nblocks * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (d) (void)hipFree(d);
    return e == hipSuccess ? SV_OK : fail(SV_E_DEVICE, "sv_debug_xcc_ids: %s", hipGetErrorString(e));
}

int sv_debug_copy_bandwidth(sv_ctx *c, size_t nbytes, uint32_t iters, double *gbps) {
    if (!c || !gbps || nbytes < 16 || iters == 0) return fail(SV_E_INVALID, "sv_debug_copy_bandwidth: bad argument");
    void *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc(&a, nbytes);
    if (e == hipSuccess) e = hipMalloc(&b, nbytes);
    if (e == hipSuccess) e = hipMemsetAsync(a, 0, nbytes, c->stream);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) {
        sv::launch_copy_calib(c, a, b, nbytes);  // warm up
        (void)hipEventRecord(e0, c->stream);
        for (uint32_t i = 0; i < iters; i++) sv::launch_copy_calib(c, a, b, nbytes);
        (void)hipEventRecord(e1, c->stream);
        e = hipStreamSynchronize(c->stream);
    }
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (e != hipSuccess) return fail(SV_E_DEVICE, "sv_debug_copy_bandwidth: %s", hipGetErrorString(e));
    *gbps = 2.0 * (double)(nbytes / 16 * 16) * iters / (ms * 1e-3) / 1e9;
    return SV_OK;
}

}  // extern "C"
