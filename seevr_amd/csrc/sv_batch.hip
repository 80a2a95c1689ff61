// sv_batch.hip - batched engine: thousands of independent scenes of uneven
// body counts, one persistent launch per substep.
//
// A workgroup takes a scene from a queue (largest scenes first) and runs the
// whole substep of physics.comp on it: gravity, per-body AABB (one wave per
// body), the all-pairs broadphase over the scene's boxes in LDS with
// row counts scanned into lexicographic slots, MPR per pair with contacts
// compacted in pair order, and the ordered Gauss-Seidel sweep with the
// scene's velocities, impulses and per-body tickets in LDS (workgroup-scope
// hand-offs), then integration.  Every float expression goes through the
// device functions the plain path uses (sv_dev.h), and every list keeps the
// plain path's canonical order and truncation, so a scene stepped in a batch
// equals the same scene stepped alone by sv_physics_step, bit for bit.
//
// The per-body (body, contact) incidence order the sweep needs comes from a
// bitonic sort in LDS of keys (local body << 12 | 2 * contact + side).
#include "sv_batch.h"

#include <algorithm>
#include <new>
#include <numeric>
#include <vector>

namespace sv {

namespace {

template <int T, bool RECORD>
__global__ void __launch_bounds__(T) k_batch_substep(BatchArgs A) {
    constexpr int W = T / 64;
    __shared__ float4 s_lo[NB_MAX], s_hi[NB_MAX];
    __shared__ uint32_t s_row[NB_MAX];
    __shared__ uint32_t s_key[INC_MAX];
    __shared__ uint16_t s_rank[INC_MAX];
    __shared__ uint16_t s_start[NB_MAX], s_deg[NB_MAX];
    __shared__ uint32_t s_ticket[NB_MAX];
    __shared__ float s_vel[3 * NB_MAX], s_omega[3 * NB_MAX];
    __shared__ float s_lambda[C_MAX];
    __shared__ uint32_t s_w[W];
    __shared__ uint32_t s_q, s_next;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    // phase timing: thread 0 stamps the wall clock after each phase's
    // barrier (every thread has finished the phase) and keeps the sums
    const bool timing = A.phase_ticks != nullptr;
    uint64_t ph[BATCH_PHASES] = {}, t_mark = timing && tid == 0 ? wall_clock64() : 0ull, n_done = 0;
    auto stamp = [&](int k) {
        if (timing && tid == 0) {
            const uint64_t now = wall_clock64();
            ph[k] += now - t_mark;
            t_mark = now;
        }
    };

    for (;;) {
        if (tid == 0) {
            uint32_t q = atomicAdd(A.queue, 1u);
            if (__hip_atomic_load(A.abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) q = A.num_scenes;
            s_q = q;
        }
        __syncthreads();
        const uint32_t q = s_q;
        __syncthreads();
        stamp(0);
        if (q >= A.num_scenes) {
            if (timing && tid == 0) {
#pragma unroll
                for (int k = 0; k < BATCH_PHASES; k++) atomicAdd(A.phase_ticks + k, (unsigned long long)ph[k]);
                atomicAdd(A.phase_ticks + BATCH_PHASES, (unsigned long long)n_done);
            }
            return;
        }
        n_done++;
        const uint32_t sc = A.sc_order[q];
        const uint32_t bs = A.sc_body[sc], ns = A.sc_body[sc + 1] - bs;
        const float gx = A.sc_grav[3 * sc], gy = A.sc_grav[3 * sc + 1], gz = A.sc_grav[3 * sc + 2];
        uint32_t *cnt = A.sc_cnt + 3 * sc;

        // step 0: gravity (the recording form first keeps the substep's
        // starting state; each lane copies the bodies it then updates)
        for (uint32_t i = tid; i < ns; i += T) {
            if (RECORD) {
                const size_t b = bs + i, N = A.n_all;
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    A.t_state[3 * b + k] = A.pos[3 * b + k];
                    A.t_state[7 * N + 3 * b + k] = A.vel[3 * b + k];
                    A.t_state[10 * N + 3 * b + k] = A.omega[3 * b + k];
                }
#pragma unroll
                for (int k = 0; k < 4; k++) A.t_state[3 * N + 4 * b + k] = A.quat[4 * b + k];
            }
            gravity_body(bs + i, A.dt, gx, gy, gz, A.ld, A.ad, A.inv_mass, A.vel, A.omega);
        }

        // AABB, rotation cache and broadphase boxes: one wave per body
        for (uint32_t i = wv; i < ns; i += W) {
            const uint32_t b = bs + i;
            m3 R;
            f3 mn, mx;
            body_aabb_wave(b, lane, A.pos, A.quat, A.hrange, A.hull, R, mn, mx);
            if (lane == 0) {
                st3(A.amin, b, mn);
                st3(A.amax, b, mx);
                st_rot(A.rot, b, R);
                float st = A.inv_mass[b] <= 0.0f ? 1.0f : 0.0f;
                s_lo[i] = make_float4(mn.x, mn.y, mn.z, st);
                s_hi[i] = make_float4(mx.x, mx.y, mx.z, 0.0f);
            }
        }
        __syncthreads();
        stamp(1);

        // broadphase: hits per row i (j > i, not both static), scanned into
        // lexicographic slots, then placed.  Pair tile: a lane per row, or a
        // group of G lanes (G | 64) per row taking G candidates a step, its
        // ballot's prefix counts keeping the row's order.
        const uint32_t G = A.pair_cols;
        const uint32_t lane = tid & 63u, gl = lane % G;
        const uint64_t gmask = G >= 64 ? ~0ull : ((1ull << G) - 1ull) << (lane - gl);
        const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
        if (G <= 1) {
            for (uint32_t i = tid; i < ns; i += T) {
                const float4 lo = s_lo[i], hi = s_hi[i];
                const bool si = lo.w != 0.0f;
                uint32_t c = 0;
                for (uint32_t j = i + 1; j < ns; j++) {
                    const float4 l2 = s_lo[j], h2 = s_hi[j];
                    if (!(si && l2.w != 0.0f) && !boxes_apart(lo, hi, l2, h2)) c++;
                }
                s_row[i] = c;
            }
        } else {
            for (uint32_t i = tid / G; i < ns; i += T / G) {
                const float4 lo = s_lo[i], hi = s_hi[i];
                const bool si = lo.w != 0.0f;
                uint32_t c = 0;
                for (uint32_t j0 = i + 1; j0 < ns; j0 += G) {
                    const uint32_t j = j0 + gl;
                    bool hit = false;
                    if (j < ns) {
                        const float4 l2 = s_lo[j], h2 = s_hi[j];
                        hit = !(si && l2.w != 0.0f) && !boxes_apart(lo, hi, l2, h2);
                    }
                    c += (uint32_t)__popcll(__ballot(hit) & gmask);
                }
                if (gl == 0) s_row[i] = c;
            }
        }
        __syncthreads();
        uint32_t npt;
        {
            const uint32_t per = (ns + T - 1) / T;
            const uint32_t r0 = min((uint32_t)tid * per, ns), r1 = min(r0 + per, ns);
            uint32_t sum = 0;
            for (uint32_t r = r0; r < r1; r++) sum += s_row[r];
            uint32_t base = block_scan<T>(sum, s_w, &npt);
            for (uint32_t r = r0; r < r1; r++) {
                uint32_t v = s_row[r];
                s_row[r] = base;
                base += v;
            }
        }
        __syncthreads();
        if (tid == 0) {
            cnt[0] = npt;
            if (npt > A.ps) cnt[2] |= SV_STATUS_PAIR_OVERFLOW;
        }
        uint32_t *pairs = A.pairs + 2 * (size_t)sc * A.ps;
        if (G <= 1) {
            for (uint32_t i = tid; i < ns; i += T) {
                uint32_t o = s_row[i];
                if (o >= A.ps) continue;
                const float4 lo = s_lo[i], hi = s_hi[i];
                const bool si = lo.w != 0.0f;
                for (uint32_t j = i + 1; j < ns && o < A.ps; j++) {
                    const float4 l2 = s_lo[j], h2 = s_hi[j];
                    if (!(si && l2.w != 0.0f) && !boxes_apart(lo, hi, l2, h2)) {
                        pairs[2 * o] = bs + i;
                        pairs[2 * o + 1] = bs + j;
                        o++;
                    }
                }
            }
        } else {
            for (uint32_t i = tid / G; i < ns; i += T / G) {
                uint32_t o = s_row[i];
                if (o >= A.ps) continue;
                const float4 lo = s_lo[i], hi = s_hi[i];
                const bool si = lo.w != 0.0f;
                for (uint32_t j0 = i + 1; j0 < ns && o < A.ps; j0 += G) {
                    const uint32_t j = j0 + gl;
                    bool hit = false;
                    if (j < ns) {
                        const float4 l2 = s_lo[j], h2 = s_hi[j];
                        hit = !(si && l2.w != 0.0f) && !boxes_apart(lo, hi, l2, h2);
                    }
                    const uint64_t m = __ballot(hit) & gmask;
                    const uint32_t q = o + (uint32_t)__popcll(m & below);
                    if (hit && q < A.ps) {
                        pairs[2 * q] = bs + i;
                        pairs[2 * q + 1] = bs + j;
                    }
                    o += (uint32_t)__popcll(m);
                }
            }
        }
        __syncthreads();
        stamp(2);

        // narrowphase: MPR per pair, contacts in pair order
        const uint32_t np = min(npt, A.ps);
        const size_t cseg = (size_t)sc * A.cs;
        const Geo g{A.hull, A.sl};
        uint32_t nct = 0;
        for (uint32_t p0 = 0; p0 < np; p0 += T) {
            const uint32_t p = p0 + tid;
            bool hit = false;
            f3 nrm = mk(0.0f, 0.0f, 0.0f), cp = mk(0.0f, 0.0f, 0.0f);
            float depth = 0.0f;
            uint32_t a = 0, b = 0;
            if (p < np) {
                a = pairs[2 * p];
                b = pairs[2 * p + 1];
                Pair P;
                P.A = load_body(a, A.pos, A.rot, A.hrange, A.sl_off);
                P.Bb = load_body(b, A.pos, A.rot, A.hrange, A.sl_off);
                hit = mpr(P, g, nrm, depth, cp) && depth > 0.0f;
            }
            uint32_t chunk;
            const uint32_t slot = nct + block_scan<T>(hit ? 1u : 0u, s_w, &chunk);
            if (hit && slot < A.cs) {
                const size_t k = cseg + slot;
                st3(A.cpos, k, cp);
                st3(A.cnrm, k, nrm);
                A.cpen[k] = depth;
                A.cga[k] = a;
                A.cgb[k] = b;
            }
            nct += chunk;
        }
        if (tid == 0) {
            cnt[1] = nct;
            if (nct > A.cs) cnt[2] |= SV_STATUS_CONTACT_OVERFLOW;
        }
        const uint32_t nc = min(nct, A.cs);
        if (RECORD && tid == 0) A.t_nc[sc] = nc;
        __syncthreads();
        stamp(3);

        // sweep constants and incidence keys; impulses start at zero (step 4)
        for (uint32_t c = tid; c < nc; c += T) {
            const size_t k = cseg + c;
            const uint32_t ga = A.cga[k], gb = A.cgb[k];
            const uint32_t a = A.gbody[ga], b = A.gbody[gb];
            const uint32_t flags =
                contact_prep(a, b, ga, gb, ld3(A.cnrm, k), ld3(A.cpos, k), A.cpen[k], A.dt, A.baumgarte, A.slop,
                             A.pos, A.vel, A.omega, A.inv_mass, A.inv_i, A.rot, A.fric, A.rest,
                             A.cprep + k * CPREP_WORDS);
            A.cprep[k * CPREP_WORDS + P_GEOM_A] = __uint_as_float(ga);
            A.cprep[k * CPREP_WORDS + P_GEOM_B] = __uint_as_float(gb);
            const bool live = !(flags & F_SKIP);
            s_key[2 * c] = (live && (flags & F_DYN_A)) ? (((a - bs) << 12) | (2 * c)) : 0xFFFFFFFFu;
            s_key[2 * c + 1] = (live && (flags & F_DYN_B)) ? (((b - bs) << 12) | (2 * c + 1)) : 0xFFFFFFFFu;
            s_lambda[c] = 0.0f;
        }
        uint32_t L = 2;
        while (L < 2 * nc) L <<= 1;
        for (uint32_t k = 2 * nc + tid; k < L; k += T) s_key[k] = 0xFFFFFFFFu;
        for (uint32_t i = tid; i < ns; i += T) {
            s_ticket[i] = 0u;
            const uint32_t b = bs + i;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                s_vel[3 * i + k] = A.vel[3 * b + k];
                s_omega[3 * i + k] = A.omega[3 * b + k];
            }
        }
        __syncthreads();
        rank_incidences<T>(s_key, s_rank, s_start, s_deg, L);
        if (tid == 0) s_next = 0u;
        __syncthreads();
        stamp(4);

        // ordered sweep (physics.comp:523-582) over items (t, c) in order
        const uint32_t total = A.iters * nc;
        {
            Item it;
            bool has = false, fin = total == 0;
            uint64_t last = wall_clock64();
            for (;;) {
                uint64_t need = __ballot(!has && !fin);
                while (need) {
                    const uint32_t cntw = (uint32_t)__popcll(need);
                    const int leader = __builtin_ctzll(need);
                    uint32_t base = 0;
                    if (lane == leader)
                        base = __hip_atomic_fetch_add(&s_next, cntw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    base = __shfl(base, leader, 64);
                    if (!has && !fin) {
                        const uint32_t q2 = base + (uint32_t)__popcll(need & lt);
                        if (q2 >= total) {
                            fin = true;
                        } else {
                            const uint32_t t = q2 / nc, c = q2 - t * nc;
                            const float *P = A.cprep + (cseg + c) * CPREP_WORDS;
                            const uint32_t flags = __float_as_uint(P[P_FLAGS]);
                            if (!(flags & F_SKIP)) {
                                const uint32_t ga = __float_as_uint(P[P_A]), gb = __float_as_uint(P[P_B]);
                                it.c = c;
                                it.q = q2;
                                it.a = ga - bs;
                                it.b = gb - bs;
                                it.dyn_a = flags & F_DYN_A;
                                it.dyn_b = flags & F_DYN_B;
                                it.ta = it.dyn_a ? t * s_deg[it.a] + s_rank[2 * c] : 0u;
                                it.tb = it.dyn_b ? t * s_deg[it.b] + s_rank[2 * c + 1] : 0u;
                                it.n = mk(P[P_N], P[P_N + 1], P[P_N + 2]);
                                it.ra = mk(P[P_RA], P[P_RA + 1], P[P_RA + 2]);
                                it.rb = mk(P[P_RB], P[P_RB + 1], P[P_RB + 2]);
                                it.kn = P[P_KN];
                                it.bias = P[P_BIAS];
                                it.mu = P[P_MU];
                                it.ima = P[P_IMA];
                                it.imb = P[P_IMB];
                                it.Ra = ld_rot(A.rot, ga);
                                it.Rb = ld_rot(A.rot, gb);
                                it.iia = ld3(A.inv_i, ga);
                                it.iib = ld3(A.inv_i, gb);
                                has = true;
                            }
                        }
                    }
                    need = __ballot(!has && !fin);
                }
                if (!__any(has)) break;
                bool progressed = false;
                if (has) {
                    const bool ready =
                        (!it.dyn_a || __hip_atomic_load(&s_ticket[it.a], __ATOMIC_RELAXED,
                                                        __HIP_MEMORY_SCOPE_WORKGROUP) == it.ta) &&
                        (!it.dyn_b || __hip_atomic_load(&s_ticket[it.b], __ATOMIC_RELAXED,
                                                        __HIP_MEMORY_SCOPE_WORKGROUP) == it.tb);
                    if (ready) {
                        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                        if (RECORD) {  // the item's inputs, for the reverse sweep
                            float *tp = A.t_items + ((size_t)sc * A.iters * A.cs + it.q) * TAPE_ITEM;
                            st3(tp, 0, ld3(s_vel, it.a));
                            st3(tp, 1, ld3(s_omega, it.a));
                            st3(tp, 2, ld3(s_vel, it.b));
                            st3(tp, 3, ld3(s_omega, it.b));
                            tp[12] = s_lambda[it.c];
                        }
                        apply_item(it, s_vel, s_omega, s_lambda);
                        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                        if (it.dyn_a)
                            __hip_atomic_fetch_add(&s_ticket[it.a], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                        if (it.dyn_b)
                            __hip_atomic_fetch_add(&s_ticket[it.b], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                        has = false;
                        progressed = true;
                    }
                }
                if (__any(progressed)) {
                    last = wall_clock64();
                } else {
                    if (__hip_atomic_load(A.abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) break;
                    if (wall_clock64() - last > STALL_TICKS) {
                        __hip_atomic_store(A.abort_flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        if (lane == 0) atomicOr(cnt + 2, (uint32_t)SV_STATUS_SOLVER_TIMEOUT);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
        }
        __syncthreads();
        stamp(5);
        for (uint32_t i = tid; i < ns; i += T) {
            const uint32_t b = bs + i;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                A.vel[3 * b + k] = s_vel[3 * i + k];
                A.omega[3 * b + k] = s_omega[3 * i + k];
                if (RECORD) A.t_w2[3 * b + k] = s_omega[3 * i + k];
            }
        }
        for (uint32_t c = tid; c < A.cs; c += T) A.lambda[cseg + c] = c < nc ? s_lambda[c] : 0.0f;
        __syncthreads();

        // step 6: integrate
        for (uint32_t i = tid; i < ns; i += T) integrate_body(bs + i, A.dt, A.inv_mass, A.pos, A.quat, A.vel, A.omega);
        __syncthreads();
        stamp(6);
    }
}

}  // namespace

typedef void (*BatchKernel)(BatchArgs);

BatchKernel batch_kernel(uint32_t threads, bool record) {
    switch (threads) {
    case 64: return record ? k_batch_substep<64, true> : k_batch_substep<64, false>;
    case 128: return record ? k_batch_substep<128, true> : k_batch_substep<128, false>;
    case 256: return record ? k_batch_substep<256, true> : k_batch_substep<256, false>;
    case 512: return record ? k_batch_substep<512, true> : k_batch_substep<512, false>;
    default: return nullptr;
    }
}

void batch_launch_substep(const BatchArgs &A, uint32_t threads, uint32_t grid, bool record, hipStream_t stream) {
    hipLaunchKernelGGL(batch_kernel(threads, record), dim3(grid), dim3(threads), 0, stream, A);
}

}  // namespace sv

namespace sv {

int batch_dev_fail(hipError_t e, const char *what) {
    char buf[256];
    snprintf(buf, sizeof(buf), "%s: %s", what, hipGetErrorString(e));
    return set_error(SV_E_DEVICE, buf);
}

BatchArgs batch_args(sv_batch *b, float dt, uint32_t *queue) {
    sv_ctx *c = b->ctx;
    auto F = [&](int id) { return reinterpret_cast<float *>(c->buf[id]); };
    auto U = [&](int id) { return reinterpret_cast<uint32_t *>(c->buf[id]); };
    BatchArgs A;
    A.sc_body = b->sc_body;
    A.sc_grav = b->sc_grav;
    A.sc_order = b->sc_order;
    A.sc_cnt = b->sc_cnt;
    A.num_scenes = b->num_scenes;
    A.queue = queue;
    A.abort_flag = b->abort_flag;
    A.pos = F(SV_BODY_POS);
    A.quat = F(SV_BODY_QUAT);
    A.vel = F(SV_BODY_VEL);
    A.omega = F(SV_BODY_OMEGA);
    A.inv_mass = F(SV_BODY_INV_MASS);
    A.inv_i = F(SV_BODY_INV_INERTIA);
    A.hull = F(SV_HULL_VERTS);
    A.fric = F(SV_GEOM_FRICTION);
    A.rest = F(SV_GEOM_RESTITUTION);
    A.hrange = c->s.hrange;
    A.sl_off = U(SV_SUPPORT_LOOKUP_OFFSET);
    A.sl = U(SV_SUPPORT_LOOKUP);
    A.gbody = U(SV_GEOM_BODY_IDX);
    A.amin = F(SV_BODY_AABB_MIN);
    A.amax = F(SV_BODY_AABB_MAX);
    A.rot = c->s.rot;
    A.pairs = U(SV_COLLISION_PAIRS);
    A.ps = b->caps.max_scene_pairs;
    A.cpos = F(SV_CONTACT_POS);
    A.cnrm = F(SV_CONTACT_NORMAL);
    A.cpen = F(SV_CONTACT_PENETRATION);
    A.lambda = F(SV_CONTACT_LAMBDA_N);
    A.cga = U(SV_CONTACT_GEOM_A);
    A.cgb = U(SV_CONTACT_GEOM_B);
    A.cs = b->caps.max_scene_contacts;
    A.cprep = c->s.cprep;
    A.dt = dt;
    A.ld = c->params.linear_damping;
    A.ad = c->params.angular_damping;
    A.baumgarte = c->params.baumgarte;
    A.slop = c->params.slop;
    A.iters = c->params.pgs_iterations;
    A.t_items = nullptr;
    A.t_w2 = nullptr;
    A.t_nc = nullptr;
    A.t_state = nullptr;
    A.n_all = b->num_bodies;
    A.phase_ticks = b->phase_timing ? b->phase_ticks : nullptr;
    A.pair_cols = b->pair_cols;
    return A;
}

// Substeps of every scene on the context stream (no render vertices).  The
// scenes' hulls are concatenated, so their offsets pass 2^24 for a few
// thousand scenes, where geom_data's float offsets (data.zig:468-473) are
// no longer exact: the batch always reads the exact SV_GEOM_HULL_RANGE (a
// scene alone has small, exact offsets, and must give the same result).
void enqueue_substeps(sv_batch *b, uint32_t substeps, float sub_dt, uint32_t threads, uint32_t wgs_per_cu) {
    sv_ctx *c = b->ctx;
    launch_hull_ranges(c, b->num_bodies, true);
    fill_u32(c, b->queue, 0u, substeps);
    uint32_t grid = (uint32_t)b->cus * wgs_per_cu;
    if (grid > b->num_scenes) grid = b->num_scenes;
    if (grid == 0) grid = 1;
    for (uint32_t s = 0; s < substeps; s++)
        batch_launch_substep(batch_args(b, sub_dt, b->queue + s), threads, grid, false, c->stream);
}

int batch_finish(sv_batch *b) {
    hipError_t e = hipStreamSynchronize(b->ctx->stream);
    if (e != hipSuccess) return batch_dev_fail(e, "batched step");
    e = hipGetLastError();
    if (e != hipSuccess) return batch_dev_fail(e, "batched step launch");
    uint32_t ab = 0;
    e = hipMemcpy(&ab, b->abort_flag, sizeof(ab), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return batch_dev_fail(e, "batched step status");
    if (ab) {
        (void)zero_device(b->abort_flag, sizeof(uint32_t));
        return set_error(SV_E_TIMEOUT, "batched step: a scene's contact sweep made no progress for 2 s");
    }
    return SV_OK;
}

}  // namespace sv

namespace {

using sv::set_error;

int dev_fail(hipError_t e, const char *what) { return sv::batch_dev_fail(e, what); }

void drop_graph(sv_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->graph_valid = false;
}

void enqueue_frame(sv_batch *b, uint32_t nv, uint32_t substeps, float dt) {
    sv_ctx *c = b->ctx;
    sv::fill_u32(c, b->sc_cnt, 0u, 3 * (size_t)b->num_scenes);
    sv::enqueue_substeps(b, substeps, dt / (float)substeps, b->threads, b->wgs_per_cu);
    sv::launch_update_verts(c, nv, b->num_bodies);
}

bool same_params(const sv_params &a, const sv_params &b) {
    return a.pgs_iterations == b.pgs_iterations && a.baumgarte == b.baumgarte && a.slop == b.slop &&
           a.linear_damping == b.linear_damping && a.angular_damping == b.angular_damping &&
           a.hull_range_u32 == b.hull_range_u32;
}

void free_batch(sv_batch *b) {
    if (!b) return;
    if (b->ctx) (void)hipStreamSynchronize(b->ctx->stream);
    drop_graph(b);
    void *ptrs[] = {b->sc_body, b->sc_order, b->sc_cnt, b->queue, b->abort_flag, b->sc_grav, b->snap, b->phase_ticks,
                    b->d_ckpt, b->d_state, b->d_w2, b->d_prep, b->d_items, b->d_nc, b->d_adj,
                    b->d_contrib, b->d_amat, b->d_amat32, b->d_adj32};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (hipEvent_t e : b->d_ev)
        if (e) (void)hipEventDestroy(e);
    if (b->d_stream) {
        (void)hipStreamSynchronize(b->d_stream);
        (void)hipStreamDestroy(b->d_stream);
    }
    if (b->ctx) sv_destroy(b->ctx);
    delete b;
}

}  // namespace

extern "C" {

int sv_batch_create(const sv_batch_caps *caps, const sv_params *params, sv_batch **out) {
    if (!caps || !out) return set_error(SV_E_INVALID, "sv_batch_create: null argument");
    *out = nullptr;
    const sv_batch_caps &k = *caps;
    if (!k.max_scenes || !k.max_bodies || !k.max_scene_bodies)
        return set_error(SV_E_INVALID, "sv_batch_create: zero capacity");
    if (k.max_scene_bodies > SV_BATCH_MAX_SCENE_BODIES || k.max_scene_contacts > SV_BATCH_MAX_SCENE_CONTACTS)
        return set_error(SV_E_INVALID, "sv_batch_create: per-scene capacity above the LDS tile");
    if ((uint64_t)k.max_scenes * k.max_scene_pairs >= 0x7FFFFFFFull ||
        (uint64_t)k.max_scenes * k.max_scene_contacts >= 0x3FFFFFFFull)
        return set_error(SV_E_INVALID, "sv_batch_create: pair or contact slots overflow 32-bit indices");
    sv_batch *b = new (std::nothrow) sv_batch();
    if (!b) return set_error(SV_E_NOMEM, "sv_batch_create: out of host memory");
    b->caps = k;
    sv_caps cc;
    sv_default_caps(&cc);
    cc.max_bodies = k.max_bodies;
    cc.max_hull_verts = k.max_hull_verts;
    cc.max_vertices = k.max_vertices;
    cc.max_triangles = k.max_triangles ? k.max_triangles : 1;
    cc.max_pairs = k.max_scenes * k.max_scene_pairs;
    cc.max_contacts = k.max_scenes * k.max_scene_contacts;
    cc.max_support_entries = k.max_support_entries;
    int rc = sv_create(&cc, params, &b->ctx);
    if (rc) {
        delete b;
        return rc;
    }
    hipError_t e = hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, b->ctx->device);
    const size_t S = k.max_scenes;
    if (e == hipSuccess) e = hipMalloc(&b->sc_body, (S + 1) * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&b->sc_order, S * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&b->sc_cnt, 3 * S * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&b->sc_grav, 3 * S * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&b->queue, sv::MAX_SUBSTEPS * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&b->abort_flag, sizeof(uint32_t));
    if (e == hipSuccess) e = sv::zero_device(b->abort_flag, sizeof(uint32_t));
    if (e == hipSuccess) e = sv::zero_device(b->sc_cnt, 3 * S * sizeof(uint32_t));
    if (e != hipSuccess) {
        free_batch(b);
        return set_error(SV_E_NOMEM, "sv_batch_create: device allocation failed");
    }
    if (b->cus <= 0) b->cus = 1;
    *out = b;
    return SV_OK;
}

int sv_batch_destroy(sv_batch *b) {
    free_batch(b);
    return SV_OK;
}

sv_ctx *sv_batch_context(sv_batch *b) { return b ? b->ctx : nullptr; }

int sv_batch_set_scenes(sv_batch *b, uint32_t num_scenes, const uint32_t *body_start, const float *gravity) {
    if (!b || !body_start || !gravity) return set_error(SV_E_INVALID, "sv_batch_set_scenes: null argument");
    if (num_scenes > b->caps.max_scenes) return set_error(SV_E_CAPACITY, "sv_batch_set_scenes: too many scenes");
    std::vector<uint32_t> order(num_scenes);
    for (uint32_t s = 0; s < num_scenes; s++) {
        if (body_start[s + 1] < body_start[s])
            return set_error(SV_E_INVALID, "sv_batch_set_scenes: body_start must not decrease");
        if (body_start[s + 1] - body_start[s] > b->caps.max_scene_bodies)
            return set_error(SV_E_CAPACITY, "sv_batch_set_scenes: a scene exceeds max_scene_bodies");
    }
    if (num_scenes && body_start[num_scenes] > b->caps.max_bodies)
        return set_error(SV_E_CAPACITY, "sv_batch_set_scenes: bodies exceed max_bodies");
    // largest scenes first, so the last workgroups to finish hold small scenes
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
        return body_start[x + 1] - body_start[x] > body_start[y + 1] - body_start[y];
    });
    hipStream_t st = b->ctx->stream;
    hipError_t e = hipMemcpyAsync(b->sc_body, body_start, (num_scenes + 1) * sizeof(uint32_t),
                                  hipMemcpyHostToDevice, st);
    if (e == hipSuccess && num_scenes)
        e = hipMemcpyAsync(b->sc_order, order.data(), num_scenes * sizeof(uint32_t), hipMemcpyHostToDevice, st);
    if (e == hipSuccess && num_scenes)
        e = hipMemcpyAsync(b->sc_grav, gravity, 3 * (size_t)num_scenes * sizeof(float), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return dev_fail(e, "sv_batch_set_scenes");
    b->num_scenes = num_scenes;
    b->num_bodies = num_scenes ? body_start[num_scenes] : 0;
    drop_graph(b);
    return SV_OK;
}

int sv_batch_set_config(sv_batch *b, uint32_t threads, uint32_t wgs_per_cu, int use_graph) {
    if (!b) return set_error(SV_E_INVALID, "sv_batch_set_config: null batch");
    if (!sv::batch_kernel(threads, false)) return set_error(SV_E_INVALID, "sv_batch_set_config: threads must be 64-512");
    if (wgs_per_cu < 1 || wgs_per_cu > 16) return set_error(SV_E_INVALID, "sv_batch_set_config: 1-16 per CU");
    b->threads = threads;
    b->wgs_per_cu = wgs_per_cu;
    b->use_graph = use_graph != 0;
    drop_graph(b);
    return SV_OK;
}

int sv_batch_step(sv_batch *b, uint32_t nv, uint32_t substeps, float dt) {
    if (!b) return set_error(SV_E_INVALID, "sv_batch_step: null batch");
    if (!b->num_scenes || !substeps) return SV_OK;
    if (substeps > sv::MAX_SUBSTEPS) return set_error(SV_E_INVALID, "sv_batch_step: too many substeps");
    if (nv > b->caps.max_vertices) return set_error(SV_E_CAPACITY, "sv_batch_step: vertices exceed capacity");
    sv_ctx *c = b->ctx;
    if (!b->use_graph) {
        enqueue_frame(b, nv, substeps, dt);
        return sv::batch_finish(b);
    }
    if (!b->graph_valid || b->g_nv != nv || b->g_substeps != substeps || b->g_dt != dt ||
        !same_params(b->g_params, c->params) || b->g_epoch != c->bind_epoch) {
        drop_graph(b);
        hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return dev_fail(e, "sv_batch_step: begin capture");
        enqueue_frame(b, nv, substeps, dt);
        e = hipStreamEndCapture(c->stream, &b->graph);
        if (e != hipSuccess) return dev_fail(e, "sv_batch_step: end capture");
        e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            drop_graph(b);
            return dev_fail(e, "sv_batch_step: instantiate");
        }
        b->graph_valid = true;
        b->g_nv = nv;
        b->g_substeps = substeps;
        b->g_dt = dt;
        b->g_params = c->params;
        b->g_epoch = c->bind_epoch;
    }
    hipError_t e = hipGraphLaunch(b->exec, c->stream);
    if (e != hipSuccess) return dev_fail(e, "sv_batch_step: graph launch");
    return sv::batch_finish(b);
}

int sv_batch_set_phase_timing(sv_batch *b, int on) {
    if (!b) return set_error(SV_E_INVALID, "sv_batch_set_phase_timing: null batch");
    if (on && !b->phase_ticks) {
        hipError_t e = hipMalloc(&b->phase_ticks, (sv::BATCH_PHASES + 1) * sizeof(unsigned long long));
        if (e == hipSuccess)
            e = sv::zero_device(b->phase_ticks, (sv::BATCH_PHASES + 1) * sizeof(unsigned long long));
        if (e != hipSuccess) {
            if (b->phase_ticks) (void)hipFree(b->phase_ticks);
            b->phase_ticks = nullptr;
            return dev_fail(e, "sv_batch_set_phase_timing");
        }
    }
    if (b->phase_timing != (on != 0)) drop_graph(b);  // the captured launches carry the pointer
    b->phase_timing = on != 0;
    return SV_OK;
}

int sv_batch_phase_times(sv_batch *b, double *us, uint64_t *scenes) {
    if (!b || !us || !scenes) return set_error(SV_E_INVALID, "sv_batch_phase_times: null argument");
    unsigned long long t[sv::BATCH_PHASES + 1] = {};
    if (b->phase_ticks) {
        hipError_t e = hipStreamSynchronize(b->ctx->stream);
        if (e == hipSuccess) e = hipMemcpy(t, b->phase_ticks, sizeof(t), hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = sv::zero_device(b->phase_ticks, sizeof(t));
        if (e != hipSuccess) return dev_fail(e, "sv_batch_phase_times");
    }
    for (int k = 0; k < sv::BATCH_PHASES; k++) us[k] = (double)t[k] * 0.01;  // 100 MHz ticks
    *scenes = t[sv::BATCH_PHASES];
    return SV_OK;
}

int sv_batch_read_counters(sv_batch *b, uint32_t *out) {
    if (!b || !out) return set_error(SV_E_INVALID, "sv_batch_read_counters: null argument");
    if (!b->num_scenes) return SV_OK;
    hipError_t e = hipMemcpy(out, b->sc_cnt, 3 * (size_t)b->num_scenes * sizeof(uint32_t), hipMemcpyDeviceToHost);
    return e == hipSuccess ? SV_OK : dev_fail(e, "sv_batch_read_counters");
}

int sv_batch_set_pair_tile(sv_batch *b, uint32_t lanes_per_row) {
    if (!b) return set_error(SV_E_INVALID, "sv_batch_set_pair_tile: null batch");
    if (lanes_per_row < 1 || lanes_per_row > 64 || (lanes_per_row & (lanes_per_row - 1)))
        return set_error(SV_E_INVALID, "sv_batch_set_pair_tile: lanes per row must be 1, 2, 4, ..., 64");
    if (b->pair_cols != lanes_per_row) drop_graph(b);  // the captured launches carry it
    b->pair_cols = lanes_per_row;
    return SV_OK;
}

int sv_batch_pair_tile(sv_batch *b, uint32_t *lanes_per_row) {
    if (!b || !lanes_per_row) return set_error(SV_E_INVALID, "sv_batch_pair_tile: null argument");
    *lanes_per_row = b->pair_cols;
    return SV_OK;
}

int sv_batch_autotune(sv_batch *b, uint32_t substeps, float dt, uint32_t *threads_out, uint32_t *wgs_out,
                      double *us_out) {
    if (!b) return set_error(SV_E_INVALID, "sv_batch_autotune: null batch");
    if (!b->num_scenes) return set_error(SV_E_INVALID, "sv_batch_autotune: no scenes");
    if (!substeps || substeps > sv::MAX_SUBSTEPS) return set_error(SV_E_INVALID, "sv_batch_autotune: substeps");
    sv_ctx *c = b->ctx;
    const size_t N = b->num_bodies, S = b->num_scenes;
    if (!b->snap && hipMalloc(&b->snap, 13 * (size_t)b->caps.max_bodies * sizeof(float) +
                                            3 * (size_t)b->caps.max_scenes * sizeof(uint32_t)) != hipSuccess)
        return set_error(SV_E_NOMEM, "sv_batch_autotune: snapshot allocation failed");
    struct Piece { void *dev; size_t bytes; };
    const Piece pieces[] = {{c->buf[SV_BODY_POS], 12 * N}, {c->buf[SV_BODY_QUAT], 16 * N},
                            {c->buf[SV_BODY_VEL], 12 * N}, {c->buf[SV_BODY_OMEGA], 12 * N},
                            {b->sc_cnt, 12 * S}};
    auto copy_all = [&](bool save) {
        char *p = reinterpret_cast<char *>(b->snap);
        hipError_t e = hipSuccess;
        for (const Piece &q : pieces) {
            if (e == hipSuccess)
                e = save ? hipMemcpyAsync(p, q.dev, q.bytes, hipMemcpyDeviceToDevice, c->stream)
                         : hipMemcpyAsync(q.dev, p, q.bytes, hipMemcpyDeviceToDevice, c->stream);
            p += q.bytes;
        }
        return e;
    };
    hipError_t e = copy_all(true);
    if (e != hipSuccess) return dev_fail(e, "sv_batch_autotune: snapshot");
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        return set_error(SV_E_DEVICE, "sv_batch_autotune: events");
    const float sub_dt = dt / (float)substeps;
    double best = 1e30;
    uint32_t bt = b->threads, bw = b->wgs_per_cu;
    const uint32_t cand_t[] = {64, 128, 256, 512};
    const uint32_t cand_g[] = {1, 4, 16, 64};  // broadphase pair tiles: lanes per row
    const uint32_t g_user = b->pair_cols;
    uint32_t bg = g_user;
    int rc = SV_OK;
    // one trial: `substeps` substeps of shape (t, w) with pair tile g, timed, state restored
    auto trial = [&](uint32_t t, uint32_t w, uint32_t g, double *us) -> int {
        b->pair_cols = g;
        sv::enqueue_substeps(b, 1, sub_dt, t, w);  // warm the shape once
        (void)hipEventRecord(e0, c->stream);
        sv::enqueue_substeps(b, substeps, sub_dt, t, w);
        (void)hipEventRecord(e1, c->stream);
        e = copy_all(false);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) return dev_fail(e, "sv_batch_autotune: trial");
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        *us = 1000.0 * ms / substeps;
        return (int)SV_OK;
    };
    // the launch shape first (threads x workgroups per CU, the pair tile as
    // set), then the pair tile at the fastest shape
    for (uint32_t t : cand_t) {
        int occ = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, reinterpret_cast<const void *>(sv::batch_kernel(t, false)),
                                                          (int)t, 0) != hipSuccess || occ < 1)
            continue;
        for (uint32_t w = 1; w <= (uint32_t)occ && w <= 8 && !rc; w++) {
            double us = 0.0;
            rc = trial(t, w, g_user, &us);
            if (!rc && us < best) {
                best = us;
                bt = t;
                bw = w;
            }
        }
        if (rc) break;
    }
    for (uint32_t g : cand_g) {
        if (rc || g == g_user) continue;
        double us = 0.0;
        rc = trial(bt, bw, g, &us);
        if (!rc && us < best) {
            best = us;
            bg = g;
        }
    }
    b->pair_cols = rc ? g_user : bg;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    rc = sv::batch_finish(b);
    if (rc) return rc;
    b->threads = bt;
    b->wgs_per_cu = bw;
    drop_graph(b);
    if (threads_out) *threads_out = bt;
    if (wgs_out) *wgs_out = bw;
    if (us_out) *us_out = best;
    return SV_OK;
}

}  // extern "C"
