// sv_physics.hip - gfx950 kernels for the per-substep stages of seeVR's
// physics.comp: gravity, per-body AABB, N^2 broadphase, MPR narrowphase,
// integration, render vertices.  The ordered PGS solver is in sv_pgs.hip.
#include "sv_internal.h"
#include "sv_math.h"
#include "sv_dev.h"

namespace sv {

namespace {

constexpr int BLOCK = 256;

template <typename T>
inline T *B(sv_ctx *c, int id) { return reinterpret_cast<T *>(c->buf[id]); }

// ---------------------------------------------------------------------------
// Hull ranges: geom_data stores (start, count) as floats (data.zig:468-473,
// read back with uint() at physics.comp:100-101); the extension buffer holds
// them as uints.  Kernels read the uint copy.
__global__ void k_hull_ranges(const float *geom_data, const uint32_t *user, int use_u32, uint32_t *out,
                              uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (use_u32) {
        out[2 * i] = user[2 * i];
        out[2 * i + 1] = user[2 * i + 1];
    } else {
        out[2 * i] = (uint32_t)geom_data[(size_t)i * 8];
        out[2 * i + 1] = (uint32_t)geom_data[(size_t)i * 8 + 1];
    }
}

// Step 0 (physics.comp:418-428): damped velocities plus gravity, dynamic only.
// With `counters` given, thread 0 also clears the pair and contact counters
// (the full step's clear, physics.zig's first dispatch, touches nothing
// gravity reads or writes).
__global__ void k_gravity(uint32_t n, float dt, float gx, float gy, float gz, float ld, float ad,
                          const float *__restrict__ inv_mass, float *__restrict__ vel, float *__restrict__ omega,
                          uint32_t *__restrict__ counters) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (counters && i == 0) {
        counters[0] = 0;
        counters[1] = 0;
    }
    if (i >= n) return;
    gravity_body(i, dt, gx, gy, gz, ld, ad, inv_mass, vel, omega);
}

// Per-body AABB, the semantics of physics.comp:431-448: min/max over
// p + R*v of the body's hull verts, seeded with +-1e10.  One DPP row (16
// lanes) per body, four bodies per wave: lanes stride the hull, six 12-byte
// loads in flight per lane, then a row reduction.
// Also caches the body's rotation columns and a packed broadphase box
// (min.xyz, static flag), (max.xyz, 0) for the later stages.
// With ga.vel set, the launch also applies k_gravity's update to the body
// (the full step's substeps, one launch fewer): the box depends on position
// and orientation only, so the two commute.
struct GravityArgs {
    float dt, gx, gy, gz, ld, ad;
    float *vel, *omega;
    uint32_t *counters;  // pair and contact counts cleared when non-null
};
__global__ void __launch_bounds__(BLOCK) k_aabb(uint32_t n, const float *__restrict__ pos,
                                                const float *__restrict__ quat, const uint32_t *__restrict__ hrange,
                                                const float *__restrict__ hull, const float *__restrict__ inv_mass,
                                                float *__restrict__ amin_out, float *__restrict__ amax_out,
                                                float *__restrict__ rot, float4 *__restrict__ bpbox, GravityArgs ga) {
    // a DPP row per body: rows only read lanes of their own row
    const uint32_t b = (blockIdx.x * blockDim.x + threadIdx.x) / AABB_ROW;
    const int g = threadIdx.x & (AABB_ROW - 1);
    if (ga.counters && blockIdx.x == 0 && threadIdx.x == 0) {
        ga.counters[0] = 0;
        ga.counters[1] = 0;
    }
    if (b >= n) return;
    m3 R;
    f3 mn, mx;
    body_aabb_row(b, g, pos, quat, hrange, hull, R, mn, mx);
    if (g == AABB_ROW - 1) {
        st3(amin_out, b, mn);
        st3(amax_out, b, mx);
        st_rot(rot, b, R);
        float st = inv_mass[b] <= 0.0f ? 1.0f : 0.0f;
        bpbox[2 * b] = make_float4(mn.x, mn.y, mn.z, st);
        bpbox[2 * b + 1] = make_float4(mx.x, mx.y, mx.z, 0.0f);
        if (ga.vel) gravity_body(b, ga.dt, ga.gx, ga.gy, ga.gz, ga.ld, ga.ad, inv_mass, ga.vel, ga.omega);
    }
}

// ---------------------------------------------------------------------------
// Broadphase (physics.comp:451-472): every (i, j > i) not both static whose
// boxes overlap (touching counts).
//
// Work is tiled in two dimensions so the triangle balances: a block takes
// 32 consecutive rows (8 per wave) against one column segment of
// bp_seg_len(n) bodies, streamed through LDS 256 boxes at a time, 64 j per
// lane-round.  Each test is branch-free (boxes_gap: six subtractions, three
// maxima, one compare); the compare's lane mask, and-ed in scalar registers
// with the lanes the row may pair with (j > i on the diagonal tile, not
// static for a static row), is the row's 64-bit hit ballot, numbered in j
// order.  Lanes past the segment or n carry an empty box.
//
// The test pass writes, per (row, segment), the hit count and the first
// BP_STAGE partners.  An exclusive scan over the counts in (row, segment)
// order gives every (row, segment) its slot range in lexicographic (i, j)
// order; k_bp_place copies staged partners there, and only tiles holding a
// (row, segment) with more than BP_STAGE hits, which the test pass lists,
// are tested again to place the rest (EMIT form).
constexpr int BP_ROWS_PER_WAVE = 8;
constexpr int BP_TILE = BLOCK;  // 256 boxes per LDS tile

template <bool EMIT>
SVD void bp_tile(uint32_t bx, uint32_t seg, uint32_t n, uint32_t nseg, uint32_t seg_len,
                 const float4 *__restrict__ box, uint32_t *__restrict__ cnt_out, const uint32_t *__restrict__ off,
                 uint32_t *__restrict__ stage, uint32_t *__restrict__ ovf_list, uint32_t *__restrict__ pairs,
                 uint32_t max_pairs, float4 *s_lo, float4 *s_hi, uint32_t *s_ovf) {
    const uint32_t blk_row0 = bx * BP_ROWS;
    const uint32_t j_lo = max(seg * seg_len, blk_row0 + 1);
    const uint32_t j_hi = min(seg * seg_len + seg_len, n);
    const uint32_t tile = bx * nseg + seg;
    if (j_lo >= j_hi) {  // no j > i in this segment: zero its counts
        uint32_t i = blk_row0 + threadIdx.x;
        if (!EMIT && threadIdx.x < BP_ROWS && i < n) cnt_out[(size_t)i * nseg + seg] = 0;
        return;
    }
    if (!EMIT && threadIdx.x == 0) *s_ovf = 0u;
    const int lane = threadIdx.x & 63;
    const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t row0 = blk_row0 + w * BP_ROWS_PER_WAVE;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const float4 empty_lo = make_float4(__builtin_inff(), __builtin_inff(), __builtin_inff(), 0.0f);
    const float4 empty_hi = make_float4(-__builtin_inff(), -__builtin_inff(), -__builtin_inff(), 0.0f);

    float4 rlo[BP_ROWS_PER_WAVE], rhi[BP_ROWS_PER_WAVE];
    uint32_t cnt[BP_ROWS_PER_WAVE];
    uint32_t base[BP_ROWS_PER_WAVE];
    uint32_t row_static = 0;  // bit r: row r is a static body (wave-uniform)
#pragma unroll
    for (int r = 0; r < BP_ROWS_PER_WAVE; r++) {
        uint32_t i = row0 + r;
        uint32_t ic = i < n ? i : 0u;
        rlo[r] = box[2 * ic];
        rhi[r] = box[2 * ic + 1];
        if (i >= n) { rlo[r] = empty_lo; rhi[r] = empty_hi; }
        row_static |= (__builtin_amdgcn_readfirstlane(__float_as_uint(rlo[r].w)) != 0u ? 1u : 0u) << r;
        // keep the row boxes in vector registers (64 scalar registers would
        // spill); the empty asm only pins the register class
        asm volatile("" : "+v"(rlo[r].x), "+v"(rlo[r].y), "+v"(rlo[r].z), "+v"(rhi[r].x), "+v"(rhi[r].y),
                     "+v"(rhi[r].z));
        cnt[r] = 0;
        base[r] = (EMIT && i < n) ? off[(size_t)i * nseg + seg] : 0u;
    }

    for (uint32_t t0 = j_lo; t0 < j_hi; t0 += BP_TILE) {
        __syncthreads();
        uint32_t jl = t0 + threadIdx.x;
        if (jl < j_hi) {
            s_lo[threadIdx.x] = box[2 * jl];
            s_hi[threadIdx.x] = box[2 * jl + 1];
        } else {
            s_lo[threadIdx.x] = empty_lo;
            s_hi[threadIdx.x] = empty_hi;
        }
        __syncthreads();
        const bool diag = t0 < row0 + BP_ROWS_PER_WAVE;  // some j <= i in this tile
#pragma unroll
        for (int ch = 0; ch < BP_TILE / 64; ch++) {
            const uint32_t j = t0 + ch * 64 + lane;
            const float4 lo = s_lo[ch * 64 + lane];
            const float4 hi = s_hi[ch * 64 + lane];
            // lanes each row may pair with: j > i on the diagonal tile, and
            // not static when the row is static (scalar masks, set per chunk)
            const uint64_t nonstatic = row_static ? ~__builtin_amdgcn_ballot_w64(lo.w != 0.0f) : ~0ull;
            uint64_t allow[BP_ROWS_PER_WAVE];
#pragma unroll
            for (int r = 0; r < BP_ROWS_PER_WAVE; r++) {
                allow[r] = diag ? __builtin_amdgcn_ballot_w64(j > row0 + r) : ~0ull;
                if (row_static & (1u << r)) allow[r] &= nonstatic;
            }
#pragma unroll
            for (int r = 0; r < BP_ROWS_PER_WAVE; r++) {
                const uint32_t i = row0 + r;
                // one compare's lane mask, and-ed with the allowed lanes in
                // scalar registers
                const uint64_t m =
                    __builtin_amdgcn_ballot_w64(!(boxes_gap(rlo[r], rhi[r], lo, hi) > 0.0f)) & allow[r];
                const bool hit = (m >> lane) & 1ull;
                if (m) {
                    const uint32_t k = cnt[r] + (uint32_t)__popcll(m & lt);
                    if (EMIT) {
                        if (hit && k >= BP_STAGE) {
                            uint32_t pos = base[r] + k;
                            if (pos < max_pairs) {
                                pairs[2 * pos] = i;
                                pairs[2 * pos + 1] = j;
                            }
                        }
                    } else if (hit && k < BP_STAGE) {
                        stage[((size_t)i * nseg + seg) * BP_STAGE + k] = j;
                    }
                    cnt[r] += (uint32_t)__popcll(m);
                }
            }
        }
    }
    if (!EMIT) {
        if (lane == 0) {
            bool ovf = false;
#pragma unroll
            for (int r = 0; r < BP_ROWS_PER_WAVE; r++) {
                uint32_t i = row0 + r;
                if (i < n) cnt_out[(size_t)i * nseg + seg] = cnt[r];
                ovf |= cnt[r] > BP_STAGE;
            }
            if (ovf) *s_ovf = 1u;
        }
        __syncthreads();
        // list the tile for the EMIT pass: ovf_list[0] counts, tiles follow
        if (threadIdx.x == 0 && *s_ovf) ovf_list[1 + atomicAdd(ovf_list, 1u)] = tile;
    }
}

// Count pass: one block per (row tile, column segment).  EMIT pass: a
// smaller grid walks the tiles the count pass listed.
template <bool EMIT>
__global__ void __launch_bounds__(BLOCK) k_broadphase(uint32_t n, uint32_t nseg, uint32_t seg_len,
                                                      const float4 *__restrict__ box, uint32_t *__restrict__ cnt_out,
                                                      const uint32_t *__restrict__ off,
                                                      uint32_t *__restrict__ stage, uint32_t *__restrict__ ovf_list,
                                                      uint32_t *__restrict__ pairs, uint32_t max_pairs) {
    __shared__ float4 s_lo[BP_TILE];
    __shared__ float4 s_hi[BP_TILE];
    __shared__ uint32_t s_ovf;
    if (!EMIT) {
        bp_tile<false>(blockIdx.x, blockIdx.y, n, nseg, seg_len, box, cnt_out, off, stage, ovf_list, pairs, max_pairs,
                       s_lo, s_hi, &s_ovf);
        return;
    }
    const uint32_t listed = ovf_list[0];
    for (uint32_t q = blockIdx.x; q < listed; q += gridDim.x) {
        const uint32_t tile = ovf_list[1 + q];
        __syncthreads();  // the previous tile's LDS reads are done
        bp_tile<true>(tile / nseg, tile % nseg, n, nseg, seg_len, box, cnt_out, off, stage, ovf_list, pairs, max_pairs,
                      s_lo, s_hi, &s_ovf);
    }
}

// Staged partners of (row i, segment s) to slots off[i*nseg+s] + k.
__global__ void __launch_bounds__(BLOCK) k_bp_place(uint32_t n, uint32_t nseg, const uint32_t *__restrict__ cnt,
                                                    const uint32_t *__restrict__ off,
                                                    const uint32_t *__restrict__ stage, uint32_t *__restrict__ pairs,
                                                    uint32_t max_pairs) {
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t rs = t / BP_STAGE;
    uint32_t k = (uint32_t)(t % BP_STAGE);
    if (rs >= (size_t)n * nseg || k >= cnt[rs]) return;
    uint32_t pos = off[rs] + k;
    if (pos < max_pairs) {
        pairs[2 * pos] = (uint32_t)(rs / nseg);
        pairs[2 * pos + 1] = stage[rs * BP_STAGE + k];
    }
}

// counters[0] = total pairs (written by the scan); clamp, flag overflow.
__global__ void k_post_pairs(uint32_t *counters, uint32_t *misc, uint32_t max_pairs) {
    uint32_t t = counters[0];
    misc[MISC_NP] = t < max_pairs ? t : max_pairs;
    if (t > max_pairs) counters[2] |= SV_STATUS_PAIR_OVERFLOW;
}

// Step 3 (physics.comp:475-491), result per pair slot: hit flag and
// (pos, normal, depth); contacts are then compacted in pair order.
//
// Divergence-sorted batches: each block takes 256 consecutive pairs (near in
// the canonical order, so they share bodies in the caches) and reorders them
// in LDS by support-path class (brute force for hulls of at most 32 verts,
// lookup above: physics.comp:107-162) before the lanes take them, so a wave
// runs pairs of one class.  Results stay in pair slots: the order of work
// has no effect on them.
__global__ void __launch_bounds__(BLOCK) k_narrow(const uint32_t *__restrict__ misc, const uint32_t *__restrict__ pairs,
                                                  const float *__restrict__ pos, const float *__restrict__ rot,
                                                  const uint32_t *__restrict__ hrange, const uint32_t *__restrict__ sl_off,
                                                  const uint32_t *__restrict__ sl, const float *__restrict__ hull,
                                                  const float *__restrict__ lxyz, uint32_t *__restrict__ flag,
                                                  float *__restrict__ res) {
    __shared__ uint32_t s_cnt[4];
    __shared__ uint32_t s_pair[BLOCK];
    uint32_t np = misc[MISC_NP];
    Geo g{hull, sl, lxyz};
    for (uint32_t p0 = blockIdx.x * blockDim.x; p0 < np; p0 += gridDim.x * blockDim.x) {
        if (threadIdx.x < 4) s_cnt[threadIdx.x] = 0;
        __syncthreads();
        const uint32_t mine = p0 + threadIdx.x;
        uint32_t k = 0, rk = 0;
        if (mine < np) {
            k = (sl_off[pairs[2 * mine]] != 0u ? 1u : 0u) | (sl_off[pairs[2 * mine + 1]] != 0u ? 2u : 0u);
            rk = atomicAdd(&s_cnt[k], 1u);
        }
        __syncthreads();
        if (mine < np) {
            uint32_t o = 0;
            for (uint32_t q = 0; q < k; q++) o += s_cnt[q];
            s_pair[o + rk] = mine;
        }
        __syncthreads();
        const uint32_t m = np - p0 < (uint32_t)BLOCK ? np - p0 : (uint32_t)BLOCK;
        if (threadIdx.x >= m) continue;
        const uint32_t p = s_pair[threadIdx.x];
        uint32_t a = pairs[2 * p], b = pairs[2 * p + 1];
        Pair P;
        P.A = load_body(a, pos, rot, hrange, sl_off);
        P.Bb = load_body(b, pos, rot, hrange, sl_off);
        f3 nrm = mk(0.0f, 0.0f, 0.0f), cp = mk(0.0f, 0.0f, 0.0f);
        float depth = 0.0f;
        bool hit = mpr(P, g, nrm, depth, cp) && depth > 0.0f;
        flag[p] = hit ? 1u : 0u;
        if (hit) {
            float *r = res + (size_t)p * 8;
            r[0] = cp.x; r[1] = cp.y; r[2] = cp.z;
            r[3] = nrm.x; r[4] = nrm.y; r[5] = nrm.z;
            r[6] = depth;
        }
    }
}

// add_contact (physics.comp:403-412) at the canonical slot.  Thread 0 also
// clamps the contact count (counters[1] holds the total) and flags an
// overflow, and with `lambda` given the kernel clears the impulses of every
// contact slot (the next step, physics.comp:494-497, which reads nothing
// written here).
__global__ void __launch_bounds__(BLOCK) k_compact_contacts(uint32_t *__restrict__ misc, uint32_t *__restrict__ counters,
                                                            const uint32_t *__restrict__ pairs,
                                                            const uint32_t *__restrict__ flag,
                                                            const uint32_t *__restrict__ idx,
                                                            const float *__restrict__ res, uint32_t max_contacts,
                                                            float *__restrict__ cpos, float *__restrict__ cnrm,
                                                            float *__restrict__ cpen, uint32_t *__restrict__ ga,
                                                            uint32_t *__restrict__ gb, float *__restrict__ lambda,
                                                            uint32_t *__restrict__ deg, uint32_t n_bodies,
                                                            uint32_t *__restrict__ ctl) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const uint32_t t = counters[1];
        const uint32_t nc = t < max_contacts ? t : max_contacts;
        misc[MISC_NC] = nc;
        if (t > max_contacts) counters[2] |= SV_STATUS_CONTACT_OVERFLOW;
        if (deg) {
            misc[MISC_NINC] = 2u * nc;
            misc[MISC_NACT] = 0u;
        }
    }
    if (deg) {  // the sweep's setup (k_sweep_setup): degrees and control words
        const uint32_t stride = gridDim.x * blockDim.x, t0 = blockIdx.x * blockDim.x + threadIdx.x;
        for (uint32_t b = t0; b < n_bodies; b += stride) deg[b] = 0u;
        for (uint32_t k = t0; k < SOLVE_CTL_WORDS; k += stride) ctl[k] = 0u;
    }
    if (lambda)
        for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < max_contacts; k += gridDim.x * blockDim.x)
            lambda[k] = 0.0f;
    uint32_t np = misc[MISC_NP];
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < np; p += gridDim.x * blockDim.x) {
        if (!flag[p]) continue;
        uint32_t k = idx[p];
        if (k >= max_contacts) continue;
        const float *r = res + (size_t)p * 8;
        cpos[3 * k] = r[0]; cpos[3 * k + 1] = r[1]; cpos[3 * k + 2] = r[2];
        cnrm[3 * k] = r[3]; cnrm[3 * k + 1] = r[4]; cnrm[3 * k + 2] = r[5];
        cpen[k] = r[6];
        ga[k] = pairs[2 * p];
        gb[k] = pairs[2 * p + 1];
    }
}

// Step 4 (physics.comp:494-497)
__global__ void k_clear_lambda(float *lambda, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) lambda[i] = 0.0f;
}

// Step 6 (physics.comp:586-604)
__global__ void k_integrate(uint32_t n, float dt, const float *__restrict__ inv_mass, float *__restrict__ pos,
                            float *__restrict__ quat, const float *__restrict__ vel,
                            const float *__restrict__ omega) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    integrate_body(i, dt, inv_mass, pos, quat, vel, omega);
}

// Step 7 (physics.comp:607-627): world = p + R * local for every render
// vertex, the owner being the last body whose body_vert_start is <= the
// vertex (the reference's binary search).  body_vert_start is a
// non-decreasing prefix (data.zig:633-640), so body b owns exactly
// [start[b], start[b + 1]) (the last body up to nv; body 0 also the
// vertices below start[0]; a body with no vertices an empty range): one row
// of 16 lanes per body walks its range, with the body's rotation computed
// once, instead of one binary search (log2 N dependent loads) per vertex.
// Same owner, same arithmetic: the same bits.
constexpr uint32_t VERT_ROW = 16, VERT_UNROLL = 8;
__global__ void __launch_bounds__(BLOCK) k_update_verts(uint32_t nv, uint32_t nb, const uint32_t *__restrict__ vstart,
                                                        const float *__restrict__ pos, const float *__restrict__ quat,
                                                        const float *__restrict__ orig, float *__restrict__ verts) {
    const uint32_t b = (blockIdx.x * blockDim.x + threadIdx.x) / VERT_ROW, r = threadIdx.x % VERT_ROW;
    if (b >= nb) return;
    const uint32_t v0 = b == 0 ? 0u : vstart[b];
    const uint32_t v1 = min(b + 1 < nb ? vstart[b + 1] : nv, nv);
    if (v0 >= v1) return;
    const m3 R = quat_mat_at(quat, b);
    const f3 p = ld3(pos, b);
    // VERT_UNROLL loads in flight per lane (none past the range's end)
    for (uint32_t v = v0 + r; v < v1; v += VERT_ROW * VERT_UNROLL) {
        f3 o[VERT_UNROLL];
#pragma unroll
        for (uint32_t k = 0; k < VERT_UNROLL; k++)
            o[k] = v + VERT_ROW * k < v1 ? ld3(orig, v + VERT_ROW * k) : mk(0.0f, 0.0f, 0.0f);
#pragma unroll
        for (uint32_t k = 0; k < VERT_UNROLL; k++) {
            if (v + VERT_ROW * k >= v1) break;
            st3(verts, v + VERT_ROW * k, p + mvec(R, o[k]));
        }
    }
}

// Step 8 (physics.comp:630-634); counters[2] (status bits) accumulates over
// a whole Physics.step call and is cleared at its start.
__global__ void k_clear_counters(uint32_t *counters) {
    counters[0] = 0;
    counters[1] = 0;
}

// Rotation cache and packed broadphase boxes for stages dispatched on their
// own (the AABB kernel writes both during a full step).
__global__ void k_rot(uint32_t n, const float *__restrict__ quat, float *__restrict__ rot) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) st_rot(rot, i, quat_mat_at(quat, i));
}

__global__ void k_pack_boxes(uint32_t n, const float *__restrict__ amin, const float *__restrict__ amax,
                             const float *__restrict__ inv_mass, float4 *__restrict__ bpbox) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float st = inv_mass[i] <= 0.0f ? 1.0f : 0.0f;
    bpbox[2 * i] = make_float4(amin[3 * i], amin[3 * i + 1], amin[3 * i + 2], st);
    bpbox[2 * i + 1] = make_float4(amax[3 * i], amax[3 * i + 1], amax[3 * i + 2], 0.0f);
}

// Plain copy of known size: the calibration kernel for HBM byte counters
// and the achievable-bandwidth reference of the roofline.
__global__ void __launch_bounds__(BLOCK) k_copy_calib(const float4 *__restrict__ src, float4 *__restrict__ dst,
                                                      size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[i];
}

// Candidate vertex of every support-lookup entry of bodies [0, n): one wave
// per body, lanes over its 3 x 64 entries (data.zig:299-353 layout).
__global__ void __launch_bounds__(BLOCK) k_sup_expand(uint32_t n, const uint32_t *__restrict__ sl_off,
                                                      const uint32_t *__restrict__ hrange,
                                                      const uint32_t *__restrict__ sl, const float *__restrict__ hull,
                                                      uint32_t n_entries, float *__restrict__ out) {
    const uint32_t b = blockIdx.x * (BLOCK / 64) + (threadIdx.x >> 6);
    if (b >= n) return;
    const uint32_t lk = sl_off[b];
    if (lk == 0) return;
    const uint32_t vs = hrange[2 * b];
    for (uint32_t e = threadIdx.x & 63; e < 192u; e += 64u) {
        if (lk + e >= n_entries) break;
        const uint32_t v = vs + sl[lk + e];
        out[3 * (size_t)(lk + e)] = hull[3 * (size_t)v];
        out[3 * (size_t)(lk + e) + 1] = hull[3 * (size_t)v + 1];
        out[3 * (size_t)(lk + e) + 2] = hull[3 * (size_t)v + 2];
    }
}

// XCC register of the first wave of each block (placement probe)
__global__ void k_xcc_ids(uint32_t *out, uint32_t n) {
    if (threadIdx.x == 0 && blockIdx.x < n) out[blockIdx.x] = xcc_id();
}

// One-way hand-off latency between two waves: workgroups 0 and `partner`
// (1: the next XCD under round-robin placement; 8: the same XCD) pass a
// device-coherent counter back and forth `rounds` times; the other
// workgroups leave.  out[0..1] = wall-clock ticks (100 MHz) of workgroup 0's
// start and end, out[2..3] = the two waves' XCC ids, out[4] = 1 when a wait
// passed 2 s (the probe then stops).
__global__ void k_handoff(uint32_t *flag, uint32_t rounds, uint32_t partner, uint64_t *out) {
    if (threadIdx.x != 0 || (blockIdx.x != 0 && blockIdx.x != partner)) return;
    const bool first = blockIdx.x == 0;
    const uint64_t limit = 200000000ull;
    out[first ? 2 : 3] = xcc_id();
    const uint64_t t0 = wall_clock64();
    bool stalled = false;
    for (uint32_t k = 0; k < rounds && !stalled; k++) {
        const uint32_t want = first ? 2u * k : 2u * k + 1u;  // the value that hands this wave the turn
        const uint64_t w0 = wall_clock64();
        for (;;) {
            const uint32_t v = __hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (v == want) break;
            if (v == 0xFFFFFFFFu || wall_clock64() - w0 > limit) {  // the other wave stopped, or no turn for 2 s
                stalled = true;
                break;
            }
        }
        if (!stalled) __hip_atomic_store(flag, want + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (stalled) {
        out[4] = 1u;
        __hip_atomic_store(flag, 0xFFFFFFFFu, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);  // release the other
    }
    if (first) {
        out[0] = t0;
        out[1] = wall_clock64();
    }
}

__global__ void k_atan2(const float *y, const float *x, float *out, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = sv_atan2f(y[i], x[i]);
}

}  // namespace

void launch_hull_ranges(sv_ctx *c, uint32_t n, bool exact) {
    if (!n) return;
    hipLaunchKernelGGL(k_hull_ranges, dim3(div_up(n, BLOCK)), dim3(BLOCK), 0, c->stream, B<float>(c, SV_GEOM_DATA),
                       B<uint32_t>(c, SV_GEOM_HULL_RANGE), (exact || c->params.hull_range_u32) ? 1 : 0, c->s.hrange, n);
}

// Scratch::sup_xyz for bodies [0, n) when it may be stale (after
// launch_hull_ranges); allocates on first use, so never call it while a
// stream is being captured.
void ensure_support_vertices(sv_ctx *c, uint32_t n) {
    const size_t L = c->caps.max_support_entries;
    if (!n || L < 2) return;
    if (!c->s.sup_xyz) {
        void *p = nullptr;
        if (hipMalloc(&p, 3 * L * sizeof(float)) != hipSuccess) return;  // supports then read through the lookup
        c->s.lazy.push_back(p);
        c->s.sup_xyz = static_cast<float *>(p);
        c->sup_dirty = true;
    }
    if (c->sup_dirty || c->sup_external || c->sup_bodies != n) {
        // one wave per body: 192 lookup entries
        hipLaunchKernelGGL(k_sup_expand, dim3(div_up(n, BLOCK / 64)), dim3(BLOCK), 0, c->stream, n,
                           B<uint32_t>(c, SV_SUPPORT_LOOKUP_OFFSET), c->s.hrange, B<uint32_t>(c, SV_SUPPORT_LOOKUP),
                           B<float>(c, SV_HULL_VERTS), (uint32_t)L, c->s.sup_xyz);
        c->sup_dirty = false;
        c->sup_bodies = n;
    }
}

void launch_gravity(sv_ctx *c, uint32_t n, float dt, const float g[3], bool clear_counters) {
    if (!n) {
        if (clear_counters) launch_clear_counters(c);
        return;
    }
    hipLaunchKernelGGL(k_gravity, dim3(div_up(n, BLOCK)), dim3(BLOCK), 0, c->stream, n, dt, g[0], g[1], g[2],
                       c->params.linear_damping, c->params.angular_damping, B<float>(c, SV_BODY_INV_MASS),
                       B<float>(c, SV_BODY_VEL), B<float>(c, SV_BODY_OMEGA),
                       clear_counters ? B<uint32_t>(c, SV_ATOMIC_COUNTERS) : (uint32_t *)nullptr);
}

static void launch_aabb_with(sv_ctx *c, uint32_t n, GravityArgs ga) {
    hipLaunchKernelGGL(k_aabb, dim3(div_up((uint64_t)n * AABB_ROW, BLOCK)), dim3(BLOCK), 0, c->stream, n,
                       B<float>(c, SV_BODY_POS), B<float>(c, SV_BODY_QUAT), c->s.hrange, B<float>(c, SV_HULL_VERTS),
                       B<float>(c, SV_BODY_INV_MASS), B<float>(c, SV_BODY_AABB_MIN), B<float>(c, SV_BODY_AABB_MAX),
                       c->s.rot, c->s.bpbox, ga);
}

void launch_aabb(sv_ctx *c, uint32_t n) {
    if (!n) return;
    launch_aabb_with(c, n, GravityArgs{0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, nullptr, nullptr, nullptr});
}

void launch_gravity_aabb(sv_ctx *c, uint32_t n, float dt, const float g[3]) {
    if (!n) {
        launch_clear_counters(c);
        return;
    }
    launch_aabb_with(c, n, GravityArgs{dt, g[0], g[1], g[2], c->params.linear_damping, c->params.angular_damping,
                                       B<float>(c, SV_BODY_VEL), B<float>(c, SV_BODY_OMEGA),
                                       B<uint32_t>(c, SV_ATOMIC_COUNTERS)});
}

int broadphase_method(const sv_ctx *c, uint32_t n) {
    if (n > BP_ALL_PAIRS_MAX_BODIES || c->broadphase_mode == SV_BROADPHASE_LBVH) return SV_BROADPHASE_LBVH;
    if (c->broadphase_mode != SV_BROADPHASE_AUTO) return (int)c->broadphase_mode;
    if (n >= BP_LBVH_MIN_BODIES) return SV_BROADPHASE_LBVH;
    return n >= BP_SWEEP_MIN_BODIES ? SV_BROADPHASE_SWEEP : SV_BROADPHASE_ALL_PAIRS;
}

void launch_broadphase(sv_ctx *c, uint32_t n) {
    uint32_t *counters = B<uint32_t>(c, SV_ATOMIC_COUNTERS);
    if (n == 0) {
        fill_u32(c, counters, 0u, 1);
        hipLaunchKernelGGL(k_post_pairs, dim3(1), dim3(1), 0, c->stream, counters, c->s.misc, c->caps.max_pairs);
        return;
    }
    // same pair list every way; without LBVH or sweep scratch the all-pairs kernel runs
    const int method = broadphase_method(c, n);
    if (method == SV_BROADPHASE_LBVH && launch_broadphase_lbvh(c, n)) return;
    if (method == SV_BROADPHASE_SWEEP && launch_broadphase_sap(c, n)) return;
    if (n > BP_ALL_PAIRS_MAX_BODIES) {  // no LBVH scratch and too many rows for the cells
        c->launch_err = hipErrorOutOfMemory;
        return;
    }
    const uint32_t seg_len = bp_seg_len(n), nseg = bp_nseg(n);
    const uint32_t nrb = div_up(n, BP_ROWS);
    uint32_t *pairs = B<uint32_t>(c, SV_COLLISION_PAIRS);
    fill_u32(c, c->s.bp_ovf, 0u, 1);
    hipLaunchKernelGGL(k_broadphase<false>, dim3(nrb, nseg), dim3(BLOCK), 0, c->stream, n, nseg, seg_len, c->s.bpbox,
                       c->s.row_count, (const uint32_t *)nullptr, c->s.bp_stage, c->s.bp_ovf, (uint32_t *)nullptr,
                       c->caps.max_pairs);
    const uint32_t ncnt = n * nseg;
    // with k_post_pairs' work (pairs to narrow, the overflow flag)
    scan_exclusive(c, c->s.row_count, c->s.row_off, nullptr, ncnt, counters, c->caps.max_pairs,
                   c->s.misc + MISC_NP, counters + 2);
    hipLaunchKernelGGL(k_bp_place, dim3(div_up((uint64_t)ncnt * BP_STAGE, BLOCK)), dim3(BLOCK), 0, c->stream, n,
                       nseg, c->s.row_count, c->s.row_off, c->s.bp_stage, pairs, c->caps.max_pairs);
    const uint32_t emit_grid = nrb * nseg < 512u ? nrb * nseg : 512u;
    hipLaunchKernelGGL(k_broadphase<true>, dim3(emit_grid), dim3(BLOCK), 0, c->stream, n, nseg, seg_len, c->s.bpbox,
                       (uint32_t *)nullptr, c->s.row_off, c->s.bp_stage, c->s.bp_ovf, pairs, c->caps.max_pairs);
}

void launch_narrowphase(sv_ctx *c, bool after_broadphase, bool clear_lambda, uint32_t sweep_bodies) {
    uint32_t *counters = B<uint32_t>(c, SV_ATOMIC_COUNTERS);
    // pairs to process: min(counters[0], max_pairs) (physics.comp:477); the
    // broadphase of the same substep has computed it already
    if (!after_broadphase)
        hipLaunchKernelGGL(k_post_pairs, dim3(1), dim3(1), 0, c->stream, counters, c->s.misc, c->caps.max_pairs);
    uint32_t mp = c->caps.max_pairs;
    uint32_t grid = div_up(mp, BLOCK);
    if (grid > 8192) grid = 8192;
    if (grid == 0) grid = 1;
    hipLaunchKernelGGL(k_narrow, dim3(grid), dim3(BLOCK), 0, c->stream, c->s.misc,
                       B<uint32_t>(c, SV_COLLISION_PAIRS), B<float>(c, SV_BODY_POS), c->s.rot, c->s.hrange,
                       B<uint32_t>(c, SV_SUPPORT_LOOKUP_OFFSET), B<uint32_t>(c, SV_SUPPORT_LOOKUP),
                       B<float>(c, SV_HULL_VERTS), c->s.sup_xyz, c->s.np_flag, c->s.np_res);
    scan_exclusive(c, c->s.np_flag, c->s.np_idx, c->s.misc + MISC_NP, mp, counters + 1);
    hipLaunchKernelGGL(k_compact_contacts, dim3(grid), dim3(BLOCK), 0, c->stream, c->s.misc, counters,
                       B<uint32_t>(c, SV_COLLISION_PAIRS), c->s.np_flag, c->s.np_idx, c->s.np_res,
                       c->caps.max_contacts, B<float>(c, SV_CONTACT_POS), B<float>(c, SV_CONTACT_NORMAL),
                       B<float>(c, SV_CONTACT_PENETRATION), B<uint32_t>(c, SV_CONTACT_GEOM_A),
                       B<uint32_t>(c, SV_CONTACT_GEOM_B),
                       clear_lambda ? B<float>(c, SV_CONTACT_LAMBDA_N) : (float *)nullptr,
                       sweep_bodies ? c->s.deg : (uint32_t *)nullptr, sweep_bodies, c->s.solve_ctl);
    c->sweep_setup_done = sweep_bodies > 0;
}

void launch_clear_lambda(sv_ctx *c) {
    uint32_t n = c->caps.max_contacts;
    if (!n) return;
    if (c->warm_factor > 0.0f) {
        launch_warm_lookup(c);
        return;
    }
    hipLaunchKernelGGL(k_clear_lambda, dim3(div_up(n, BLOCK)), dim3(BLOCK), 0, c->stream,
                       B<float>(c, SV_CONTACT_LAMBDA_N), n);
}

void launch_integrate(sv_ctx *c, uint32_t n, float dt) {
    if (!n) return;
    hipLaunchKernelGGL(k_integrate, dim3(div_up(n, BLOCK)), dim3(BLOCK), 0, c->stream, n, dt,
                       B<float>(c, SV_BODY_INV_MASS), B<float>(c, SV_BODY_POS), B<float>(c, SV_BODY_QUAT),
                       B<float>(c, SV_BODY_VEL), B<float>(c, SV_BODY_OMEGA));
}

void launch_update_verts(sv_ctx *c, uint32_t nv, uint32_t nb) {
    if (!nv || !nb) return;
    hipLaunchKernelGGL(k_update_verts, dim3(div_up((size_t)nb * VERT_ROW, BLOCK)), dim3(BLOCK), 0, c->stream, nv, nb,
                       B<uint32_t>(c, SV_BODY_VERT_START), B<float>(c, SV_BODY_POS), B<float>(c, SV_BODY_QUAT),
                       B<float>(c, SV_ORIGINAL_VERTICES), B<float>(c, SV_VERTICES));
}

void launch_clear_counters(sv_ctx *c) {
    hipLaunchKernelGGL(k_clear_counters, dim3(1), dim3(1), 0, c->stream, B<uint32_t>(c, SV_ATOMIC_COUNTERS));
}

void launch_rot(sv_ctx *c, uint32_t n) {
    if (!n) return;
    hipLaunchKernelGGL(k_rot, dim3(div_up(n, BLOCK)), dim3(BLOCK), 0, c->stream, n, B<float>(c, SV_BODY_QUAT),
                       c->s.rot);
}

void launch_pack_boxes(sv_ctx *c, uint32_t n) {
    if (!n) return;
    hipLaunchKernelGGL(k_pack_boxes, dim3(div_up(n, BLOCK)), dim3(BLOCK), 0, c->stream, n,
                       B<float>(c, SV_BODY_AABB_MIN), B<float>(c, SV_BODY_AABB_MAX), B<float>(c, SV_BODY_INV_MASS),
                       c->s.bpbox);
}

void launch_copy_calib(sv_ctx *c, const void *src, void *dst, size_t nbytes) {
    size_t n = nbytes / sizeof(float4);
    if (!n) return;
    uint32_t grid = div_up(n, BLOCK);
    if (grid > 16384) grid = 16384;
    hipLaunchKernelGGL(k_copy_calib, dim3(grid), dim3(BLOCK), 0, c->stream, (const float4 *)src, (float4 *)dst, n);
}

void launch_xcc_ids(sv_ctx *c, uint32_t *out, uint32_t n) {
    if (!n) return;
    hipLaunchKernelGGL(k_xcc_ids, dim3(n), dim3(64), 0, c->stream, out, n);
}

void launch_handoff(sv_ctx *c, uint32_t *flag, uint32_t rounds, uint32_t partner, uint64_t *out) {
    hipLaunchKernelGGL(k_handoff, dim3(partner + 1), dim3(64), 0, c->stream, flag, rounds, partner, out);
}

void launch_debug_atan2(sv_ctx *c, const float *y, const float *x, float *out, uint32_t n) {
    if (!n) return;
    hipLaunchKernelGGL(k_atan2, dim3(div_up(n, BLOCK)), dim3(BLOCK), 0, c->stream, y, x, out, n);
}

}  // namespace sv
