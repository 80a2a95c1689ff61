// sv_physics.hip - gfx950 kernels for the per-substep stages of seeVR's
// physics.comp: gravity, per-body AABB, N^2 broadphase, MPR narrowphase,
// integration, render vertices.  The ordered PGS solver is in sv_pgs.hip.
#include "sv_internal.h"
#include "sv_math.h"

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
__global__ void k_gravity(uint32_t n, float dt, float gx, float gy, float gz, float ld, float ad,
                          const float *__restrict__ inv_mass, float *__restrict__ vel, float *__restrict__ omega) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (inv_mass[i] <= 0.0f) return;
    vel[i * 3 + 0] = vel[i * 3 + 0] * ld + gx * dt;
    vel[i * 3 + 1] = vel[i * 3 + 1] * ld + gy * dt;
    vel[i * 3 + 2] = vel[i * 3 + 2] * ld + gz * dt;
    omega[i * 3 + 0] *= ad;
    omega[i * 3 + 1] *= ad;
    omega[i * 3 + 2] *= ad;
}

// Per-body AABB, the semantics of physics.comp:431-448: min/max over
// p + R*v of the body's hull verts, seeded with +-1e10.  One wave per body,
// lanes stride the hull (coalesced 12-byte loads), then a shuffle tree.
// Also caches the body's rotation columns and a packed broadphase box
// (min.xyz, static flag), (max.xyz, 0) for the later stages.
__global__ void __launch_bounds__(BLOCK) k_aabb(uint32_t n, const float *__restrict__ pos,
                                                const float *__restrict__ quat, const uint32_t *__restrict__ hrange,
                                                const float *__restrict__ hull, const float *__restrict__ inv_mass,
                                                float *__restrict__ amin_out, float *__restrict__ amax_out,
                                                float *__restrict__ rot, float4 *__restrict__ bpbox) {
    uint32_t b = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    int lane = threadIdx.x & 63;
    if (b >= n) return;
    m3 R = quat_mat_at(quat, b);
    f3 p = ld3(pos, b);
    uint32_t vs = hrange[2 * b], vc = hrange[2 * b + 1];
    f3 mn = mk(1e10f, 1e10f, 1e10f), mx = mk(-1e10f, -1e10f, -1e10f);
    for (uint32_t v = lane; v < vc; v += 64) {
        f3 wv = p + mvec(R, ld3(hull, vs + v));
        mn = vmin3(mn, wv);
        mx = vmax3(mx, wv);
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        f3 omn = mk(__shfl_xor(mn.x, off, 64), __shfl_xor(mn.y, off, 64), __shfl_xor(mn.z, off, 64));
        f3 omx = mk(__shfl_xor(mx.x, off, 64), __shfl_xor(mx.y, off, 64), __shfl_xor(mx.z, off, 64));
        mn = vmin3(mn, omn);
        mx = vmax3(mx, omx);
    }
    if (lane == 0) {
        st3(amin_out, b, mn);
        st3(amax_out, b, mx);
        st_rot(rot, b, R);
        float st = inv_mass[b] <= 0.0f ? 1.0f : 0.0f;
        bpbox[2 * b] = make_float4(mn.x, mn.y, mn.z, st);
        bpbox[2 * b + 1] = make_float4(mx.x, mx.y, mx.z, 0.0f);
    }
}

// ---------------------------------------------------------------------------
// Broadphase (physics.comp:451-472): every (i, j > i) not both static whose
// boxes overlap (touching counts).
//
// Work is tiled in two dimensions so the triangle balances: a block takes
// 32 consecutive rows (8 per wave) against one column segment of
// bp_seg_len(n) bodies, streamed through LDS 256 boxes at a time, 64 j per
// lane-round.  Each test is branch-free: for finite IEEE floats
// max(lo_j - hi_i, lo_i - hi_j) > 0 exactly when hi_i < lo_j or hi_j < lo_i
// (a difference of two floats is positive iff the first is larger, with
// gradual underflow), so row i overlaps lane j iff the largest of the three
// axis separations is not positive.  Lanes past the segment or n carry an
// empty box.  A row's hits in one round form a 64-bit ballot numbered in j
// order.
//
// The test pass writes, per (row, segment), the hit count and the first
// BP_STAGE partners.  An exclusive scan over the counts in (row, segment)
// order gives every (row, segment) its slot range in lexicographic (i, j)
// order; k_bp_place copies staged partners there, and only tiles holding a
// (row, segment) with more than BP_STAGE hits are tested again to place the
// rest (EMIT form).
constexpr int BP_ROWS_PER_WAVE = 8;
constexpr int BP_TILE = BLOCK;  // 256 boxes per LDS tile

template <bool EMIT>
__global__ void __launch_bounds__(BLOCK) k_broadphase(uint32_t n, uint32_t nseg, uint32_t seg_len,
                                                      const float4 *__restrict__ box, uint32_t *__restrict__ cnt_out,
                                                      const uint32_t *__restrict__ off,
                                                      uint32_t *__restrict__ stage, uint32_t *__restrict__ tile_ovf,
                                                      uint32_t *__restrict__ pairs, uint32_t max_pairs) {
    __shared__ float4 s_lo[BP_TILE];
    __shared__ float4 s_hi[BP_TILE];
    const uint32_t seg = blockIdx.y;
    const uint32_t blk_row0 = blockIdx.x * BP_ROWS;
    const uint32_t j_lo = max(seg * seg_len, blk_row0 + 1);
    const uint32_t j_hi = min(seg * seg_len + seg_len, n);
    const uint32_t tile = blockIdx.x * nseg + seg;
    if (EMIT) {
        if (j_lo >= j_hi || tile_ovf[tile] == 0) return;
    } else if (j_lo >= j_hi) {  // no j > i in this segment: zero its counts
        uint32_t i = blk_row0 + threadIdx.x;
        if (threadIdx.x < BP_ROWS && i < n) cnt_out[(size_t)i * nseg + seg] = 0;
        return;
    }
    const int lane = threadIdx.x & 63;
    const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t row0 = blk_row0 + w * BP_ROWS_PER_WAVE;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const float4 empty_lo = make_float4(__builtin_inff(), __builtin_inff(), __builtin_inff(), 0.0f);
    const float4 empty_hi = make_float4(-__builtin_inff(), -__builtin_inff(), -__builtin_inff(), 0.0f);

    float4 rlo[BP_ROWS_PER_WAVE], rhi[BP_ROWS_PER_WAVE];
    uint32_t cnt[BP_ROWS_PER_WAVE];
    uint32_t base[BP_ROWS_PER_WAVE];
    bool any_static = false;
#pragma unroll
    for (int r = 0; r < BP_ROWS_PER_WAVE; r++) {
        uint32_t i = row0 + r;
        uint32_t ic = i < n ? i : 0u;
        rlo[r] = box[2 * ic];
        rhi[r] = box[2 * ic + 1];
        if (i >= n) { rlo[r] = empty_lo; rhi[r] = empty_hi; }
        any_static |= rlo[r].w != 0.0f;
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
#pragma unroll
            for (int r = 0; r < BP_ROWS_PER_WAVE; r++) {
                const uint32_t i = row0 + r;
                bool hit = !boxes_apart(rlo[r], rhi[r], lo, hi);
                if (diag) hit = hit && j > i;
                if (any_static) hit = hit && !(rlo[r].w != 0.0f && lo.w != 0.0f);
                const uint64_t m = __ballot(hit);
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
    if (!EMIT && lane == 0) {
        bool ovf = false;
#pragma unroll
        for (int r = 0; r < BP_ROWS_PER_WAVE; r++) {
            uint32_t i = row0 + r;
            if (i < n) cnt_out[(size_t)i * nseg + seg] = cnt[r];
            ovf |= cnt[r] > BP_STAGE;
        }
        if (ovf) tile_ovf[tile] = 1u;
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

// ---------------------------------------------------------------------------
// MPR narrowphase (physics.comp:107-399), one lane per pair.
struct Body {
    f3 p;
    m3 R;
    uint32_t vs, vc, lk;
};

struct Geo {
    const float *hull;
    const uint32_t *lookup;
};

SVD Body load_body(uint32_t b, const float *pos, const float *rot, const uint32_t *hrange,
                   const uint32_t *sl_off) {
    Body B;
    B.p = ld3(pos, b);
    B.R = ld_rot(rot, b);
    B.vs = hrange[2 * b];
    B.vc = hrange[2 * b + 1];
    B.lk = sl_off[b];
    return B;
}

// physics.comp:107-162: farthest hull vertex along world direction d.
SVD f3 support(const Body &B, f3 d, const Geo &g) {
    f3 ld = mtvec(B.R, d);
    float best = -1e10f;
    f3 bv = mk(0.0f, 0.0f, 0.0f);
    if (B.lk == 0) {
        for (uint32_t i = 0; i < B.vc; i++) {
            f3 v = ld3(g.hull, B.vs + i);
            float dd = dot3(v, ld);
            if (dd > best) {
                best = dd;
                bv = v;
            }
        }
        return B.p + mvec(B.R, bv);
    }
    float ang[3];
    ang[0] = sv_atan2f(ld.z, ld.y);
    ang[1] = sv_atan2f(ld.z, ld.x);
    ang[2] = sv_atan2f(ld.y, ld.x);
#pragma unroll
    for (uint32_t ax = 0; ax < 3; ax++) {
        float a = ang[ax];
        if (a < 0.0f) a += 6.2831853f;
        float bin_f = a / 6.2831853f * 64.0f;
        uint32_t b0 = ((uint32_t)bin_f) % 64u;
        uint32_t b1 = (b0 + 1u) % 64u;
        uint32_t i0 = g.lookup[B.lk + ax * 64u + b0];
        f3 v0 = ld3(g.hull, B.vs + i0);
        float d0 = dot3(v0, ld);
        if (d0 > best) {
            best = d0;
            bv = v0;
        }
        uint32_t i1 = g.lookup[B.lk + ax * 64u + b1];
        f3 v1 = ld3(g.hull, B.vs + i1);
        float d1 = dot3(v1, ld);
        if (d1 > best) {
            best = d1;
            bv = v1;
        }
    }
    return B.p + mvec(B.R, bv);
}

constexpr float CCD_EPS = 1e-6f;
constexpr float CCD_TOL = 1e-6f;
constexpr int CCD_ITERS = 50;

// Minkowski portal: Minkowski points and the support points on A and B.
struct Portal {
    f3 v[4], sa[4], sb[4];
};

struct Pair {
    Body A, Bb;
};

SVD void mink(const Pair &P, const Geo &g, f3 d, f3 &v, f3 &sa, f3 &sb) {
    sa = support(P.A, d, g);
    sb = support(P.Bb, -d, g);
    v = sa - sb;
}

SVD f3 portal_dir(f3 v1, f3 v2, f3 v3) { return norm3(cross3(v2 - v1, v3 - v1)); }

SVD bool reach_tol(f3 v1, f3 v2, f3 v3, f3 v4, f3 dir) {
    float dv1 = dot3(v1, dir), dv2 = dot3(v2, dir), dv3 = dot3(v3, dir), dv4 = dot3(v4, dir);
    float d1 = gmin(gmin(dv4 - dv1, dv4 - dv2), dv4 - dv3);
    return d1 < CCD_TOL + CCD_EPS * gmax(1.0f, d1);
}

// physics.comp:200-221
SVD void expand_portal(Portal &s, f3 v4, f3 s4a, f3 s4b) {
    f3 v4v0 = cross3(v4, s.v[0]);
    int slot;
    if (dot3(s.v[1], v4v0) > 0.0f) {
        slot = dot3(s.v[2], v4v0) > 0.0f ? 1 : 3;
    } else {
        slot = dot3(s.v[3], v4v0) > 0.0f ? 2 : 1;
    }
    if (slot == 1) { s.v[1] = v4; s.sa[1] = s4a; s.sb[1] = s4b; }
    else if (slot == 2) { s.v[2] = v4; s.sa[2] = s4a; s.sb[2] = s4b; }
    else { s.v[3] = v4; s.sa[3] = s4a; s.sb[3] = s4b; }
}

// physics.comp:224-247: barycentric contact point, averaged over A and B.
SVD f3 find_pos(const Portal &s) {
    f3 v0 = s.v[0], v1 = s.v[1], v2 = s.v[2], v3 = s.v[3];
    float b0 = dot3(cross3(v1, v2), v3);
    float b1 = -dot3(cross3(v0, v2), v3);
    float b2 = dot3(cross3(v0, v1), v3);
    float b3 = -dot3(cross3(v0, v1), v2);
    float sum = b0 + b1 + b2 + b3;
    if (__builtin_fabsf(sum) < CCD_EPS) {
        f3 dir = portal_dir(v1, v2, v3);
        b0 = 0.0f;
        b1 = dot3(cross3(v2, v3), dir);
        b2 = dot3(cross3(v3, v1), dir);
        b3 = dot3(cross3(v1, v2), dir);
        sum = b0 + b1 + b2 + b3;
    }
    f3 p1 = ((b0 * s.sa[0] + b1 * s.sa[1]) + b2 * s.sa[2]) + b3 * s.sa[3];
    f3 p2 = ((b0 * s.sb[0] + b1 * s.sb[1]) + b2 * s.sb[2]) + b3 * s.sb[3];
    return (0.5f / sum) * (p1 + p2);
}

// physics.comp:250-320: -1 separated, 0 portal found, 1 touching, 2 collinear
SVD int discover_portal(const Pair &P, const Geo &g, Portal &s) {
    s.sa[0] = P.A.p;
    s.sb[0] = P.Bb.p;
    s.v[0] = s.sa[0] - s.sb[0];
    if (__builtin_fabsf(s.v[0].x) < CCD_EPS && __builtin_fabsf(s.v[0].y) < CCD_EPS &&
        __builtin_fabsf(s.v[0].z) < CCD_EPS)
        s.v[0].x += 10.0f * CCD_EPS;
    f3 dir = norm3(-s.v[0]);
    mink(P, g, dir, s.v[1], s.sa[1], s.sb[1]);
    if (dot3(s.v[1], dir) < CCD_EPS) return -1;
    dir = cross3(s.v[0], s.v[1]);
    if (dot3(dir, dir) < CCD_EPS) {
        if (__builtin_fabsf(s.v[1].x) < CCD_EPS && __builtin_fabsf(s.v[1].y) < CCD_EPS &&
            __builtin_fabsf(s.v[1].z) < CCD_EPS)
            return 1;
        return 2;
    }
    dir = norm3(dir);
    mink(P, g, dir, s.v[2], s.sa[2], s.sb[2]);
    if (dot3(s.v[2], dir) < CCD_EPS) return -1;
    f3 va = s.v[1] - s.v[0];
    f3 vb = s.v[2] - s.v[0];
    dir = norm3(cross3(va, vb));
    if (dot3(dir, s.v[0]) > 0.0f) {
        f3 t = s.v[1]; s.v[1] = s.v[2]; s.v[2] = t;
        t = s.sa[1]; s.sa[1] = s.sa[2]; s.sa[2] = t;
        t = s.sb[1]; s.sb[1] = s.sb[2]; s.sb[2] = t;
        dir = -dir;
    }
    for (int trials = 0; trials < 15;) {
        f3 v3c, s3a, s3b;
        mink(P, g, dir, v3c, s3a, s3b);
        if (dot3(v3c, dir) < CCD_EPS) return -1;
        if (dot3(cross3(s.v[1], v3c), s.v[0]) < -CCD_EPS) {
            s.v[2] = v3c; s.sa[2] = s3a; s.sb[2] = s3b;
            dir = norm3(cross3(s.v[1] - s.v[0], s.v[2] - s.v[0]));
            trials++;
            continue;
        }
        if (dot3(cross3(v3c, s.v[2]), s.v[0]) < -CCD_EPS) {
            s.v[1] = v3c; s.sa[1] = s3a; s.sb[1] = s3b;
            dir = norm3(cross3(s.v[1] - s.v[0], s.v[2] - s.v[0]));
            trials++;
            continue;
        }
        s.v[3] = v3c; s.sa[3] = s3a; s.sb[3] = s3b;
        return 0;
    }
    return -1;
}

// physics.comp:323-343
SVD int refine_portal(const Pair &P, const Geo &g, Portal &s) {
    for (int it = 0; it < CCD_ITERS; it++) {
        f3 dir = portal_dir(s.v[1], s.v[2], s.v[3]);
        if (dot3(s.v[1], dir) > -CCD_EPS) return 0;
        f3 v4, s4a, s4b;
        mink(P, g, dir, v4, s4a, s4b);
        if (!(dot3(v4, dir) > -CCD_EPS) || reach_tol(s.v[1], s.v[2], s.v[3], v4, dir)) return -1;
        expand_portal(s, v4, s4a, s4b);
    }
    return 1;
}

// physics.comp:346-365 (always terminates with a result by the last iteration)
SVD void find_penetration(const Pair &P, const Geo &g, Portal &s, f3 &normal, float &depth) {
    for (int it = 0; it < CCD_ITERS; it++) {
        f3 dir = portal_dir(s.v[1], s.v[2], s.v[3]);
        f3 v4, s4a, s4b;
        mink(P, g, dir, v4, s4a, s4b);
        if (reach_tol(s.v[1], s.v[2], s.v[3], v4, dir) || it >= CCD_ITERS - 1) {
            depth = dot3(dir, s.v[1]);
            normal = dir;
            return;
        }
        expand_portal(s, v4, s4a, s4b);
    }
}

// physics.comp:368-399
SVD bool mpr(const Pair &P, const Geo &g, f3 &normal, float &depth, f3 &cp) {
    Portal s;
    int res = discover_portal(P, g, s);
    if (res == 1) {
        normal = norm3(s.v[0]);
        depth = 0.0f;
        cp = 0.5f * (s.sa[1] + s.sb[1]);
        return true;
    }
    if (res == 2) {
        depth = len3(s.v[1]);
        normal = norm3(s.v[1]);
        cp = 0.5f * (s.sa[1] + s.sb[1]);
        return true;
    }
    if (res != 0) return false;
    if (refine_portal(P, g, s) < 0) return false;
    find_penetration(P, g, s, normal, depth);
    cp = find_pos(s);
    return true;
}

// Step 3 (physics.comp:475-491), result per pair slot: hit flag and
// (pos, normal, depth); contacts are then compacted in pair order.
__global__ void __launch_bounds__(BLOCK) k_narrow(const uint32_t *__restrict__ misc, const uint32_t *__restrict__ pairs,
                                                  const float *__restrict__ pos, const float *__restrict__ rot,
                                                  const uint32_t *__restrict__ hrange, const uint32_t *__restrict__ sl_off,
                                                  const uint32_t *__restrict__ sl, const float *__restrict__ hull,
                                                  uint32_t *__restrict__ flag, float *__restrict__ res) {
    uint32_t np = misc[MISC_NP];
    Geo g{hull, sl};
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < np; p += gridDim.x * blockDim.x) {
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

// counters[1] = total contacts (written by the scan); clamp, flag overflow.
__global__ void k_post_contacts(uint32_t *counters, uint32_t *misc, uint32_t max_contacts) {
    uint32_t t = counters[1];
    misc[MISC_NC] = t < max_contacts ? t : max_contacts;
    if (t > max_contacts) counters[2] |= SV_STATUS_CONTACT_OVERFLOW;
}

// add_contact (physics.comp:403-412) at the canonical slot.
__global__ void __launch_bounds__(BLOCK) k_compact_contacts(const uint32_t *__restrict__ misc,
                                                            const uint32_t *__restrict__ pairs,
                                                            const uint32_t *__restrict__ flag,
                                                            const uint32_t *__restrict__ idx,
                                                            const float *__restrict__ res, uint32_t max_contacts,
                                                            float *__restrict__ cpos, float *__restrict__ cnrm,
                                                            float *__restrict__ cpen, uint32_t *__restrict__ ga,
                                                            uint32_t *__restrict__ gb) {
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
    if (inv_mass[i] <= 0.0f) return;
    pos[i * 3 + 0] += vel[i * 3 + 0] * dt;
    pos[i * 3 + 1] += vel[i * 3 + 1] * dt;
    pos[i * 3 + 2] += vel[i * 3 + 2] * dt;
    float ox = omega[i * 3], oy = omega[i * 3 + 1], oz = omega[i * 3 + 2];
    float qx = quat[i * 4], qy = quat[i * 4 + 1], qz = quat[i * 4 + 2], qw = quat[i * 4 + 3];
    float dx = ox * qw + oy * qz - oz * qy;
    float dy = -ox * qz + oy * qw + oz * qx;
    float dz = ox * qy - oy * qx + oz * qw;
    float dw = -ox * qx - oy * qy - oz * qz;
    qx += 0.5f * dx * dt;
    qy += 0.5f * dy * dt;
    qz += 0.5f * dz * dt;
    qw += 0.5f * dw * dt;
    float l = __builtin_sqrtf(qx * qx + qy * qy + qz * qz + qw * qw);
    if (l > 1e-8f) {
        qx /= l;
        qy /= l;
        qz /= l;
        qw /= l;
    }
    quat[i * 4] = qx;
    quat[i * 4 + 1] = qy;
    quat[i * 4 + 2] = qz;
    quat[i * 4 + 3] = qw;
}

// Step 7 (physics.comp:607-627): owner body by binary search on
// body_vert_start, world = p + R * local.
__global__ void __launch_bounds__(BLOCK) k_update_verts(uint32_t nv, uint32_t nb, const uint32_t *__restrict__ vstart,
                                                        const float *__restrict__ pos, const float *__restrict__ quat,
                                                        const float *__restrict__ orig, float *__restrict__ verts) {
    uint32_t vid = blockIdx.x * blockDim.x + threadIdx.x;
    if (vid >= nv) return;
    uint32_t lo = 0, hi = nb;
    while (lo + 1 < hi) {
        uint32_t mid = (lo + hi) / 2;
        if (vstart[mid] <= vid) lo = mid; else hi = mid;
    }
    m3 R = quat_mat_at(quat, lo);
    f3 wp = ld3(pos, lo) + mvec(R, ld3(orig, vid));
    st3(verts, vid, wp);
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

__global__ void k_atan2(const float *y, const float *x, float *out, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = sv_atan2f(y[i], x[i]);
}

}  // namespace

void launch_hull_ranges(sv_ctx *c, uint32_t n) {
    if (!n) return;
    hipLaunchKernelGGL(k_hull_ranges, dim3(div_up(n, BLOCK)), dim3(BLOCK), 0, c->stream, B<float>(c, SV_GEOM_DATA),
                       B<uint32_t>(c, SV_GEOM_HULL_RANGE), (int)c->params.hull_range_u32, c->s.hrange, n);
}

void launch_gravity(sv_ctx *c, uint32_t n, float dt, const float g[3]) {
    if (!n) return;
    hipLaunchKernelGGL(k_gravity, dim3(div_up(n, BLOCK)), dim3(BLOCK), 0, c->stream, n, dt, g[0], g[1], g[2],
                       c->params.linear_damping, c->params.angular_damping, B<float>(c, SV_BODY_INV_MASS),
                       B<float>(c, SV_BODY_VEL), B<float>(c, SV_BODY_OMEGA));
}

void launch_aabb(sv_ctx *c, uint32_t n) {
    if (!n) return;
    hipLaunchKernelGGL(k_aabb, dim3(div_up((uint64_t)n * 64, BLOCK)), dim3(BLOCK), 0, c->stream, n,
                       B<float>(c, SV_BODY_POS), B<float>(c, SV_BODY_QUAT), c->s.hrange, B<float>(c, SV_HULL_VERTS),
                       B<float>(c, SV_BODY_INV_MASS), B<float>(c, SV_BODY_AABB_MIN), B<float>(c, SV_BODY_AABB_MAX),
                       c->s.rot, c->s.bpbox);
}

void launch_broadphase(sv_ctx *c, uint32_t n) {
    uint32_t *counters = B<uint32_t>(c, SV_ATOMIC_COUNTERS);
    if (n == 0) {
        fill_u32(c, counters, 0u, 1);
        hipLaunchKernelGGL(k_post_pairs, dim3(1), dim3(1), 0, c->stream, counters, c->s.misc, c->caps.max_pairs);
        return;
    }
    const bool lbvh = c->broadphase_mode == 2 || (c->broadphase_mode == 0 && n >= BP_LBVH_MIN_BODIES);
    // same pair list either way; without LBVH scratch the all-pairs kernel runs
    if (lbvh && launch_broadphase_lbvh(c, n)) return;
    const uint32_t seg_len = bp_seg_len(n), nseg = bp_nseg(n);
    const uint32_t nrb = div_up(n, BP_ROWS);
    uint32_t *pairs = B<uint32_t>(c, SV_COLLISION_PAIRS);
    fill_u32(c, c->s.bp_ovf, 0u, (size_t)nrb * nseg);
    hipLaunchKernelGGL(k_broadphase<false>, dim3(nrb, nseg), dim3(BLOCK), 0, c->stream, n, nseg, seg_len, c->s.bpbox,
                       c->s.row_count, (const uint32_t *)nullptr, c->s.bp_stage, c->s.bp_ovf, (uint32_t *)nullptr,
                       c->caps.max_pairs);
    const uint32_t ncnt = n * nseg;
    scan_exclusive(c, c->s.row_count, c->s.row_off, nullptr, ncnt, counters);
    hipLaunchKernelGGL(k_post_pairs, dim3(1), dim3(1), 0, c->stream, counters, c->s.misc, c->caps.max_pairs);
    hipLaunchKernelGGL(k_bp_place, dim3(div_up((uint64_t)ncnt * BP_STAGE, BLOCK)), dim3(BLOCK), 0, c->stream, n,
                       nseg, c->s.row_count, c->s.row_off, c->s.bp_stage, pairs, c->caps.max_pairs);
    hipLaunchKernelGGL(k_broadphase<true>, dim3(nrb, nseg), dim3(BLOCK), 0, c->stream, n, nseg, seg_len, c->s.bpbox,
                       (uint32_t *)nullptr, c->s.row_off, c->s.bp_stage, c->s.bp_ovf, pairs, c->caps.max_pairs);
}

void launch_narrowphase(sv_ctx *c) {
    uint32_t *counters = B<uint32_t>(c, SV_ATOMIC_COUNTERS);
    // pairs to process: min(counters[0], max_pairs) (physics.comp:477)
    hipLaunchKernelGGL(k_post_pairs, dim3(1), dim3(1), 0, c->stream, counters, c->s.misc, c->caps.max_pairs);
    uint32_t mp = c->caps.max_pairs;
    uint32_t grid = div_up(mp, BLOCK);
    if (grid > 8192) grid = 8192;
    if (grid == 0) grid = 1;
    hipLaunchKernelGGL(k_narrow, dim3(grid), dim3(BLOCK), 0, c->stream, c->s.misc,
                       B<uint32_t>(c, SV_COLLISION_PAIRS), B<float>(c, SV_BODY_POS), c->s.rot, c->s.hrange,
                       B<uint32_t>(c, SV_SUPPORT_LOOKUP_OFFSET), B<uint32_t>(c, SV_SUPPORT_LOOKUP),
                       B<float>(c, SV_HULL_VERTS), c->s.np_flag, c->s.np_res);
    scan_exclusive(c, c->s.np_flag, c->s.np_idx, c->s.misc + MISC_NP, mp, counters + 1);
    hipLaunchKernelGGL(k_post_contacts, dim3(1), dim3(1), 0, c->stream, counters, c->s.misc, c->caps.max_contacts);
    hipLaunchKernelGGL(k_compact_contacts, dim3(grid), dim3(BLOCK), 0, c->stream, c->s.misc,
                       B<uint32_t>(c, SV_COLLISION_PAIRS), c->s.np_flag, c->s.np_idx, c->s.np_res,
                       c->caps.max_contacts, B<float>(c, SV_CONTACT_POS), B<float>(c, SV_CONTACT_NORMAL),
                       B<float>(c, SV_CONTACT_PENETRATION), B<uint32_t>(c, SV_CONTACT_GEOM_A),
                       B<uint32_t>(c, SV_CONTACT_GEOM_B));
}

void launch_clear_lambda(sv_ctx *c) {
    uint32_t n = c->caps.max_contacts;
    if (!n) return;
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
    hipLaunchKernelGGL(k_update_verts, dim3(div_up(nv, BLOCK)), dim3(BLOCK), 0, c->stream, nv, nb,
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

void launch_debug_atan2(sv_ctx *c, const float *y, const float *x, float *out, uint32_t n) {
    if (!n) return;
    hipLaunchKernelGGL(k_atan2, dim3(div_up(n, BLOCK)), dim3(BLOCK), 0, c->stream, y, x, out, n);
}

}  // namespace sv
