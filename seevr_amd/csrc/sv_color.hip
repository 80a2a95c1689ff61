// sv_color.hip - graph-coloured contact sweep (beside the ordered one) and
// the constraint residual that judges it.
//
// The coloured schedule runs the same Gauss-Seidel item (apply_item,
// physics.comp:525-580) in a different order: contacts are coloured so that
// no two contacts of a colour share a dynamic body, and every iteration visits
// the colours in turn, all contacts of a colour at once.  That is the serial
// sweep over the contacts sorted by (colour, index); it changes the result
// relative to the reference's order (so it is judged by the residual, not by
// parity) and it is deterministic.
//
// Colouring: rounds of an independent-set selection with hashed priorities
// (Luby).  In round r every uncoloured live contact offers the 64-bit key
// (r, hash(c), c) to each of its dynamic bodies with atomicMax; a contact
// whose key is the maximum at both bodies takes colour r.  Keys of earlier
// rounds are smaller, so the claim array is never cleared within a substep.
// Contacts still uncoloured after COLOR_ROUNDS rounds form a last class that
// one lane sweeps in index order.  A stable radix sort by colour gives the
// order; one cooperative launch then runs iterations x colours with a grid
// barrier between colours.
#include <type_traits>

#include "sv_internal.h"
#include "sv_math.h"
#include "sv_dev.h"

#include <cmath>
#include <cstring>

namespace sv {

namespace {

constexpr int BLOCK = 256;
constexpr uint32_t COLOR_ROUNDS = 32;                // colours 0..31; 32 = remainder
constexpr uint32_t COLOR_BINS = COLOR_ROUNDS + 1;
constexpr uint32_t UNCOLORED = 0xFFFFFFFFu;
constexpr uint64_t STALL_TICKS = 200000000ull;       // 2 s of the 100 MHz wall clock

template <typename T>
inline T *B(sv_ctx *c, int id) { return reinterpret_cast<T *>(c->buf[id]); }

SVD uint32_t fmix32(uint32_t h) {  // MurmurHash3 finaliser
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    h ^= h >> 16;
    return h;
}
SVD uint64_t color_key(uint32_t r, uint32_t c) {
    return ((uint64_t)r << 56) | ((uint64_t)(fmix32(c * 0x9E3779B9u) >> 8) << 24) | (uint64_t)(c & 0xFFFFFFu);
}

__global__ void __launch_bounds__(BLOCK) k_color_init(const uint32_t *__restrict__ misc,
                                                      const float *__restrict__ prep, uint32_t *__restrict__ color) {
    const uint32_t nc = misc[MISC_NC];
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += gridDim.x * blockDim.x) {
        const uint32_t flags = __float_as_uint(prep[(size_t)c * CPREP_WORDS + P_FLAGS]);
        color[c] = (flags & F_SKIP) ? COLOR_BINS : UNCOLORED;  // skipped contacts sort last and are not swept
    }
}

__global__ void __launch_bounds__(BLOCK) k_color_claim(const uint32_t *__restrict__ misc, uint32_t r,
                                                       const float *__restrict__ prep,
                                                       const uint32_t *__restrict__ color,
                                                       unsigned long long *claim) {
    const uint32_t nc = misc[MISC_NC];
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += gridDim.x * blockDim.x) {
        if (color[c] != UNCOLORED) continue;
        const float *P = prep + (size_t)c * CPREP_WORDS;
        const uint32_t flags = __float_as_uint(P[P_FLAGS]);
        const unsigned long long key = color_key(r, c);
        if (flags & F_DYN_A) atomicMax(claim + __float_as_uint(P[P_A]), key);
        if (flags & F_DYN_B) atomicMax(claim + __float_as_uint(P[P_B]), key);
    }
}

__global__ void __launch_bounds__(BLOCK) k_color_pick(const uint32_t *__restrict__ misc, uint32_t r,
                                                      const float *__restrict__ prep, uint32_t *__restrict__ color,
                                                      const unsigned long long *__restrict__ claim) {
    const uint32_t nc = misc[MISC_NC];
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += gridDim.x * blockDim.x) {
        if (color[c] != UNCOLORED) continue;
        const float *P = prep + (size_t)c * CPREP_WORDS;
        const uint32_t flags = __float_as_uint(P[P_FLAGS]);
        const unsigned long long key = color_key(r, c);
        const bool wa = !(flags & F_DYN_A) || claim[__float_as_uint(P[P_A])] == key;
        const bool wb = !(flags & F_DYN_B) || claim[__float_as_uint(P[P_B])] == key;
        if (wa && wb) color[c] = r;
    }
}

// sort keys: colour (remainder = COLOR_ROUNDS, skipped = COLOR_BINS), values: contact
__global__ void __launch_bounds__(BLOCK) k_color_keys(const uint32_t *__restrict__ misc, uint32_t *__restrict__ color,
                                                      uint32_t *__restrict__ key, uint32_t *__restrict__ val,
                                                      uint32_t *__restrict__ cnt) {
    const uint32_t nc = misc[MISC_NC];
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += gridDim.x * blockDim.x) {
        uint32_t k = color[c];
        if (k == UNCOLORED) k = COLOR_ROUNDS;
        color[c] = k;
        key[c] = k;
        val[c] = c;
        if (k < COLOR_BINS) atomicAdd(cnt + k, 1u);
    }
}

__global__ void k_color_offsets(const uint32_t *__restrict__ cnt, uint32_t *__restrict__ off) {
    uint32_t run = 0;
    for (uint32_t k = 0; k < COLOR_BINS; k++) {
        off[k] = run;
        run += cnt[k];
    }
    off[COLOR_BINS] = run;
}

// Grid-wide barrier for a cooperative launch: generation counter bar[1],
// arrivals bar[0]; device-scope release before arriving, acquire after.
// COH: the sweep's mutable words are accessed at the point of coherence
// (CohMem), so the barrier only waits for every lane's stores to complete
// and needs no L2 write-back or invalidation.
template <bool COH>
SVD bool grid_sync(uint32_t *bar, uint32_t *abort_flag, uint32_t *counters) {
    if (COH) __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    __shared__ uint32_t s_ok;
    if (threadIdx.x == 0) {
        if (!COH) release_agent();
        const uint32_t gen = __hip_atomic_load(bar + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t arrived = __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1u;
        uint32_t ok = 1u;
        if (arrived == gridDim.x) {
            __hip_atomic_store(bar, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (COH) __builtin_amdgcn_s_waitcnt(0);
            else release_agent();
            __hip_atomic_fetch_add(bar + 1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            const uint64_t t0 = wall_clock64();
            while (__hip_atomic_load(bar + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen) {
                if (__hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) {
                    ok = 0u;
                    break;
                }
                if (wall_clock64() - t0 > STALL_TICKS) {
                    __hip_atomic_store(abort_flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    atomicOr(counters + 2, (uint32_t)SV_STATUS_SOLVER_TIMEOUT);
                    ok = 0u;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        if (COH) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        else __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        s_ok = ok;
    }
    __syncthreads();
    return s_ok != 0u;
}

SVD void load_constants(Item &it, const float *__restrict__ P, uint32_t c, const float *__restrict__ rot,
                        const float *__restrict__ inv_i) {
    const uint32_t flags = __float_as_uint(P[P_FLAGS]);
    it.c = c;
    it.a = __float_as_uint(P[P_A]);
    it.b = __float_as_uint(P[P_B]);
    it.dyn_a = flags & F_DYN_A;
    it.dyn_b = flags & F_DYN_B;
    it.n = mk(P[P_N], P[P_N + 1], P[P_N + 2]);
    it.ra = mk(P[P_RA], P[P_RA + 1], P[P_RA + 2]);
    it.rb = mk(P[P_RB], P[P_RB + 1], P[P_RB + 2]);
    it.kn = P[P_KN];
    it.bias = P[P_BIAS];
    it.mu = P[P_MU];
    it.ima = P[P_IMA];
    it.imb = P[P_IMB];
    it.Ra = ld_rot(rot, it.a);
    it.Rb = ld_rot(rot, it.b);
    it.iia = ld3(inv_i, it.a);
    it.iib = ld3(inv_i, it.b);
}

// iterations x colours; colour k = order[off[k] .. off[k+1]) (ascending
// contact index); the remainder class is swept by one lane.
template <bool COH>
__global__ void __launch_bounds__(BLOCK) k_solve_colored(const uint32_t *__restrict__ order,
                                                         const uint32_t *__restrict__ off, uint32_t iters,
                                                         const float *__restrict__ prep, float *vel, float *omega,
                                                         const float *__restrict__ rot,
                                                         const float *__restrict__ inv_i, float *lambda,
                                                         uint32_t *bar, uint32_t *abort_flag, uint32_t *counters) {
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x, gstride = gridDim.x * blockDim.x;
    uint32_t last = 0;
    for (uint32_t k = 0; k < COLOR_ROUNDS; k++)
        if (off[k + 1] > off[k]) last = k + 1;
    for (uint32_t t = 0; t < iters; t++) {
        for (uint32_t k = 0; k < last; k++) {
            const uint32_t b0 = off[k], b1 = off[k + 1];
            for (uint32_t i = b0 + gtid; i < b1; i += gstride) {
                const uint32_t c = order[i];
                Item it;
                load_constants(it, prep + (size_t)c * CPREP_WORDS, c, rot, inv_i);
                apply_item<typename std::conditional<COH, CohMem, PlainMem>::type>(it, vel, omega, lambda);
            }
            if (!grid_sync<COH>(bar, abort_flag, counters)) return;
        }
        const uint32_t r0 = off[COLOR_ROUNDS], r1 = off[COLOR_BINS];
        if (r1 > r0) {
            if (gtid == 0)
                for (uint32_t i = r0; i < r1; i++) {
                    const uint32_t c = order[i];
                    Item it;
                    load_constants(it, prep + (size_t)c * CPREP_WORDS, c, rot, inv_i);
                    apply_item<typename std::conditional<COH, CohMem, PlainMem>::type>(it, vel, omega, lambda);
                }
            if (!grid_sync<COH>(bar, abort_flag, counters)) return;
        }
    }
}

// Complementarity residual of the normal constraints after a sweep:
// active contacts (lambda > 0) should reach the target normal velocity
// (v_n = bias), inactive ones should not approach faster than it.  Per
// contact r = lambda > 0 ? |v_n - bias| : max(bias - v_n, 0); out[0] += r^2,
// out[1] = max r (as uint bits of a non-negative float), out[2] += 1.
__global__ void __launch_bounds__(BLOCK) k_residual(const uint32_t *__restrict__ misc, const float *__restrict__ prep,
                                                    const float *__restrict__ vel, const float *__restrict__ omega,
                                                    const float *__restrict__ lambda, float *sum, uint32_t *mx,
                                                    uint32_t *count) {
    const uint32_t nc = misc[MISC_NC];
    float s = 0.0f, m = 0.0f;
    uint32_t n = 0;
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += gridDim.x * blockDim.x) {
        const float *P = prep + (size_t)c * CPREP_WORDS;
        if (__float_as_uint(P[P_FLAGS]) & F_SKIP) continue;
        const uint32_t a = __float_as_uint(P[P_A]), b = __float_as_uint(P[P_B]);
        const f3 nn = mk(P[P_N], P[P_N + 1], P[P_N + 2]);
        const f3 ra = mk(P[P_RA], P[P_RA + 1], P[P_RA + 2]), rb = mk(P[P_RB], P[P_RB + 1], P[P_RB + 2]);
        const f3 va = ld3(vel, a) + cross3(ld3(omega, a), ra), vb = ld3(vel, b) + cross3(ld3(omega, b), rb);
        const float vn = dot3(vb - va, nn), bias = P[P_BIAS];
        const float r = lambda[c] > 0.0f ? __builtin_fabsf(vn - bias) : gmax(bias - vn, 0.0f);
        s += r * r;
        m = gmax(m, r);
        n++;
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        s += __shfl_xor(s, o, 64);
        m = gmax(m, __shfl_xor(m, o, 64));
        n += __shfl_xor(n, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(sum, s);
        atomicMax(mx, __float_as_uint(m));
        atomicAdd(count, n);
    }
}

}  // namespace

namespace {

template <typename T>
bool lazy(sv_ctx *c, T *&p, size_t count) {
    if (p) return true;
    void *q = nullptr;
    if (hipMalloc(&q, (count ? count : 1) * sizeof(T)) != hipSuccess) return false;
    c->s.lazy.push_back(q);
    p = static_cast<T *>(q);
    return true;
}

}  // namespace

bool launch_solve_colored(sv_ctx *c, uint32_t n_bodies) {
    Scratch &s = c->s;
    const size_t N = c->caps.max_bodies, C = c->caps.max_contacts;
    if (C >= (1u << 24)) return false;  // contact index in 24 bits of the colour key
    if (!lazy(c, s.col_claim, N + 1) || !lazy(c, s.col_color, C + 1) || !lazy(c, s.col_cnt, COLOR_BINS + 2) ||
        !lazy(c, s.col_off, COLOR_BINS + 2) || !lazy(c, s.col_bar, 4))
        return false;
    static int coop = -1;
    if (coop < 0) {
        int v = 0;
        coop = hipDeviceGetAttribute(&v, hipDeviceAttributeCooperativeLaunch, c->device) == hipSuccess && v ? 1 : 0;
    }
    if (!coop) return false;
    // coherent accesses pay off when barriers dominate (1M-body scene, 6.3M
    // contact slots: 6.05 ms against 7.39 ms per sweep) but not on smaller
    // lists, where the plain form's L2 hits win (32k-body pile, 197k slots:
    // 4.87 ms against 6.46 ms)
    const bool coh = c->sweep_variant != SWEEP_FENCED && c->caps.max_contacts >= (1u << 20);
    const void *kern = coh ? reinterpret_cast<const void *>(k_solve_colored<true>)
                           : reinterpret_cast<const void *>(k_solve_colored<false>);
    int occ = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, BLOCK, 0) != hipSuccess ||
        occ < 1 || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess)
        return false;
    hipStream_t st = c->stream;
    uint32_t grid = div_up(c->caps.max_contacts, BLOCK);
    if (grid > 8192) grid = 8192;
    if (grid == 0) grid = 1;
    hipLaunchKernelGGL(k_color_init, dim3(grid), dim3(BLOCK), 0, st, s.misc, s.cprep, s.col_color);
    fill_u32(c, reinterpret_cast<uint32_t *>(s.col_claim), 0u, 2 * (size_t)n_bodies);
    for (uint32_t r = 0; r < COLOR_ROUNDS; r++) {
        hipLaunchKernelGGL(k_color_claim, dim3(grid), dim3(BLOCK), 0, st, s.misc, r, s.cprep, s.col_color,
                           s.col_claim);
        hipLaunchKernelGGL(k_color_pick, dim3(grid), dim3(BLOCK), 0, st, s.misc, r, s.cprep, s.col_color,
                           s.col_claim);
    }
    fill_u32(c, s.col_cnt, 0u, COLOR_BINS + 1);
    hipLaunchKernelGGL(k_color_keys, dim3(grid), dim3(BLOCK), 0, st, s.misc, s.col_color, s.inc_key[0], s.inc_val[0],
                       s.col_cnt);
    int w = radix_sort(c, s.inc_key[0], s.inc_val[0], s.inc_key[1], s.inc_val[1], s.misc + MISC_NC, c->caps.max_contacts,
                       6);
    hipLaunchKernelGGL(k_color_offsets, dim3(1), dim3(1), 0, st, s.col_cnt, s.col_off);
    fill_u32(c, s.col_bar, 0u, 4);
    fill_u32(c, s.solve_ctl, 0u, 4);
    const uint32_t *order = s.inc_val[w];
    const uint32_t *off = s.col_off;
    uint32_t iters = c->params.pgs_iterations;
    float *vel = B<float>(c, SV_BODY_VEL), *omega = B<float>(c, SV_BODY_OMEGA);
    const float *prep = s.cprep, *rot = s.rot, *inv_i = B<float>(c, SV_BODY_INV_INERTIA);
    float *lam = B<float>(c, SV_CONTACT_LAMBDA_N);
    uint32_t *bar = s.col_bar, *abort_flag = s.solve_ctl, *counters = B<uint32_t>(c, SV_ATOMIC_COUNTERS);
    void *args[] = {&order, &off, &iters, &prep, &vel, &omega, &rot, &inv_i, &lam, &bar, &abort_flag, &counters};
    // workgroups: sv_set_solver's count, else one per 8192 contact slots
    // within [32, 256]: every grid barrier costs an arrival per workgroup, so
    // small lists want few workgroups (32 was fastest on a 50k-contact pile)
    uint32_t maxg = (uint32_t)occ * (uint32_t)cus;
    uint32_t cgrid = c->solver_workgroups ? c->solver_workgroups : c->caps.max_contacts / 8192u;
    if (!c->solver_workgroups) cgrid = cgrid < 32u ? 32u : (cgrid > 256u ? 256u : cgrid);
    if (cgrid < 1) cgrid = 1;
    if (cgrid > maxg) cgrid = maxg;
    StageTimer sweep_timer(c, "phys/pgs_sweep");
    return hipLaunchCooperativeKernel(kern, dim3(cgrid), dim3(BLOCK), args, 0, st) == hipSuccess;
}

int color_counts(sv_ctx *c, uint32_t *out, uint32_t cap) {
    if (!c->s.col_cnt) return SV_E_INVALID;
    uint32_t n = cap < COLOR_BINS ? cap : COLOR_BINS;
    if (hipMemcpy(out, c->s.col_cnt, n * sizeof(uint32_t), hipMemcpyDeviceToHost) != hipSuccess) return SV_E_DEVICE;
    return (int)n;
}

int solver_residual(sv_ctx *c, double *rms, double *max_abs, uint32_t *count) {
    Scratch &s = c->s;
    if (!lazy(c, s.res_acc, 4)) return SV_E_NOMEM;
    fill_u32(c, s.res_acc, 0u, 4);
    uint32_t grid = div_up(c->caps.max_contacts, BLOCK);
    if (grid > 4096) grid = 4096;
    if (grid == 0) grid = 1;
    hipLaunchKernelGGL(k_residual, dim3(grid), dim3(BLOCK), 0, c->stream, s.misc, s.cprep, B<float>(c, SV_BODY_VEL),
                       B<float>(c, SV_BODY_OMEGA), B<float>(c, SV_CONTACT_LAMBDA_N),
                       reinterpret_cast<float *>(s.res_acc), s.res_acc + 1, s.res_acc + 2);
    uint32_t h[4];
    if (hipMemcpyAsync(h, s.res_acc, sizeof(h), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return SV_E_DEVICE;
    float sum, mx;
    memcpy(&sum, &h[0], 4);
    memcpy(&mx, &h[1], 4);
    if (count) *count = h[2];
    if (rms) *rms = h[2] ? std::sqrt((double)sum / h[2]) : 0.0;
    if (max_abs) *max_abs = mx;
    return SV_OK;
}

}  // namespace sv
