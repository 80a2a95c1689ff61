// sv_prims.hip - device-wide exclusive scan and stable LSD radix sort used
// by the broadphase, narrowphase compaction, the ordered solver and the
// LBVH.  Wave64 scans use cross-lane shuffles; the sort ranks keys with
// wave ballots (8 ballots give each lane the mask of lanes holding the same
// 8-bit digit), keeps per-wave digit counters in LDS and scatters stably.
#include "sv_internal.h"

namespace sv {

namespace {

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 4;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;  // 1024
static_assert(SCAN_TILE == SCAN_TILE_VALUES, "scan scratch sizing");

constexpr int SORT_THREADS = 256;
constexpr int SORT_WAVES = SORT_THREADS / 64;
constexpr int SORT_ROUNDS = 8;
constexpr int SORT_TILE = SORT_THREADS * SORT_ROUNDS;  // 2048
static_assert(SORT_TILE == SORT_TILE_KEYS, "sort scratch sizing");

__device__ __forceinline__ uint32_t read_n(const uint32_t *n_dev, uint32_t n_max) {
    if (!n_dev) return n_max;
    uint32_t n = *n_dev;
    return n < n_max ? n : n_max;
}

// Inclusive scan across the 64 lanes of a wave.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t t = __shfl_up(v, d, 64);
        if (lane >= d) v += t;
    }
    return v;
}

// Exclusive scan of one value per thread over a block of SCAN_THREADS;
// returns the exclusive prefix, writes the block total to *total.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *total) {
    __shared__ uint32_t wsum[SCAN_THREADS / 64];
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t inc = wave_incl_scan(v);
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    uint32_t base = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < SCAN_THREADS / 64; i++) {
        uint32_t s = wsum[i];
        if (i < w) base += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return base + inc - v;
}

__global__ void __launch_bounds__(SCAN_THREADS) k_scan_partials(const uint32_t *in, const uint32_t *n_dev,
                                                                uint32_t n_max, uint32_t *part) {
    uint32_t n = read_n(n_dev, n_max);
    uint32_t base = blockIdx.x * SCAN_TILE;
    uint32_t s = 0;
    if (base < n) {
#pragma unroll
        for (int k = 0; k < SCAN_ITEMS; k++) {
            uint32_t i = base + k * SCAN_THREADS + threadIdx.x;
            if (i < n) s += in[i];
        }
    }
    uint32_t tot;
    block_excl_scan(s, &tot);
    if (threadIdx.x == 0) part[blockIdx.x] = base < n ? tot : 0u;
}

// One block scans all partials in place (exclusive) and writes the total.
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_spine(uint32_t *part, uint32_t nparts, uint32_t *total_dev) {
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nparts; base += SCAN_THREADS) {
        uint32_t i = base + threadIdx.x;
        uint32_t v = i < nparts ? part[i] : 0u;
        uint32_t tot;
        uint32_t ex = block_excl_scan(v, &tot);
        if (i < nparts) part[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0 && total_dev) *total_dev = carry;
}

__global__ void __launch_bounds__(SCAN_THREADS) k_scan_apply(const uint32_t *in, uint32_t *out, const uint32_t *n_dev,
                                                             uint32_t n_max, const uint32_t *part) {
    uint32_t n = read_n(n_dev, n_max);
    uint32_t base = blockIdx.x * SCAN_TILE;
    if (base >= n) return;
    // thread t owns elements base + t*ITEMS .. +ITEMS (contiguous per thread)
    uint32_t v[SCAN_ITEMS];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        uint32_t i = base + threadIdx.x * SCAN_ITEMS + k;
        v[k] = i < n ? in[i] : 0u;
        s += v[k];
    }
    uint32_t tot;
    uint32_t ex = block_excl_scan(s, &tot) + part[blockIdx.x];
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        uint32_t i = base + threadIdx.x * SCAN_ITEMS + k;
        if (i < n) out[i] = ex;
        ex += v[k];
    }
}

__global__ void k_fill(uint32_t *p, uint32_t value, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = value;
}

// Per-tile 256-bin digit histogram, digit-major over the tiles that hold
// keys: hist[d * nb + blk], nb = tiles of the key count read on the device
// (<= nblocks, the launch), so the scan between the two kernels covers
// 256 * nb entries, whose number block 0 writes to *hist_n.
__global__ void __launch_bounds__(SORT_THREADS) k_sort_hist(const uint32_t *keys, const uint32_t *n_dev, uint32_t n_max,
                                                            int shift, uint32_t *hist, uint32_t *hist_n) {
    __shared__ uint32_t h[256];
    const uint32_t n = read_n(n_dev, n_max), nb = (n + SORT_TILE - 1) / SORT_TILE;
    if (blockIdx.x == 0 && threadIdx.x == 0) *hist_n = 256u * nb;
    if (blockIdx.x >= nb) return;
    h[threadIdx.x] = 0;
    __syncthreads();
    uint32_t base = blockIdx.x * SORT_TILE;
    for (int k = 0; k < SORT_ROUNDS; k++) {
        uint32_t i = base + k * SORT_THREADS + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * nb + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter.  Element order inside a tile: wave w owns
// [w*ROUNDS*64, (w+1)*ROUNDS*64), round r the next 64, lane = offset.
__global__ void __launch_bounds__(SORT_THREADS) k_sort_scatter(const uint32_t *kin, const uint32_t *vin, uint32_t *kout,
                                                               uint32_t *vout, const uint32_t *n_dev, uint32_t n_max,
                                                               int shift, const uint32_t *offs) {
    __shared__ uint32_t wcnt[SORT_WAVES][256];
    __shared__ uint32_t gbase[256];
    const uint32_t n = read_n(n_dev, n_max), nb = (n + SORT_TILE - 1) / SORT_TILE;
    if (blockIdx.x >= nb) return;
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int d = threadIdx.x; d < 256; d += SORT_THREADS) {
        gbase[d] = offs[d * nb + blockIdx.x];
#pragma unroll
        for (int q = 0; q < SORT_WAVES; q++) wcnt[q][d] = 0;
    }
    __syncthreads();
    uint32_t base = blockIdx.x * SORT_TILE + w * (SORT_ROUNDS * 64);
    uint32_t key[SORT_ROUNDS], val[SORT_ROUNDS], loc[SORT_ROUNDS];
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
    for (int r = 0; r < SORT_ROUNDS; r++) {
        uint32_t i = base + r * 64 + lane;
        bool valid = i < n;
        key[r] = valid ? kin[i] : 0u;
        val[r] = valid ? vin[i] : 0u;
        uint32_t d = (key[r] >> shift) & 255u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            bool bit = (d >> b) & 1u;
            uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        uint32_t rank = (uint32_t)__popcll(peers & lt);
        uint32_t cur = wcnt[w][d];
        loc[r] = cur + rank;
        __builtin_amdgcn_wave_barrier();
        if (valid && rank == 0) wcnt[w][d] = cur + (uint32_t)__popcll(peers);
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    for (int d = threadIdx.x; d < 256; d += SORT_THREADS) {
        uint32_t run = 0;
#pragma unroll
        for (int q = 0; q < SORT_WAVES; q++) {
            uint32_t t = wcnt[q][d];
            wcnt[q][d] = run;
            run += t;
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < SORT_ROUNDS; r++) {
        uint32_t i = base + r * 64 + lane;
        if (i < n) {
            uint32_t d = (key[r] >> shift) & 255u;
            uint32_t pos = gbase[d] + wcnt[w][d] + loc[r];
            kout[pos] = key[r];
            vout[pos] = val[r];
        }
    }
}

}  // namespace

// Grow a scratch buffer (only at creation or on first use of a larger size;
// never inside graph capture because prims_reserve sizes everything up front).
void ensure_u32(uint32_t *&p, uint32_t &cap, uint32_t need) {
    if (need <= cap) return;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    if (hipMalloc(&p, (size_t)need * sizeof(uint32_t)) == hipSuccess) cap = need;
}

// Scratch of at least `need` words: grown when the context owns it; a
// fixed (caller-provided) one that is too small is a launch error instead.
static bool reserve_fixed(sv_ctx *c, uint32_t *&p, uint32_t &cap, uint32_t need) {
    if (c->scratch_fixed) {
        if (need <= cap) return true;
        c->launch_err = hipErrorInvalidValue;
        return false;
    }
    ensure_u32(p, cap, need);
    return p != nullptr && need <= cap;
}

void prims_reserve(sv_ctx *c, uint32_t max_scan_n, uint32_t max_sort_n) {
    uint32_t sort_blocks = div_up(max_sort_n, SORT_TILE);
    uint32_t hist_n = 256u * sort_blocks;
    uint32_t scan_n = max_scan_n > hist_n ? max_scan_n : hist_n;
    ensure_u32(c->s.scan_part, c->s.scan_part_cap, div_up(scan_n, SCAN_TILE) + 1);
    ensure_u32(c->s.sort_hist, c->s.sort_hist_cap, 2u * hist_n + 1);
}

void scan_exclusive(sv_ctx *c, const uint32_t *in, uint32_t *out, const uint32_t *n_dev, uint32_t n_max,
                    uint32_t *total_dev) {
    if (n_max == 0) {
        if (total_dev) fill_u32(c, total_dev, 0u, 1);
        return;
    }
    uint32_t nparts = div_up(n_max, SCAN_TILE);
    if (!reserve_fixed(c, c->s.scan_part, c->s.scan_part_cap, nparts)) return;
    hipLaunchKernelGGL(k_scan_partials, dim3(nparts), dim3(SCAN_THREADS), 0, c->stream, in, n_dev, n_max,
                       c->s.scan_part);
    hipLaunchKernelGGL(k_scan_spine, dim3(1), dim3(SCAN_THREADS), 0, c->stream, c->s.scan_part, nparts, total_dev);
    hipLaunchKernelGGL(k_scan_apply, dim3(nparts), dim3(SCAN_THREADS), 0, c->stream, in, out, n_dev, n_max,
                       c->s.scan_part);
}

int radix_sort(sv_ctx *c, uint32_t *k0, uint32_t *v0, uint32_t *k1, uint32_t *v1, const uint32_t *n_dev,
               uint32_t n_max, int key_bits) {
    if (n_max == 0 || key_bits <= 0) return 0;
    uint32_t nblocks = div_up(n_max, SORT_TILE);
    uint32_t hist_n = 256u * nblocks;
    if (!reserve_fixed(c, c->s.sort_hist, c->s.sort_hist_cap, 2u * hist_n + 1)) return 0;
    uint32_t *hist = c->s.sort_hist, *offs = c->s.sort_hist + hist_n, *used = c->s.sort_hist + 2u * hist_n;
    int passes = (key_bits + 7) / 8;
    uint32_t *ka = k0, *va = v0, *kb = k1, *vb = v1;
    for (int p = 0; p < passes; p++) {
        int shift = p * 8;
        hipLaunchKernelGGL(k_sort_hist, dim3(nblocks), dim3(SORT_THREADS), 0, c->stream, ka, n_dev, n_max, shift,
                           hist, used);
        scan_exclusive(c, hist, offs, used, hist_n, nullptr);
        hipLaunchKernelGGL(k_sort_scatter, dim3(nblocks), dim3(SORT_THREADS), 0, c->stream, ka, va, kb, vb, n_dev,
                           n_max, shift, offs);
        uint32_t *t = ka; ka = kb; kb = t;
        t = va; va = vb; vb = t;
    }
    return passes & 1;
}

void fill_u32(sv_ctx *c, uint32_t *p, uint32_t value, size_t n) {
    if (n == 0) return;
    uint32_t blocks = div_up(n, 256);
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_fill, dim3(blocks), dim3(256), 0, c->stream, p, value, n);
}

}  // namespace sv
