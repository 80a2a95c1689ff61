// sv_prims.hip - device-wide exclusive scan and stable LSD radix sort used
// by the broadphase, narrowphase compaction, the ordered solver and the
// LBVH.  Wave64 scans use cross-lane shuffles; the sort takes one launch for
// every pass's digit histograms and one launch per pass (decoupled
// look-back), ranks keys with wave ballots (8 ballots give each lane the
// mask of lanes holding the same 8-bit digit) and writes each tile out
// digit by digit from LDS.
#include "sv_internal.h"

namespace sv {

namespace {

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 4;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;  // 1024
static_assert(SCAN_TILE == SCAN_TILE_VALUES, "scan scratch sizing");

// Radix-sort tiles: 256 threads x 8 rounds (2048 keys) up to
// SORT_SMALL_MAX keys, where short look-back chains and more workgroups
// matter; 1024 threads x 8 rounds (8192 keys) above: half the tiles of
// 512-thread tiles for the same keys in flight, so the look-back walks half
// as far, with each wave's ranking chain as short (8.57M keys: 358 -> 344 us
// for the 4-pass sort; 512 threads x 16 rounds had been slower, 89 against
// 80 us per pass, from the longer ranking chain).  The choice goes by the
// launch size, which a count read on the device can be far below: the
// 32k-body pile's incidence sort (capacity 395k keys, ~110k sorted) took
// 14.0 us per pass on 8192-key tiles, 11.9 on 4096-key ones.
static_assert(256 * 8 == SORT_TILE_SMALL && 1024 * 8 == SORT_TILE_LARGE, "sort scratch sizing");
constexpr int HIST_THREADS = 256;
constexpr int HIST_ITEMS = 16;                         // keys per thread and iteration
constexpr int HIST_CHUNK = HIST_THREADS * HIST_ITEMS;  // 4096
constexpr uint32_t HIST_BLOCKS = 512;
constexpr int LB_WINDOW = 16;                          // predecessors read at once in the look-back

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

// Single-pass exclusive scan (decoupled look-back).  A workgroup takes the
// next tile index from the control word (tiles are taken in launch order, so
// every tile a workgroup waits for is held by a running workgroup), scans its
// SCAN_TILE values, publishes the tile total as an "aggregate" word, and one
// wave looks back over the predecessors' words (64 at a time) until an
// inclusive prefix, which it then publishes for its own tile.  Status words
// are 64-bit: flag (SCAN_FLAG_A / SCAN_FLAG_P) above a 32-bit value.
//
// Scratch `ctl`: word 0 = the half in use (bit 31) | the next tile index,
// then two halves of `tiles_cap` status words.  A launch uses one half and
// zeroes the other, the one the next launch uses; the workgroup with the last
// tile index resets the index and flips the half, so the state lives on the
// device (graph replays see it too).  With a count read on the device, only
// the first ceil(n / SCAN_TILE) workgroups take tile indices (one each) and
// clear; the others leave at once.
// A wait is bounded (2 s, a broken invariant): it then sets
// SV_STATUS_SCAN_TIMEOUT in *timeout_status and takes the prefix so far.
//
// Optional output: the total to *total_dev; with post_np, also k_post_pairs'
// work: *post_np = min(total, post_cap) and SV_STATUS_PAIR_OVERFLOW in
// *post_status when the total exceeds post_cap.
constexpr uint64_t SCAN_FLAG_A = 1ull << 32, SCAN_FLAG_P = 2ull << 32;
constexpr uint64_t SCAN_STALL_TICKS = 200000000ull;  // 2 s of the 100 MHz wall clock

__device__ __forceinline__ uint64_t ld_agent64(const uint64_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent64(uint64_t *p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void __launch_bounds__(SCAN_THREADS) k_scan(const uint32_t *__restrict__ in, uint32_t *__restrict__ out,
                                                       const uint32_t *n_dev, uint32_t n_max, uint32_t *ctl,
                                                       uint32_t tiles_cap, uint32_t *total_dev, uint32_t post_cap,
                                                       uint32_t *post_np, uint32_t *post_status,
                                                       uint32_t *timeout_status) {
    __shared__ uint32_t s_word, s_excl;
    const uint32_t tid = threadIdx.x, lane = tid & 63u;
    const uint32_t n = read_n(n_dev, n_max);
    const uint32_t nt = (n + SCAN_TILE - 1) / SCAN_TILE;  // tiles with values (<= g)
    // only the first nt workgroups work (a count read on the device may be
    // below the launch size): each takes one tile index, and they zero the
    // other half (with no tile the halves stay as they are: the one in use
    // is still clean)
    if (blockIdx.x >= nt) {
        if (nt == 0 && blockIdx.x == 0 && tid == 0) {
            if (total_dev) *total_dev = 0u;
            if (post_np) *post_np = 0u;
        }
        return;
    }
    if (tid == 0) s_word = atomicAdd(ctl, 1u);
    __syncthreads();
    const uint32_t half = s_word >> 31, tile = s_word & 0x7FFFFFFFu;
    if (tile >= nt || tile >= tiles_cap) {  // a broken invariant (the counter was not at zero): fail, in bounds
        if (tid == 0 && timeout_status) atomicOr(timeout_status, (uint32_t)SV_STATUS_SCAN_TIMEOUT);
        return;
    }
    uint64_t *stat = reinterpret_cast<uint64_t *>(ctl + 2) + (size_t)half * tiles_cap;
    uint64_t *other = reinterpret_cast<uint64_t *>(ctl + 2) + (size_t)(half ^ 1u) * tiles_cap;
    for (uint32_t i = tile * SCAN_THREADS + tid; i < tiles_cap; i += nt * SCAN_THREADS) other[i] = 0ull;
    if (tile == nt - 1 && tid == 0)  // every working workgroup has taken its index
        __hip_atomic_store(ctl, (half ^ 1u) << 31, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // thread t owns elements base + t*ITEMS .. +ITEMS (contiguous per thread)
    const uint32_t base = tile * SCAN_TILE;
    uint32_t v[SCAN_ITEMS];
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        const uint32_t i = base + tid * SCAN_ITEMS + k;
        v[k] = i < n ? in[i] : 0u;
        sum += v[k];
    }
    uint32_t agg;
    uint32_t ex = block_excl_scan(sum, &agg);
    if (tile == 0) {
        if (tid == 0) {
            st_agent64(stat, SCAN_FLAG_P | agg);
            s_excl = 0u;
        }
    } else if (tid < 64) {
        if (tid == 0) st_agent64(stat + tile, SCAN_FLAG_A | agg);
        // look back: lane k reads the word of tile t - k
        uint32_t excl = 0;
        int32_t t = (int32_t)tile - 1;
        uint64_t start = 0;
        for (;;) {
            const int32_t q = t - (int32_t)lane;
            const uint64_t w = q >= 0 ? ld_agent64(stat + q) : SCAN_FLAG_P;
            const uint64_t pm = __ballot((w & SCAN_FLAG_P) != 0ull);
            const uint64_t em = __ballot((w & (SCAN_FLAG_A | SCAN_FLAG_P)) == 0ull);
            // the lanes up to the nearest inclusive prefix (or all 64)
            const uint64_t span = pm ? (pm & (0ull - pm)) * 2ull - 1ull : ~0ull;
            if ((em & span) == 0ull) {
                uint32_t x = (span >> lane) & 1ull ? (uint32_t)w : 0u;
#pragma unroll
                for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, 64);
                excl += x;
                if (pm) break;
                t -= 64;
                continue;
            }
            // a tile in the span has not published: its workgroup is running
            if (start == 0) start = wall_clock64();
            if (wall_clock64() - start > SCAN_STALL_TICKS) {
                if (lane == 0 && timeout_status) atomicOr(timeout_status, (uint32_t)SV_STATUS_SCAN_TIMEOUT);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        if (tid == 0) {
            st_agent64(stat + tile, SCAN_FLAG_P | (uint64_t)(excl + agg));
            s_excl = excl;
        }
    }
    __syncthreads();
    ex += s_excl;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        const uint32_t i = base + tid * SCAN_ITEMS + k;
        if (i < n) out[i] = ex;
        ex += v[k];
    }
    if (tile == nt - 1 && tid == 0) {
        const uint32_t total = s_excl + agg;
        if (total_dev) *total_dev = total;
        if (post_np) {
            *post_np = total < post_cap ? total : post_cap;
            if (total > post_cap) atomicOr(post_status, (uint32_t)SV_STATUS_PAIR_OVERFLOW);
        }
    }
}

__global__ void k_fill(uint32_t *p, uint32_t value, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = value;
}

// ---------------------------------------------------------------------------
// Stable LSD radix sort, 8-bit digits, one launch per pass (decoupled
// look-back).  Element order inside a tile of THREADS x ROUNDS keys: wave w owns
// [w * ROUNDS * 64, (w + 1) * ROUNDS * 64), round r the next 64, lane =
// offset, so a wave's ranks follow element order.
//
// k_os_hist: every digit histogram of the sort in one read of the keys
// (a digit histogram does not depend on the order of the keys).  Blocks
// count in LDS (one addition per run of equal digits in a wave) and add
// their counts to per-pass accumulators with device-scope atomics; the last
// block to finish (a device-scope counter) turns them into each pass's
// exclusive digit offsets and leaves the accumulators, the pass tile
// counters and the done counter at zero for the next sort.  The same launch
// zeroes the look-back status words the passes use.
//
// k_os_pass: a workgroup takes the next tile index from the pass counter
// (tiles are taken in launch order, so every tile a workgroup waits for is
// held by a running workgroup), ranks its keys stably (8 wave ballots give
// each lane the lanes holding its digit), publishes its digit counts as
// "aggregate" words, and finds its base per digit by looking back over the
// predecessors' words (flag and count in one 32-bit word, device-coherent,
// LB_WINDOW of them read at once) until an inclusive prefix; then publishes
// its own inclusive prefix.  A wait is bounded (2 s, a broken invariant):
// it then sets the abort word and the SV_STATUS_SORT_TIMEOUT bit of the
// context's status word.  Keys, then values, are staged digit-sorted in LDS
// and written out in runs, so consecutive lanes write consecutive addresses.
constexpr uint32_t OS_FLAG_A = 1u << 30, OS_FLAG_P = 2u << 30, OS_VALUE = (1u << 30) - 1u;
constexpr uint64_t OS_STALL_TICKS = 200000000ull;  // 2 s of the 100 MHz wall clock

__device__ __forceinline__ uint32_t ld_agent(const uint32_t *p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ __forceinline__ void st_agent(uint32_t *p, uint32_t v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

// Exclusive scan of one value per digit, held by threads 0..255 of a
// block of nthreads >= 256 (every thread calls it).
template <int NTHREADS>
__device__ __forceinline__ uint32_t digit_excl_scan(uint32_t v) {
    __shared__ uint32_t wsum[4];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    uint32_t inc = 0;
    if (tid < 256) {
        inc = wave_incl_scan(v);
        if (lane == 63) wsum[w] = inc;
    }
    __syncthreads();
    uint32_t base = 0;
    if (tid < 256) {
#pragma unroll
        for (int i = 0; i < 4; i++)
            if (i < w) base += wsum[i];
    }
    __syncthreads();
    return base + inc - v;
}

// Stable ranks of a wave's ROUNDS x 64 keys (element base + r * 64 + lane)
// by their digit: the ballots of every round first (no dependence between
// rounds), then the wave's running digit counters `wc` (zeroed by the
// caller).  pos[r] = rank of the key among the wave's keys of its digit,
// all ones past the n keys; wc[d] ends as the wave's count of digit d.
template <int ROUNDS>
__device__ __forceinline__ void os_rank(const uint32_t (&key)[ROUNDS], uint32_t (&pos)[ROUNDS], uint32_t base,
                                        uint32_t n, int shift, uint32_t dmask, uint32_t *wc) {
    const int lane = threadIdx.x & 63;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
    for (int r = 0; r < ROUNDS; r++) {
        const bool valid = base + r * 64 + lane < n;
        const uint32_t d = (key[r] >> shift) & dmask;
        // lanes whose digit bits all equal this lane's: bal XNOR (bit ? ~0 : 0)
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; bit++) {
            const uint32_t b = (d >> bit) & 1u;
            const uint64_t bal = __ballot(b);
            const uint64_t m = (uint64_t)(int64_t)(-(int32_t)b);
            peers &= ~(bal ^ m);
        }
        // rank among the round's lanes with this digit | their number << 8;
        // all ones for a lane past the keys
        pos[r] = valid ? (uint32_t)__popcll(peers & lt) | ((uint32_t)__popcll(peers) << 8) : 0xFFFFFFFFu;
    }
#pragma unroll
    for (int r = 0; r < ROUNDS; r++) {
        const uint32_t d = (key[r] >> shift) & dmask;
        const bool valid = pos[r] != 0xFFFFFFFFu;
        const uint32_t rank = pos[r] & 255u, num = pos[r] >> 8;
        const uint32_t cur = wc[d];
        if (valid) pos[r] = cur + rank;
        __builtin_amdgcn_wave_barrier();
        if (valid && rank == 0) wc[d] = cur + num;
        __builtin_amdgcn_wave_barrier();
    }
}

// Look-back for one digit: the sum of the counts of tiles 0..tile-1 from
// their status words col[t * 256] (an inclusive prefix ends the walk).
__device__ __forceinline__ uint32_t os_lookback(const uint32_t *col, uint32_t tile, uint32_t *ctl, uint32_t *counters) {
    uint32_t excl = 0;
    int32_t t = (int32_t)tile - 1;  // nearest predecessor not yet added
    uint64_t start = 0;
    while (t >= 0) {
        uint32_t v[LB_WINDOW];
#pragma unroll
        for (int j = 0; j < LB_WINDOW; j++) v[j] = (t - j >= 0) ? ld_agent(col + (size_t)(t - j) * 256) : OS_FLAG_P;
        int j = 0;
        bool done = false;
#pragma unroll
        for (int q = 0; q < LB_WINDOW; q++) {
            if (j == q) {
                const uint32_t x = v[q];
                if (x & OS_FLAG_P) {
                    excl += x & OS_VALUE;
                    done = true;
                } else if (x & OS_FLAG_A) {
                    excl += x & OS_VALUE;
                    j = q + 1;
                }
            }
        }
        if (done) break;
        t -= j;
        if (j == LB_WINDOW) continue;
        // tile t has not published: the workgroup holding it is running
        if (start == 0) start = wall_clock64();
        if (wall_clock64() - start > OS_STALL_TICKS) {
            st_agent(ctl + SORT_CTL_ABORT, 1u);
            if (counters) atomicOr(counters + 2, (uint32_t)SV_STATUS_SORT_TIMEOUT);
            break;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    return excl;
}

template <int PASSES>
__global__ void __launch_bounds__(HIST_THREADS) k_os_hist(const uint32_t *__restrict__ keys, const uint32_t *n_dev,
                                                          uint32_t n_max, uint32_t top_mask, uint32_t tile_keys,
                                                          uint32_t *acc,
                                                          uint32_t *__restrict__ goff, uint32_t *ctl,
                                                          uint32_t *__restrict__ status) {
    constexpr int passes = PASSES;
    __shared__ uint32_t h[SORT_MAX_PASSES][256];
    __shared__ uint32_t s_last;
    const uint32_t n = read_n(n_dev, n_max), nb = (n + tile_keys - 1) / tile_keys;
    const int tid = threadIdx.x, lane = tid & 63;
    // blocks beyond the keys on the device (at least one block works) leave at once
    uint32_t active = (n + HIST_CHUNK - 1) / HIST_CHUNK;
    active = active < 1u ? 1u : active > gridDim.x ? gridDim.x : active;
    if (blockIdx.x >= active) return;
    // the passes' look-back words start at zero
    const size_t nstat = (size_t)passes * nb * 256;
    for (size_t i = (size_t)blockIdx.x * HIST_THREADS + tid; i < nstat; i += (size_t)active * HIST_THREADS)
        status[i] = 0u;
    for (int p = 0; p < (int)SORT_MAX_PASSES; p++) h[p][tid] = 0u;
    __syncthreads();
    for (uint32_t base = blockIdx.x * HIST_CHUNK; base < n; base += active * HIST_CHUNK) {
        uint32_t key[HIST_ITEMS];
#pragma unroll
        for (int k = 0; k < HIST_ITEMS; k++) {
            const uint32_t i = base + k * HIST_THREADS + tid;
            key[k] = i < n ? keys[i] : 0u;
        }
        // one addition per run of equal digits in consecutive lanes (the
        // lanes holding keys are a prefix of the wave)
#pragma unroll
        for (int k = 0; k < HIST_ITEMS; k++) {
            const bool valid = base + k * HIST_THREADS + tid < n;
            const uint32_t nvalid = (uint32_t)__popcll(__ballot(valid));
#pragma unroll
            for (int p = 0; p < passes; p++) {
                const uint32_t d = (key[k] >> (8 * p)) & (p == passes - 1 ? top_mask : 255u);
                const uint32_t prev = __shfl_up(d, 1, 64);
                const bool head = valid && (lane == 0 || d != prev);
                const uint64_t heads = __ballot(head);
                if (head) {
                    const uint64_t above = lane == 63 ? 0ull : heads & (~0ull << (lane + 1));
                    const uint32_t next = above ? (uint32_t)__builtin_ctzll(above) : nvalid;
                    atomicAdd(&h[p][d], next - (uint32_t)lane);
                }
            }
        }
    }
    __syncthreads();
    for (int p = 0; p < passes; p++) {
        const uint32_t c = h[p][tid];
        if (c) atomicAdd(acc + p * 256 + tid, c);
    }
    __builtin_amdgcn_s_waitcnt(0);  // this thread's additions have been performed
    __syncthreads();                // ... and every thread's
    if (tid == 0) s_last = atomicAdd(ctl + SORT_CTL_DONE, 1u) == active - 1;
    __syncthreads();
    if (!s_last) return;
    for (int p = 0; p < passes; p++) {
        const uint32_t tot = ld_agent(acc + p * 256 + tid);
        st_agent(acc + p * 256 + tid, 0u);
        uint32_t all;
        goff[p * 256 + tid] = block_excl_scan(tot, &all);
    }
    if (tid == 0) {
        ctl[SORT_CTL_DONE] = 0u;
        for (int p = 0; p < (int)SORT_MAX_PASSES; p++) ctl[SORT_CTL_TILE + p] = 0u;
    }
}

template <int THREADS, int ROUNDS>
__global__ void __launch_bounds__(THREADS) k_os_pass(const uint32_t *__restrict__ kin, const uint32_t *__restrict__ vin,
                                                     uint32_t *__restrict__ kout, uint32_t *__restrict__ vout,
                                                     const uint32_t *n_dev, uint32_t n_max, int pass, uint32_t dmask,
                                                     const uint32_t *__restrict__ goff, uint32_t *ctl,
                                                     uint32_t *status, uint32_t *counters) {
    constexpr int WAVES = THREADS / 64, TILE = THREADS * ROUNDS;
    __shared__ uint32_t s_buf[TILE];
    __shared__ uint32_t wcnt[WAVES][256];
    __shared__ uint32_t s_lo[256], s_gb[256];
    __shared__ uint32_t s_tile;
    const uint32_t n = read_n(n_dev, n_max), nb = (n + TILE - 1) / TILE;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int shift = 8 * pass;
    // exactly the workgroups below the device's tile count take tiles
    if (blockIdx.x >= nb) return;
    if (tid == 0) s_tile = atomicAdd(ctl + SORT_CTL_TILE + pass, 1u);
    for (int i = tid; i < WAVES * 256; i += THREADS) (&wcnt[0][0])[i] = 0u;
    __syncthreads();
    const uint32_t tile = s_tile;

    // stable ranks inside the tile: the ballots of every round first (no
    // dependence between rounds), then the wave's running digit counters
    const uint32_t base = tile * TILE + w * (ROUNDS * 64);
    uint32_t key[ROUNDS], val[ROUNDS], pos[ROUNDS];
#pragma unroll
    for (int r = 0; r < ROUNDS; r++) {
        const uint32_t i = base + r * 64 + lane;
        const bool valid = i < n;
        key[r] = valid ? kin[i] : 0u;
        val[r] = valid ? vin[i] : 0u;
    }
    os_rank<ROUNDS>(key, pos, base, n, shift, dmask, wcnt[w]);
    __syncthreads();
    // digit d = tid (< 256): the waves' offsets and the tile's count
    const uint32_t d = (uint32_t)tid;
    uint32_t cnt = 0;
    uint32_t *mine = status + ((size_t)pass * nb + tile) * 256 + d;
    if (tid < 256) {
#pragma unroll
        for (int q = 0; q < WAVES; q++) {
            const uint32_t t = wcnt[q][d];
            wcnt[q][d] = cnt;
            cnt += t;
        }
        st_agent(mine, (tile == 0 ? OS_FLAG_P : OS_FLAG_A) | cnt);
    }
    const uint32_t lo = digit_excl_scan<THREADS>(cnt);
    if (tid < 256) {
        s_lo[d] = lo;
        const uint32_t excl = os_lookback(status + (size_t)pass * nb * 256 + d, tile, ctl, counters);
        if (tile > 0) st_agent(mine, OS_FLAG_P | (excl + cnt));
        s_gb[d] = goff[pass * 256 + d] + excl;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < ROUNDS; r++) {
        if (pos[r] != 0xFFFFFFFFu) {
            const uint32_t dd = (key[r] >> shift) & dmask;
            pos[r] += s_lo[dd] + wcnt[w][dd];
        }
    }
    // keys staged digit-sorted in LDS (element order kept per digit) and
    // written out in runs; then the values through the same buffer
    const uint32_t left = n - tile * TILE;
    const uint32_t m = left < (uint32_t)TILE ? left : (uint32_t)TILE;
#pragma unroll
    for (int r = 0; r < ROUNDS; r++)
        if (pos[r] != 0xFFFFFFFFu) s_buf[pos[r]] = key[r];
    __syncthreads();
    uint32_t dst[ROUNDS];
#pragma unroll
    for (int k = 0; k < ROUNDS; k++) {
        const uint32_t i = k * THREADS + tid;
        dst[k] = 0xFFFFFFFFu;
        if (i < m) {
            const uint32_t kk = s_buf[i];
            const uint32_t dd = (kk >> shift) & dmask;
            dst[k] = s_gb[dd] + (i - s_lo[dd]);
            kout[dst[k]] = kk;
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < ROUNDS; r++)
        if (pos[r] != 0xFFFFFFFFu) s_buf[pos[r]] = val[r];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < ROUNDS; k++)
        if (dst[k] != 0xFFFFFFFFu) vout[dst[k]] = s_buf[k * THREADS + tid];
}


// Grid barrier of k_os_one: every thread's stores have completed (its own
// wait, then the workgroup barrier) before thread 0 adds the workgroup's
// arrival, and the workgroup goes on once `target` arrivals are counted
// (bounded like the look-back: 2 s, then the abort word and status bit).
__device__ __forceinline__ void os_grid_sync(uint32_t *ctl, uint32_t target, uint32_t *counters) {
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicAdd(ctl + SORT_CTL_BAR, 1u);
        uint64_t start = 0;
        while (ld_agent(ctl + SORT_CTL_BAR) < target) {
            if (start == 0) start = wall_clock64();
            if (wall_clock64() - start > OS_STALL_TICKS) {
                st_agent(ctl + SORT_CTL_ABORT, 1u);
                if (counters) atomicOr(counters + 2, (uint32_t)SV_STATUS_SORT_TIMEOUT);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    __syncthreads();
}

// k_os_one: the whole sort in one launch, for at most SORT_ONE_MAX keys,
// where the launches of the multi-launch form (a histogram launch, then one
// per pass, each a few microseconds from dispatch to its end) cost more than
// the work.  Workgroup t sorts tile t (2048 keys over eight waves of four
// rounds: a shorter ranking and write-out chain per wave than four waves of
// eight), and the at most 256 working workgroups are resident together (two
// fit on a CU): they count their tile's digits of every pass into the
// accumulators, meet at a grid barrier, scan the totals into each pass's
// digit offsets, and run the passes (ranking, look-back, digit-ordered
// write-out as in k_os_pass) with a grid barrier after each but the last.  What one pass writes for the
// next (keys, values) and the look-back words go to the point of coherence
// and are read from there (device-scope relaxed stores and loads), so no
// cache write-back or invalidation is needed across the XCDs' L2 caches.
// Each workgroup zeroes its own look-back words before the first barrier;
// the workgroup that leaves last zeroes the accumulators and the barrier
// counters for the next sort.  Output in (k1, v1) for an odd pass count,
// else (k0, v0), as with the multi-launch form.
constexpr int ONE_THREADS = 512, ONE_ROUNDS = 4;  // 2048-key tiles over eight waves
template <int PASSES>
__global__ void __launch_bounds__(ONE_THREADS) k_os_one(uint32_t *k0, uint32_t *v0, uint32_t *k1, uint32_t *v1,
                                                        const uint32_t *n_dev, uint32_t n_max, uint32_t top_mask,
                                                        uint32_t *acc, uint32_t *ctl, uint32_t *status,
                                                        uint32_t *counters) {
    constexpr int THREADS = ONE_THREADS, ROUNDS = ONE_ROUNDS, WAVES = THREADS / 64, TILE = THREADS * ROUNDS;
    static_assert(TILE == SORT_TILE_SMALL, "tile of the look-back scratch");
    __shared__ uint32_t s_buf[TILE];
    __shared__ uint32_t wcnt[WAVES][256];
    __shared__ uint32_t s_goff[PASSES][256];
    __shared__ uint32_t s_lo[256], s_gb[256];
    __shared__ uint32_t s_last;
    const uint32_t n = read_n(n_dev, n_max), nb = (n + TILE - 1) / TILE;
    const uint32_t tile = blockIdx.x;
    if (tile >= nb) return;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const bool digit = tid < 256;  // threads 0..255 each hold digit tid
    const uint32_t base = tile * TILE + w * (ROUNDS * 64);
    if (digit) {
        for (int p = 0; p < PASSES; p++) {
            st_agent(status + ((size_t)p * nb + tile) * 256 + tid, 0u);
            s_goff[p][tid] = 0u;
        }
    }
    uint32_t key[ROUNDS], val[ROUNDS], pos[ROUNDS];
#pragma unroll
    for (int r = 0; r < ROUNDS; r++) {
        const uint32_t i = base + r * 64 + lane;
        const bool valid = i < n;
        key[r] = valid ? k0[i] : 0u;
        val[r] = valid ? v0[i] : 0u;
    }
    __syncthreads();
    // the tile's digit counts of every pass (one addition per run of equal
    // digits in a round; the lanes holding keys are a prefix of the wave)
#pragma unroll
    for (int r = 0; r < ROUNDS; r++) {
        const bool valid = base + r * 64 + lane < n;
        const uint32_t nvalid = (uint32_t)__popcll(__ballot(valid));
#pragma unroll
        for (int p = 0; p < PASSES; p++) {
            const uint32_t d = (key[r] >> (8 * p)) & (p == PASSES - 1 ? top_mask : 255u);
            const uint32_t prev = __shfl_up(d, 1, 64);
            const bool head = valid && (lane == 0 || d != prev);
            const uint64_t heads = __ballot(head);
            if (head) {
                const uint64_t above = lane == 63 ? 0ull : heads & (~0ull << (lane + 1));
                const uint32_t next = above ? (uint32_t)__builtin_ctzll(above) : nvalid;
                atomicAdd(&s_goff[p][d], next - (uint32_t)lane);
            }
        }
    }
    __syncthreads();
    if (digit) {
#pragma unroll
        for (int p = 0; p < PASSES; p++) {
            const uint32_t c = s_goff[p][tid];
            if (c) atomicAdd(acc + p * 256 + tid, c);
        }
    }
    os_grid_sync(ctl, nb, counters);
    uint32_t tot[PASSES];
#pragma unroll
    for (int p = 0; p < PASSES; p++) tot[p] = digit ? ld_agent(acc + p * 256 + tid) : 0u;  // one round trip
#pragma unroll
    for (int p = 0; p < PASSES; p++) {
        const uint32_t ex = digit_excl_scan<THREADS>(tot[p]);
        if (digit) s_goff[p][tid] = ex;
    }

#pragma unroll 1
    for (int p = 0; p < PASSES; p++) {
        const bool last = p == PASSES - 1;
        const uint32_t *kin = (p & 1) ? k1 : k0, *vin = (p & 1) ? v1 : v0;
        uint32_t *kout = (p & 1) ? k0 : k1, *vout = (p & 1) ? v0 : v1;
        const int shift = 8 * p;
        const uint32_t dmask = last ? top_mask : 255u;
        if (p > 0) {
#pragma unroll
            for (int r = 0; r < ROUNDS; r++) {
                const uint32_t i = base + r * 64 + lane;
                const bool valid = i < n;
                key[r] = valid ? ld_agent(kin + i) : 0u;
                val[r] = valid ? ld_agent(vin + i) : 0u;
            }
        }
        for (int i = tid; i < WAVES * 256; i += THREADS) (&wcnt[0][0])[i] = 0u;
        __syncthreads();
        os_rank<ROUNDS>(key, pos, base, n, shift, dmask, wcnt[w]);
        __syncthreads();
        // digit d = tid: the waves' offsets, the tile's count, its base
        const uint32_t d = (uint32_t)tid;
        uint32_t cnt = 0;
        uint32_t *mine = status + ((size_t)p * nb + tile) * 256 + d;
        if (digit) {
#pragma unroll
            for (int q = 0; q < WAVES; q++) {
                const uint32_t t = wcnt[q][d];
                wcnt[q][d] = cnt;
                cnt += t;
            }
            st_agent(mine, (tile == 0 ? OS_FLAG_P : OS_FLAG_A) | cnt);
        }
        const uint32_t lo = digit_excl_scan<THREADS>(cnt);
        if (digit) {
            s_lo[d] = lo;
            const uint32_t excl = os_lookback(status + (size_t)p * nb * 256 + d, tile, ctl, counters);
            if (tile > 0) st_agent(mine, OS_FLAG_P | (excl + cnt));
            s_gb[d] = s_goff[p][d] + excl;
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < ROUNDS; r++) {
            if (pos[r] != 0xFFFFFFFFu) {
                const uint32_t dd = (key[r] >> shift) & dmask;
                pos[r] += s_lo[dd] + wcnt[w][dd];
            }
        }
        // keys, then values, staged digit-sorted in LDS and written out in
        // runs; device-scope stores for the next pass, plain ones at the end
        const uint32_t left = n - tile * TILE;
        const uint32_t m = left < (uint32_t)TILE ? left : (uint32_t)TILE;
#pragma unroll
        for (int r = 0; r < ROUNDS; r++)
            if (pos[r] != 0xFFFFFFFFu) s_buf[pos[r]] = key[r];
        __syncthreads();
        uint32_t dst[ROUNDS];
#pragma unroll
        for (int k = 0; k < ROUNDS; k++) {
            const uint32_t i = k * THREADS + tid;
            dst[k] = 0xFFFFFFFFu;
            if (i < m) {
                const uint32_t kk = s_buf[i];
                const uint32_t dd = (kk >> shift) & dmask;
                dst[k] = s_gb[dd] + (i - s_lo[dd]);
                if (last) kout[dst[k]] = kk;
                else st_agent(kout + dst[k], kk);
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < ROUNDS; r++)
            if (pos[r] != 0xFFFFFFFFu) s_buf[pos[r]] = val[r];
        __syncthreads();
#pragma unroll
        for (int k = 0; k < ROUNDS; k++) {
            if (dst[k] != 0xFFFFFFFFu) {
                if (last) vout[dst[k]] = s_buf[k * THREADS + tid];
                else st_agent(vout + dst[k], s_buf[k * THREADS + tid]);
            }
        }
        if (!last) os_grid_sync(ctl, (uint32_t)(p + 2) * nb, counters);
    }
    // every workgroup has passed its last barrier and read the accumulators
    // once the exit count reaches nb: the last one leaves them at zero
    if (tid == 0) s_last = atomicAdd(ctl + SORT_CTL_EXIT, 1u) == nb - 1;
    __syncthreads();
    if (s_last) {
        if (digit) {
#pragma unroll
            for (int p = 0; p < PASSES; p++) st_agent(acc + p * 256 + tid, 0u);
        }
        if (tid == 0) {
            st_agent(ctl + SORT_CTL_BAR, 0u);
            st_agent(ctl + SORT_CTL_EXIT, 0u);
        }
    }
}

}  // namespace

// Grow a scratch buffer (only at creation or on first use of a larger size;
// never inside graph capture because prims_reserve sizes everything up front).
void ensure_u32(uint32_t *&p, uint32_t &cap, uint32_t need) {
    if (need <= cap) return;
    if (p) {
        (void)hipDeviceSynchronize();  // launches queued on a context stream may still use it
        (void)hipFree(p);
    }
    p = nullptr;
    cap = 0;
    // zeroed before any launch uses it: the sort's and the scan's control
    // words must start at zero
    if (hipMalloc(&p, (size_t)need * sizeof(uint32_t)) == hipSuccess) {
        if (zero_device(p, (size_t)need * sizeof(uint32_t)) == hipSuccess) {
            cap = need;
        } else {
            (void)hipFree(p);
            p = nullptr;
        }
    }
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
    ensure_u32(c->s.scan_part, c->s.scan_part_cap, (uint32_t)scan_part_words(max_scan_n));
    ensure_u32(c->s.sort_hist, c->s.sort_hist_cap, (uint32_t)sort_hist_words(max_sort_n));
}

void scan_exclusive(sv_ctx *c, const uint32_t *in, uint32_t *out, const uint32_t *n_dev, uint32_t n_max,
                    uint32_t *total_dev, uint32_t post_cap, uint32_t *post_np, uint32_t *post_status) {
    if (n_max == 0) {
        if (total_dev) fill_u32(c, total_dev, 0u, 1);
        if (post_np) fill_u32(c, post_np, 0u, 1);
        return;
    }
    const uint32_t tiles = div_up(n_max, SCAN_TILE);
    if (!reserve_fixed(c, c->s.scan_part, c->s.scan_part_cap, (uint32_t)scan_part_words(n_max))) return;
    const uint32_t tiles_cap = scan_tiles_cap(c->s.scan_part_cap);
    uint32_t *counters = reinterpret_cast<uint32_t *>(c->buf[SV_ATOMIC_COUNTERS]);
    hipLaunchKernelGGL(k_scan, dim3(tiles), dim3(SCAN_THREADS), 0, c->stream, in, out, n_dev, n_max, c->s.scan_part,
                       tiles_cap, total_dev, post_cap, post_np, post_status, counters ? counters + 2 : nullptr);
}

int radix_sort(sv_ctx *c, uint32_t *k0, uint32_t *v0, uint32_t *k1, uint32_t *v1, const uint32_t *n_dev,
               uint32_t n_max, int key_bits) {
    if (n_max == 0 || key_bits <= 0) return 0;
    if (key_bits > 32 || n_max > SORT_MAX_KEYS) {
        c->launch_err = hipErrorInvalidValue;
        return 0;
    }
    const bool small = n_max <= SORT_SMALL_MAX;
    const uint32_t tile_keys = small ? SORT_TILE_SMALL : SORT_TILE_LARGE;
    const uint32_t nblocks = div_up(n_max, tile_keys);
    if (!reserve_fixed(c, c->s.sort_hist, c->s.sort_hist_cap, (uint32_t)sort_hist_words(n_max))) return 0;
    uint32_t *acc = c->s.sort_hist + SORT_ACC_OFF, *goff = c->s.sort_hist + SORT_GOFF_OFF,
             *ctl = c->s.sort_hist + SORT_CTL_OFF, *status = c->s.sort_hist + SORT_STATUS_OFF;
    uint32_t *counters = reinterpret_cast<uint32_t *>(c->buf[SV_ATOMIC_COUNTERS]);
    const int passes = (key_bits + 7) / 8;
    // the last pass's digit keeps only the bits below key_bits
    const uint32_t top_bits = (uint32_t)key_bits - 8u * (uint32_t)(passes - 1);
    const uint32_t top_mask = top_bits >= 8 ? 255u : (1u << top_bits) - 1u;
    if (n_max <= SORT_ONE_MAX) {
        auto one = passes == 1 ? k_os_one<1> : passes == 2 ? k_os_one<2> : passes == 3 ? k_os_one<3> : k_os_one<4>;
        hipLaunchKernelGGL(one, dim3(nblocks), dim3(ONE_THREADS), 0, c->stream, k0, v0, k1, v1, n_dev, n_max, top_mask, acc,
                           ctl, status, counters);
        return passes & 1;
    }
    const uint32_t chunks = div_up(n_max, HIST_CHUNK);
    const uint32_t hb = chunks < HIST_BLOCKS ? chunks : HIST_BLOCKS;
    auto hist = passes == 1 ? k_os_hist<1> : passes == 2 ? k_os_hist<2> : passes == 3 ? k_os_hist<3> : k_os_hist<4>;
    hipLaunchKernelGGL(hist, dim3(hb), dim3(HIST_THREADS), 0, c->stream, k0, n_dev, n_max, top_mask, tile_keys, acc,
                       goff, ctl, status);
    uint32_t *ka = k0, *va = v0, *kb = k1, *vb = v1;
    for (int p = 0; p < passes; p++) {
        const uint32_t dmask = p == passes - 1 ? top_mask : 255u;
        if (small)
            hipLaunchKernelGGL((k_os_pass<256, 8>), dim3(nblocks), dim3(256), 0, c->stream, ka, va, kb, vb, n_dev,
                               n_max, p, dmask, goff, ctl, status, counters);
        else
            hipLaunchKernelGGL((k_os_pass<1024, 8>), dim3(nblocks), dim3(1024), 0, c->stream, ka, va, kb, vb, n_dev,
                               n_max, p, dmask, goff, ctl, status, counters);
        uint32_t *t = ka; ka = kb; kb = t;
        t = va; va = vb; vb = t;
    }
    return passes & 1;
}

bool sort_take_abort(sv_ctx *c) {
    if (!c->s.sort_hist || c->s.sort_hist_cap < SORT_STATUS_OFF) return false;
    uint32_t *w = c->s.sort_hist + SORT_CTL_OFF + SORT_CTL_ABORT;
    uint32_t v = 0;
    if (hipMemcpy(&v, w, sizeof(v), hipMemcpyDeviceToHost) != hipSuccess || v == 0) return false;
    const uint32_t zero = 0;
    (void)hipMemcpy(w, &zero, sizeof(zero), hipMemcpyHostToDevice);
    return true;
}

void fill_u32(sv_ctx *c, uint32_t *p, uint32_t value, size_t n) {
    if (n == 0) return;
    uint32_t blocks = div_up(n, 256);
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_fill, dim3(blocks), dim3(256), 0, c->stream, p, value, n);
}

}  // namespace sv
