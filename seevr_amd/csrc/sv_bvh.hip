// sv_bvh.hip - LBVH build of bvh.comp on gfx950: centroids, centroid
// bounds, 30-bit Morton codes, stable LSD radix sort (8-bit digits, wave
// ballot ranking; any digit width gives the same stable order as the
// reference's eight 4-bit passes), Karras 2012 hierarchy, leaf boxes and the
// bottom-up refit (the reference's arrival climb, one launch).
#include "sv_internal.h"
#include "sv_math.h"

namespace sv {

namespace {

constexpr int BLOCK = 256;
constexpr int BOUNDS_ITEMS = 8;
static_assert(BLOCK * BOUNDS_ITEMS == BOUNDS_TILE, "bounds partials sizing");

template <typename T>
inline T *B(sv_ctx *c, int id) { return reinterpret_cast<T *>(c->buf[id]); }

// bvh.comp:69-79 (triangles) / body AABB centres (extension).
__global__ void k_centroids(uint32_t n, const float *__restrict__ verts, const uint32_t *__restrict__ idx,
                            float *__restrict__ cent, uint32_t *__restrict__ sort_idx, uint32_t *__restrict__ prim) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t i0 = idx[i * 3], i1 = idx[i * 3 + 1], i2 = idx[i * 3 + 2];
#pragma unroll
    for (int k = 0; k < 3; k++) cent[i * 3 + k] = (verts[i0 * 3 + k] + verts[i1 * 3 + k] + verts[i2 * 3 + k]) / 3.0f;
    sort_idx[i] = i;
    prim[i] = i;
}

__global__ void k_centroids_bodies(uint32_t n, const float *__restrict__ amin, const float *__restrict__ amax,
                                   float *__restrict__ cent, uint32_t *__restrict__ sort_idx,
                                   uint32_t *__restrict__ prim) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
#pragma unroll
    for (int k = 0; k < 3; k++) cent[i * 3 + k] = (amin[i * 3 + k] + amax[i * 3 + k]) * 0.5f;
    sort_idx[i] = i;
    prim[i] = i;
}

// bvh.comp:88-165: min/max of centroids, idle lanes padded with +-1e30.
__device__ void block_minmax(f3 &mn, f3 &mx) {
    __shared__ float s[6][BLOCK / 64];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        mn = vmin3(mn, mk(__shfl_xor(mn.x, off, 64), __shfl_xor(mn.y, off, 64), __shfl_xor(mn.z, off, 64)));
        mx = vmax3(mx, mk(__shfl_xor(mx.x, off, 64), __shfl_xor(mx.y, off, 64), __shfl_xor(mx.z, off, 64)));
    }
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
        s[0][w] = mn.x; s[1][w] = mn.y; s[2][w] = mn.z;
        s[3][w] = mx.x; s[4][w] = mx.y; s[5][w] = mx.z;
    }
    __syncthreads();
    mn = mk(s[0][0], s[1][0], s[2][0]);
    mx = mk(s[3][0], s[4][0], s[5][0]);
#pragma unroll
    for (int q = 1; q < BLOCK / 64; q++) {
        mn = vmin3(mn, mk(s[0][q], s[1][q], s[2][q]));
        mx = vmax3(mx, mk(s[3][q], s[4][q], s[5][q]));
    }
}

__global__ void __launch_bounds__(BLOCK) k_bounds_partial(uint32_t n, const float *__restrict__ cent,
                                                          float *__restrict__ part) {
    f3 mn = mk(1e30f, 1e30f, 1e30f), mx = mk(-1e30f, -1e30f, -1e30f);
    uint32_t base = blockIdx.x * (BLOCK * BOUNDS_ITEMS);
#pragma unroll
    for (int k = 0; k < BOUNDS_ITEMS; k++) {
        uint32_t i = base + k * BLOCK + threadIdx.x;
        if (i < n) {
            f3 c = ld3(cent, i);
            mn = vmin3(mn, c);
            mx = vmax3(mx, c);
        }
    }
    block_minmax(mn, mx);
    if (threadIdx.x == 0) {
        float *p = part + blockIdx.x * 6;
        p[0] = mn.x; p[1] = mn.y; p[2] = mn.z; p[3] = mx.x; p[4] = mx.y; p[5] = mx.z;
    }
}

__global__ void __launch_bounds__(BLOCK) k_bounds_final(uint32_t nparts, const float *__restrict__ part,
                                                        float *__restrict__ bounds) {
    f3 mn = mk(1e30f, 1e30f, 1e30f), mx = mk(-1e30f, -1e30f, -1e30f);
    for (uint32_t i = threadIdx.x; i < nparts; i += BLOCK) {
        const float *p = part + i * 6;
        mn = vmin3(mn, mk(p[0], p[1], p[2]));
        mx = vmax3(mx, mk(p[3], p[4], p[5]));
    }
    block_minmax(mn, mx);
    if (threadIdx.x == 0) {
        bounds[0] = mn.x; bounds[1] = mn.y; bounds[2] = mn.z;
        bounds[3] = mx.x; bounds[4] = mx.y; bounds[5] = mx.z;
    }
}

// bvh.comp:32-46
SVD uint32_t expand_bits(uint32_t v) {
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}
SVD uint32_t quant10(float x) { return (uint32_t)gmin(gmax(x * 1024.0f, 0.0f), 1023.0f); }

// bvh.comp:169-183
__global__ void k_morton(uint32_t n, const float *__restrict__ bounds, const float *__restrict__ cent,
                         uint32_t *__restrict__ morton) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    f3 bmin = mk(bounds[0], bounds[1], bounds[2]);
    f3 ext = mk(bounds[3], bounds[4], bounds[5]) - bmin;
    f3 inv = mk(ext.x > 1e-6f ? 1.0f / ext.x : 0.0f, ext.y > 1e-6f ? 1.0f / ext.y : 0.0f,
                ext.z > 1e-6f ? 1.0f / ext.z : 0.0f);
    f3 c = hmul(ld3(cent, i) - bmin, inv);
    morton[i] = (expand_bits(quant10(c.x)) << 2) | (expand_bits(quant10(c.y)) << 1) | expand_bits(quant10(c.z));
}

// bvh.comp:48-65: clz with clz(0) = 32; equal keys fall back to the index bits.
SVD int clz32(uint32_t x) { return x == 0u ? 32 : __builtin_clz(x); }
SVD int delta(const uint32_t *__restrict__ m, int i, int j, int n) {
    if (j < 0 || j >= n) return -1;
    uint32_t a = m[i], b = m[j];
    if (a == b) return 32 + clz32((uint32_t)i ^ (uint32_t)j);
    return clz32(a ^ b);
}

// bvh.comp:299-348 (Karras 2012): leaves n-1+i, internal nodes 0..n-2.
// With `box` given (the broadphase tree) thread i also writes leaf i's float4
// box, body ids[i]'s (lo.w = the body index, hi.w = 1 if static), and clears
// internal node i's refit arrival counter.
__global__ void k_hierarchy(uint32_t nt, const uint32_t *__restrict__ m, uint32_t *__restrict__ left,
                            uint32_t *__restrict__ right, uint32_t *__restrict__ count,
                            uint32_t *__restrict__ parent, uint2 *__restrict__ rng,
                            const uint32_t *__restrict__ ids = nullptr, const float4 *__restrict__ box = nullptr,
                            float4 *__restrict__ nlo = nullptr, float4 *__restrict__ nhi = nullptr,
                            uint32_t *__restrict__ flags = nullptr) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    int n = (int)nt;
    if (i < nt) {
        uint32_t leaf = nt - 1 + i;
        left[leaf] = i;
        count[leaf] = 1;
        if (nt == 1) parent[0] = 0xFFFFFFFFu;  // single leaf is the root
        if (box) {
            const uint32_t b = ids[i];
            const float4 lo = box[2 * b], hi = box[2 * b + 1];
            nlo[leaf] = make_float4(lo.x, lo.y, lo.z, __uint_as_float(b));
            nhi[leaf] = make_float4(hi.x, hi.y, hi.z, __uint_as_float(lo.w != 0.0f ? 1u : 0u));
            if (i + 1 < nt) flags[i] = 0u;
        }
    }
    if (i + 1 >= nt) return;
    int ii = (int)i;
    int dl = delta(m, ii, ii - 1, n), dr = delta(m, ii, ii + 1, n);
    int d = dr > dl ? 1 : -1;
    int dmin = delta(m, ii, ii - d, n);
    int lmax = 2;
    while (delta(m, ii, ii + lmax * d, n) > dmin) lmax *= 2;
    int l = 0, t = lmax / 2;
    while (t >= 1) {
        if (delta(m, ii, ii + (l + t) * d, n) > dmin) l += t;
        t /= 2;
    }
    int j = ii + l * d;
    int first = ii < j ? ii : j, last = ii < j ? j : ii;
    int dn = delta(m, first, last, n);
    int split = first, stride = last - first;
    while (stride > 1) {
        stride = (stride + 1) / 2;
        int mid = split + stride;
        if (mid < last && delta(m, first, mid, n) > dn) split = mid;
    }
    uint32_t lc = (uint32_t)split, rc = (uint32_t)(split + 1);
    if (first == split) lc = (uint32_t)split + nt - 1;
    if (last == split + 1) rc = (uint32_t)(split + 1) + nt - 1;
    left[i] = lc;
    right[i] = rc;
    count[i] = 0;
    parent[lc] = i;
    parent[rc] = i;
    if (rng) rng[i] = make_uint2((uint32_t)first, (uint32_t)last);  // for the block-local refit
    if (i == 0) parent[0] = 0xFFFFFFFFu;
}

// bvh.comp:352-368
__global__ void k_leaf_tris(uint32_t n, const uint32_t *__restrict__ sort_idx, const float *__restrict__ verts,
                            const uint32_t *__restrict__ idx, uint32_t *__restrict__ prim,
                            float *__restrict__ nmin, float *__restrict__ nmax) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t leaf = n - 1 + i, pi = sort_idx[i];
    prim[i] = pi;
    f3 a = ld3(verts, idx[pi * 3]), b = ld3(verts, idx[pi * 3 + 1]), c = ld3(verts, idx[pi * 3 + 2]);
    st3(nmin, leaf, vmin3(vmin3(a, b), c));
    st3(nmax, leaf, vmax3(vmax3(a, b), c));
}

__global__ void k_leaf_bodies(uint32_t n, const uint32_t *__restrict__ sort_idx, const float *__restrict__ amin,
                              const float *__restrict__ amax, uint32_t *__restrict__ prim,
                              float *__restrict__ nmin, float *__restrict__ nmax) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t leaf = n - 1 + i, pi = sort_idx[i];
    prim[i] = pi;
    st3(nmin, leaf, ld3(amin, pi));
    st3(nmax, leaf, ld3(amax, pi));
}

// bvh.comp:372-375
__global__ void k_clear_flags(uint32_t n, uint32_t *flags) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i + 1 < n) flags[i] = 0;
}

// ---------------------------------------------------------------------------
// Bottom-up refit, bvh.comp:379-397: a thread per leaf climbs; at each parent
// the first arrival stops and the second writes the union of the two child
// boxes (left child first, as the reference's min/max read them) and goes on.
// One launch, no host read.  The box hand-off between the two arrivals uses
// device-coherent accesses (agent-scope relaxed atomics: SC1 loads and stores
// served at the point of coherence, so another XCD's L2 holds no stale copy):
// a climber waits until its box stores have completed before it bumps the
// parent's arrival counter, and the second arrival reads the sibling's box
// only after its own bump returned the first's.  No cache write-back or
// invalidation is needed, which made the fenced climb slow (3.8 ms at 1M
// leaves).  The continuing thread keeps its own child's box in registers.
// A node's key prefix strictly lengthens toward the leaves (at most 64
// bits), so a climb takes at most 65 steps; the bound only guards a corrupt
// parent array.
SVD float ldc(const float *p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
SVD void stc(float *p, float v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
SVD f3 ld3c(const float *p, uint32_t i) { return mk(ldc(p + 3 * (size_t)i), ldc(p + 3 * (size_t)i + 1), ldc(p + 3 * (size_t)i + 2)); }
SVD void st3c(float *p, uint32_t i, f3 v) {
    stc(p + 3 * (size_t)i, v.x);
    stc(p + 3 * (size_t)i + 1, v.y);
    stc(p + 3 * (size_t)i + 2, v.z);
}
SVD float4 ld4c(const float4 *p, uint32_t i) {
    const float *q = reinterpret_cast<const float *>(p + i);
    return make_float4(ldc(q), ldc(q + 1), ldc(q + 2), ldc(q + 3));
}
SVD void st4c(float4 *p, uint32_t i, float4 v) {
    float *q = reinterpret_cast<float *>(p + i);
    stc(q, v.x); stc(q + 1, v.y); stc(q + 2, v.z); stc(q + 3, v.w);
}
// Arrival at parent p after this thread's stores completed: true for the
// second of the two arrivals.  The caller's topology reads are inputs of the
// first empty asm, so they are issued before the bump and overlap the store
// completion; the second keeps the sibling loads below the bump's return.
SVD bool arrive_second(uint32_t *flag, uint32_t li, uint32_t ri, uint32_t pp) {
    asm volatile("" ::"v"(li), "v"(ri), "v"(pp) : "memory");
    __builtin_amdgcn_s_waitcnt(0);
    const uint32_t old = __hip_atomic_fetch_add(flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    asm volatile("" ::: "memory");
    return old != 0u;
}

// float3 node arrays of the reference layout (SV_BVH_AABB_MIN/MAX)
__global__ void __launch_bounds__(BLOCK) k_propagate(uint32_t n, const uint32_t *__restrict__ parent_of,
                                                     const uint32_t *__restrict__ left,
                                                     const uint32_t *__restrict__ right, uint32_t *flags,
                                                     float *nmin, float *nmax) {
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    uint32_t node = n - 1 + i;
    f3 mn = ld3(nmin, node), mx = ld3(nmax, node);  // leaf box of the previous launch
    uint32_t p = parent_of[node];
    for (int depth = 0; p != 0xFFFFFFFFu && depth < 80; depth++) {
        // topology is immutable: read it before the arrival, so the
        // sibling's box is the only round trip after it
        const uint32_t li = left[p], ri = right[p], pp = parent_of[p];
        if (!arrive_second(flags + p, li, ri, pp)) return;
        const bool is_left = li == node;
        const uint32_t sib = is_left ? ri : li;
        const f3 smn = ld3c(nmin, sib), smx = ld3c(nmax, sib);
        mn = is_left ? vmin3(mn, smn) : vmin3(smn, mn);
        mx = is_left ? vmax3(mx, smx) : vmax3(smx, mx);
        st3c(nmin, p, mn);
        st3c(nmax, p, mx);
        node = p;
        p = pp;
    }
}

// Two-launch refit, the same boxes as the climb above.  Internal node i
// covers a leaf range with i at one end (Karras).  Launch 1: workgroup k
// takes leaves [kT, kT + T) (T = REFIT_T); a node whose range lies inside
// them has its index there too, so its arrival counter and box live in LDS
// (index i - kT), both its arrivals are threads of this workgroup (an LDS
// counter and LDS boxes order them), and the workgroup stages its leaves'
// boxes and parents and the topology and leaf ranges of internal nodes
// [kT, kT + T) in LDS with one round of loads.  A thread whose node's parent
// crosses the workgroup's leaves stops there and lists the node in the
// workgroup's own slots of the list (T slots and a count per workgroup: no
// shared counter); the workgroup then writes the boxes of the nodes it
// completed in index order (coalesced).  Launch 2: one thread per listed
// node climbs the crossing nodes, all of which are ancestors of listed
// nodes, with the device-coherent hand-off of k_propagate (the listed
// nodes' boxes are visible to it across the launch boundary).  No thread
// waits, and the long climbs of launch 2 run side by side instead of holding
// launch 1's workgroups.  Leaf ranges come from k_hierarchy.
constexpr int REFIT_T = 512;
constexpr int REFIT_LANES = 32;  // launch 2: threads per listed workgroup
// List words for n leaves: T slots and a count per workgroup of launch 1.
inline size_t refit_list_words(size_t n) { return (n + REFIT_T - 1) / REFIT_T * (REFIT_T + 1) + 1; }
struct Box3 {
    f3 mn, mx;
};
struct Box4 {
    float4 lo, hi;
};
SVD Box3 box_load(const float *mn, const float *mx, uint32_t i) { return {ld3(mn, i), ld3(mx, i)}; }
SVD Box3 box_load_coh(const float *mn, const float *mx, uint32_t i) { return {ld3c(mn, i), ld3c(mx, i)}; }
SVD void box_store(float *mn, float *mx, uint32_t i, const Box3 &b) { st3(mn, i, b.mn); st3(mx, i, b.mx); }
SVD void box_store_coh(float *mn, float *mx, uint32_t i, const Box3 &b) { st3c(mn, i, b.mn); st3c(mx, i, b.mx); }
SVD Box3 box_union(const Box3 &l, const Box3 &r) { return {vmin3(l.mn, r.mn), vmax3(l.mx, r.mx)}; }  // left first
SVD Box4 box_load(const float4 *lo, const float4 *hi, uint32_t i) { return {lo[i], hi[i]}; }
SVD Box4 box_load_coh(const float4 *lo, const float4 *hi, uint32_t i) { return {ld4c(lo, i), ld4c(hi, i)}; }
SVD void box_store(float4 *lo, float4 *hi, uint32_t i, const Box4 &b) { lo[i] = b.lo; hi[i] = b.hi; }
SVD void box_store_coh(float4 *lo, float4 *hi, uint32_t i, const Box4 &b) { st4c(lo, i, b.lo); st4c(hi, i, b.hi); }
// The broadphase tree's node record (k_bb_pairs): the two children's boxes
// side by side, written by the thread that completes the node.
// The walk's 64-byte record of internal node p, per child (left, then
// right): its box, with lo.w = the smallest body index below it, bit 31 set
// when every body below is static (indices are < 2^31), and hi.w = the
// child's node index, so that a visit needs no other load.
SVD float4 rec_lo(const Box4 &b) {
    return make_float4(b.lo.x, b.lo.y, b.lo.z,
                       __uint_as_float(__float_as_uint(b.lo.w) | (__float_as_uint(b.hi.w) << 31)));
}
SVD void rec_pair(float4 *rec, uint32_t p, uint32_t li, const Box4 &l, uint32_t ri, const Box4 &r) {
    if (!rec) return;
    float4 *q = rec + 4 * (size_t)p;
    q[0] = rec_lo(l);
    q[1] = make_float4(l.hi.x, l.hi.y, l.hi.z, __uint_as_float(li));
    q[2] = rec_lo(r);
    q[3] = make_float4(r.hi.x, r.hi.y, r.hi.z, __uint_as_float(ri));
}
SVD void rec_pair(float4 *, uint32_t, uint32_t, const Box3 &, uint32_t, const Box3 &) {}
SVD Box4 box_union(const Box4 &l, const Box4 &r) {  // see k_bb_propagate
    const uint32_t mi = min(__float_as_uint(l.lo.w), __float_as_uint(r.lo.w));
    const uint32_t st = __float_as_uint(l.hi.w) & __float_as_uint(r.hi.w);
    return {make_float4(gmin(l.lo.x, r.lo.x), gmin(l.lo.y, r.lo.y), gmin(l.lo.z, r.lo.z), __uint_as_float(mi)),
            make_float4(gmax(l.hi.x, r.hi.x), gmax(l.hi.y, r.hi.y), gmax(l.hi.z, r.hi.z), __uint_as_float(st))};
}

template <typename T, typename BoxT>
__global__ void __launch_bounds__(REFIT_T) k_refit_blocks(uint32_t n, const uint32_t *__restrict__ parent_of,
                                                          const uint32_t *__restrict__ left,
                                                          const uint32_t *__restrict__ right,
                                                          const uint2 *__restrict__ rng, T *na, T *nb,
                                                          uint32_t *__restrict__ list, float4 *__restrict__ rec) {
    constexpr uint32_t TH = REFIT_T;
    __shared__ uint32_t lflag[TH];
    __shared__ BoxT lbox[2 * TH];  // [0, TH): internal nodes b0 + k; [TH, 2 TH): leaves b0 + k
    __shared__ uint32_t lleft[TH], lright[TH], lpar[TH], lleafpar[TH];
    __shared__ uint2 lrng[TH];
    __shared__ uint32_t s_ntop;
    const uint32_t b0 = blockIdx.x * TH, t = threadIdx.x, i = b0 + t;
    lflag[t] = 0u;
    if (t == 0) s_ntop = 0u;
    BoxT box;
    if (i < n) {
        box = box_load(na, nb, n - 1 + i);  // leaf box of the previous launch
        lbox[TH + t] = box;
        lleafpar[t] = parent_of[n - 1 + i];
    }
    if (i + 1 < n) {
        lleft[t] = left[i];
        lright[t] = right[i];
        lpar[t] = parent_of[i];
        lrng[t] = rng[i];
    }
    __syncthreads();
    if (i < n) {
        uint32_t node = n - 1 + i;
        uint32_t p = lleafpar[t];
        for (int depth = 0; depth < 80; depth++) {
            if (p == 0xFFFFFFFFu) break;  // the root, inside this workgroup's leaves (its box is in LDS)
            bool plocal = false;
            if (p >= b0 && p - b0 < TH) {
                const uint2 r = lrng[p - b0];
                plocal = r.x >= b0 && r.y < b0 + TH;
            }
            if (!plocal) {  // p crosses the workgroup's leaves: launch 2 goes on from node
                list[(size_t)blockIdx.x * (TH + 1) + 1 + atomicAdd(&s_ntop, 1u)] = node;
                break;
            }
            // p is local, so its topology is staged.  LDS only: this
            // thread's box write has completed before its bump (LDS is one
            // memory for the workgroup, and a wave's LDS accesses run in
            // order), and the sibling's box is read after the bump returned.
            const uint32_t li = lleft[p - b0], ri = lright[p - b0], pp = lpar[p - b0];
            asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
            const uint32_t old = __hip_atomic_fetch_add(&lflag[p - b0], 1u, __ATOMIC_RELAXED,
                                                        __HIP_MEMORY_SCOPE_WORKGROUP);
            asm volatile("" ::: "memory");
            if (old == 0u) break;
            const bool is_left = li == node;
            const uint32_t sib = is_left ? ri : li;
            const BoxT sbox = sib >= n - 1 ? lbox[TH + (sib - (n - 1) - b0)] : lbox[sib - b0];
            box = is_left ? box_union(box, sbox) : box_union(sbox, box);
            lbox[p - b0] = box;
            node = p;
            p = pp;
        }
    }
    __syncthreads();
    // the internal nodes completed here (both arrivals counted), written out
    // in index order: consecutive threads store consecutive boxes; such a
    // node covers leaves of this workgroup only, so its children are leaves
    // or internal nodes of this workgroup and their boxes are in LDS (the
    // node's record of them goes out the same way)
    if (i + 1 < n && lflag[t] == 2u) {
        box_store(na, nb, i, lbox[t]);
        if (rec) {
            const uint32_t li = lleft[t], ri = lright[t];
            rec_pair(rec, i, li, li >= n - 1 ? lbox[TH + (li - (n - 1) - b0)] : lbox[li - b0], ri,
                     ri >= n - 1 ? lbox[TH + (ri - (n - 1) - b0)] : lbox[ri - b0]);
        }
    }
    if (t == 0) list[(size_t)blockIdx.x * (TH + 1)] = s_ntop;
}

template <typename T, typename BoxT>
__global__ void __launch_bounds__(BLOCK) k_refit_cross(uint32_t n, const uint32_t *__restrict__ parent_of,
                                                       const uint32_t *__restrict__ left,
                                                       const uint32_t *__restrict__ right, uint32_t *flags, T *na,
                                                       T *nb, const uint32_t *__restrict__ list, float4 *__restrict__ rec) {
    const uint32_t nw = (n + REFIT_T - 1) / REFIT_T;  // workgroups of launch 1
    const uint32_t w = blockIdx.x * (BLOCK / REFIT_LANES) + threadIdx.x / REFIT_LANES;
    if (w >= nw) return;
    const uint32_t *slot = list + (size_t)w * (REFIT_T + 1);
    const uint32_t m = slot[0];
    for (uint32_t k = threadIdx.x % REFIT_LANES; k < m; k += REFIT_LANES) {
        uint32_t node = slot[1 + k];
        BoxT box = box_load(na, nb, node);  // written by launch 1
        uint32_t p = parent_of[node];
        for (int depth = 0; p != 0xFFFFFFFFu && depth < 80; depth++) {
            const uint32_t li = left[p], ri = right[p], pp = parent_of[p];
            if (!arrive_second(flags + p, li, ri, pp)) break;
            const bool is_left = li == node;
            const uint32_t sib = is_left ? ri : li;
            const BoxT sbox = box_load_coh(na, nb, sib);
            const BoxT own = box;
            box = is_left ? box_union(own, sbox) : box_union(sbox, own);
            rec_pair(rec, p, li, is_left ? own : sbox, ri, is_left ? sbox : own);
            box_store_coh(na, nb, p, box);
            node = p;
            p = pp;
        }
    }
}

// Both launches of the two-launch refit, after k_clear_flags (which clears
// the arrival counters); `list` holds refit_list_words(n) words.
template <typename T, typename BoxT>
void launch_refit(hipStream_t st, uint32_t n, const uint32_t *parent_of, const uint32_t *left, const uint32_t *right,
                  const uint2 *rng, uint32_t *flags, T *na, T *nb, uint32_t *list, float4 *rec = nullptr) {
    const uint32_t nw = div_up(n, REFIT_T);
    hipLaunchKernelGGL((k_refit_blocks<T, BoxT>), dim3(nw), dim3(REFIT_T), 0, st, n, parent_of, left, right, rng, na,
                       nb, list, rec);
    hipLaunchKernelGGL((k_refit_cross<T, BoxT>), dim3(div_up(nw, BLOCK / REFIT_LANES)), dim3(BLOCK), 0, st, n,
                       parent_of, left, right, flags, na, nb, (const uint32_t *)list, rec);
}

__global__ void k_copy_u32(uint32_t n, const uint32_t *src, uint32_t *dst) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// ---------------------------------------------------------------------------
// Broadphase by LBVH traversal (extension beside physics.comp:451-472).
//
// The tree is the body LBVH above (box centres, 30-bit Morton codes, Karras
// hierarchy) with float4 nodes: lo.w carries the smallest body index below
// the node and hi.w whether every body below is static.  Each body j walks
// the tree with its own box and reports every leaf body i < j whose box is
// not apart from its own and that is not static together with j.  A node is
// skipped when j's box is apart from the node's box (the node box is the
// exact union of the leaves: a bound of j's box below the node's lower bound
// is below every leaf's, so no leaf below can overlap), when its smallest index is >= j, or when j and the
// whole subtree are static.  The reported set is exactly the all-pairs set.
// Pairs are counted per row i, the row counts scanned into offsets (total =
// counters[0]), the pairs of rows that start below max_pairs appended in any
// order, and that part sorted by (i, j): the first max_pairs entries are the
// all-pairs kernel's.

// Box centres, the identity body list, and the centres' bounds in one launch:
// each workgroup reduces its BOUNDS_TILE centres to a partial as
// k_bounds_partial does, and the workgroup that arrives last at `done`
// reduces the partials as k_bounds_final does (same partition, same order)
// and puts the counter back to zero for the next launch.  Partials go out
// and come back with device-coherent accesses; the arrival waits for the
// partial's stores to complete.  Also zeroes the row counts the counting
// walk bumps and the append counter of the pair list.
__global__ void __launch_bounds__(BLOCK) k_bb_centroids(uint32_t n, const float4 *__restrict__ box,
                                                        float *__restrict__ cent, uint32_t *__restrict__ ids,
                                                        float *part, float *__restrict__ bounds, uint32_t *done,
                                                        uint32_t *__restrict__ cnt, uint32_t *__restrict__ tail) {
    __shared__ uint32_t s_last;
    f3 mn = mk(1e30f, 1e30f, 1e30f), mx = mk(-1e30f, -1e30f, -1e30f);
    const uint32_t base = blockIdx.x * BOUNDS_TILE;
    if (blockIdx.x == 0 && threadIdx.x == 0) tail[0] = 0u;
#pragma unroll
    for (int k = 0; k < BOUNDS_ITEMS; k++) {
        const uint32_t i = base + k * BLOCK + threadIdx.x;
        if (i < n) {
            const float4 lo = box[2 * i], hi = box[2 * i + 1];
            const f3 c = mk((lo.x + hi.x) * 0.5f, (lo.y + hi.y) * 0.5f, (lo.z + hi.z) * 0.5f);
            st3(cent, i, c);
            ids[i] = i;
            cnt[i] = 0u;
            mn = vmin3(mn, c);
            mx = vmax3(mx, c);
        }
    }
    block_minmax(mn, mx);
    if (threadIdx.x == 0) {
        float *p = part + blockIdx.x * 6;
        stc(p + 0, mn.x); stc(p + 1, mn.y); stc(p + 2, mn.z);
        stc(p + 3, mx.x); stc(p + 4, mx.y); stc(p + 5, mx.z);
        __builtin_amdgcn_s_waitcnt(0);
        const uint32_t old = __hip_atomic_fetch_add(done, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_last = old + 1u == gridDim.x ? 1u : 0u;
    }
    __syncthreads();
    if (!s_last) return;
    asm volatile("" ::: "memory");
    mn = mk(1e30f, 1e30f, 1e30f);
    mx = mk(-1e30f, -1e30f, -1e30f);
    for (uint32_t q = threadIdx.x; q < gridDim.x; q += BLOCK) {
        const float *p = part + q * 6;
        mn = vmin3(mn, mk(ldc(p + 0), ldc(p + 1), ldc(p + 2)));
        mx = vmax3(mx, mk(ldc(p + 3), ldc(p + 4), ldc(p + 5)));
    }
    block_minmax(mn, mx);
    if (threadIdx.x == 0) {
        bounds[0] = mn.x; bounds[1] = mn.y; bounds[2] = mn.z;
        bounds[3] = mx.x; bounds[4] = mx.y; bounds[5] = mx.z;
        __hip_atomic_store(done, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// The broadphase tree's refit (see k_propagate): the union of the children,
// the smaller index minimum and the AND of the all-static flags.
__global__ void __launch_bounds__(BLOCK) k_bb_propagate(uint32_t n, const uint32_t *__restrict__ parent_of,
                                                        const uint32_t *__restrict__ left,
                                                        const uint32_t *__restrict__ right, uint32_t *flags,
                                                        float4 *nlo, float4 *nhi) {
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    uint32_t node = n - 1 + i;
    float4 lo = nlo[node], hi = nhi[node];
    uint32_t p = parent_of[node];
    for (int depth = 0; p != 0xFFFFFFFFu && depth < 80; depth++) {
        const uint32_t li = left[p], ri = right[p], pp = parent_of[p];
        if (!arrive_second(flags + p, li, ri, pp)) return;
        const bool is_left = li == node;
        const uint32_t sib = is_left ? ri : li;
        const float4 sl = ld4c(nlo, sib), sh = ld4c(nhi, sib);
        const float4 al = is_left ? lo : sl, ah = is_left ? hi : sh, bl = is_left ? sl : lo, bh = is_left ? sh : hi;
        const uint32_t mi = min(__float_as_uint(al.w), __float_as_uint(bl.w));
        const uint32_t st = __float_as_uint(ah.w) & __float_as_uint(bh.w);
        lo = make_float4(gmin(al.x, bl.x), gmin(al.y, bl.y), gmin(al.z, bl.z), __uint_as_float(mi));
        hi = make_float4(gmax(ah.x, bh.x), gmax(ah.y, bh.y), gmax(ah.z, bh.z), __uint_as_float(st));
        st4c(nlo, p, lo);
        st4c(nhi, p, hi);
        node = p;
        p = pp;
    }
}

constexpr int BB_STACK = 64;  // > tree depth: node prefixes strictly grow downward and lie in [2, 63]

// Workgroups go to the 8 XCDs round robin; block b is given the logical
// block that makes each XCD's share one contiguous range (so each XCD's L2
// holds the subtrees of one region of Morton order).  A bijection of
// [0, gridDim.x).
SVD uint32_t xcd_block(uint32_t b, uint32_t g) {
    const uint32_t x = b & 7u, q = b >> 3, per = g >> 3, extra = g & 7u;
    return x < extra ? x * (per + 1u) + q : extra * (per + 1u) + (x - extra) * per + q;
}

// The first BB_LDS stack entries of each lane live in LDS ([entry][lane],
// conflict free); deeper ones in a private array.
constexpr int BB_LDS = 16;

// Rows a counting lane stages for k_bb_scatter (more: that lane walks again).
constexpr int BB_STAGE = 32;

// Place the pairs (mi, j) of the lanes with `found` set whose row mi starts
// below max_pairs (the rows the capped list keeps) in their row: at
// off[mi] + the row's fill count, in no particular order (k_rows_short and
// k_rows_long sort each row).  The lanes holding the row of the first such
// lane take their slots with one atomic (the ground's row gathers many
// lanes of a wave), the others one each.
SVD void bb_place(bool found, uint32_t mi, uint32_t j, uint64_t lt, const uint32_t *__restrict__ off,
                  uint32_t *fill, uint32_t *__restrict__ pk, uint32_t max_pairs) {
    const uint32_t o = found ? off[mi] : 0u;
    const bool keep = found && o < max_pairs;
    const uint64_t km = __ballot(keep);
    if (!km) return;
    const int first = __builtin_ctzll(km);
    const uint32_t lead = __shfl(mi, first, 64);
    const uint64_t same = __ballot(keep && mi == lead);
    uint32_t base = 0;
    if (keep && (km & lt) == 0ull) base = atomicAdd(fill + lead, (uint32_t)__popcll(same));
    base = __shfl(base, first, 64);
    if (!keep) return;
    const uint32_t slot = mi == lead ? base + (uint32_t)__popcll(same & lt) : atomicAdd(fill + mi, 1u);
    pk[o + slot] = j;
}

// Append the pairs (mi, j) of the lanes with `found` set whose row mi starts
// below max_pairs (the rows the capped list keeps) at the end of the pair
// buffer: one atomic on the append counter per wave.  The pairs land in no
// particular order; the pair sort that follows orders them.
SVD void bb_append(bool found, uint32_t mi, uint32_t j, uint64_t lt, const uint32_t *__restrict__ off,
                   uint32_t *tail, uint32_t *__restrict__ pk, uint32_t *__restrict__ pv, uint32_t max_pairs, int bits,
                   bool combined) {
    const bool keep = found && off[mi] < max_pairs;
    const uint64_t km = __ballot(keep);
    if (!km) return;
    const int first = __builtin_ctzll(km);
    uint32_t base = 0;
    if (keep && __builtin_amdgcn_mbcnt_hi((uint32_t)(km >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)km, 0u)) == 0u)
        base = atomicAdd(tail, (uint32_t)__popcll(km));
    base = __shfl(base, first, 64);
    if (!keep) return;
    const uint32_t s = base + (uint32_t)__popcll(km & lt);
    if (combined) {
        pk[s] = (mi << bits) | j;
    } else {
        pk[s] = j;
        pv[s] = mi;
    }
}

// One lane per body j, in Morton order so neighbouring lanes walk similar
// paths.  COUNT: bump the row count of every reported i, and, with a
// staging buffer, also keep the first BB_STAGE reported rows of lane t in
// stage[k * n + t] and their number in scnt[t], so that k_bb_scatter appends
// them without a second walk.  EMIT: append the pairs whose row starts below
// max_pairs; with scnt given, only lanes that found more than BB_STAGE rows
// walk (k_bb_scatter appended the others).
template <bool EMIT>
__global__ void __launch_bounds__(BLOCK) k_bb_pairs(uint32_t n, const uint32_t *__restrict__ ids,
                                                    const float4 *__restrict__ box, const float4 *__restrict__ rec,
                                                    uint32_t *cnt,
                                                    const uint32_t *__restrict__ off, uint32_t *tail,
                                                    uint32_t *__restrict__ pk, uint32_t *__restrict__ pv,
                                                    uint32_t max_pairs, int bits, bool combined,
                                                    uint32_t *__restrict__ stage, uint32_t *__restrict__ scnt) {
    __shared__ uint32_t s_stack[BB_LDS * BLOCK];
    uint32_t t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (t >= n) return;
    if (EMIT && scnt != nullptr && scnt[t] <= (uint32_t)BB_STAGE) return;
    const uint32_t j = ids[t];
    const float4 jl = box[2 * j], jh = box[2 * j + 1];
    // a child is skipped when its tag (smallest index, bit 31 = all static)
    // under this mask is >= j: index >= j, or j and the subtree static
    const uint32_t jmask = jl.w != 0.0f ? 0xFFFFFFFFu : 0x7FFFFFFFu;
    uint32_t *lstack = s_stack + threadIdx.x;
    uint32_t stack[BB_STACK - BB_LDS];
    const int lane = threadIdx.x & 63;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    int sp = 0;
    uint32_t node = 0, nf = 0;
    for (;;) {
        const float4 *rn = rec + 4 * (size_t)node;
        const float4 cb[4] = {rn[0], rn[1], rn[2], rn[3]};  // the children's boxes, tags and indices
        // every field is wanted: this keeps the four loads together (the
        // compiler otherwise loaded a child's box only after its tag passed,
        // a second dependent round trip per visit)
        asm volatile("" ::"v"(cb[0].x), "v"(cb[0].y), "v"(cb[0].z), "v"(cb[0].w), "v"(cb[1].x), "v"(cb[1].y),
                     "v"(cb[1].z), "v"(cb[1].w), "v"(cb[2].x), "v"(cb[2].y), "v"(cb[2].z), "v"(cb[2].w),
                     "v"(cb[3].x), "v"(cb[3].y), "v"(cb[3].z), "v"(cb[3].w));
        bool go0 = false, go1 = false;
        uint32_t ch0 = 0, ch1 = 0;
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const float4 lo = cb[2 * q], hi = cb[2 * q + 1];
            const uint32_t c = __float_as_uint(hi.w), tag = __float_as_uint(lo.w);
            const uint32_t mi = tag & 0x7FFFFFFFu;
            const bool keep = ((tag & jmask) < j) & !boxes_apart(jl, jh, lo, hi);
            const bool leaf = keep && c >= n - 1;
            if (q == 0) {
                go0 = keep && !leaf;
                ch0 = c;
            } else {
                go1 = keep && !leaf;
                ch1 = c;
            }
            if (EMIT) {
                bb_append(leaf, mi, j, lt, off, tail, pk, pv, max_pairs, bits, combined);
            } else {
                // one atomic for the first finding lane's row and its lanes
                const uint64_t lm = __ballot(leaf);
                if (lm) {
                    const int first = __builtin_ctzll(lm);
                    const uint32_t lead = __shfl(mi, first, 64);
                    const uint64_t same = __ballot(leaf && mi == lead);
                    if (lane == first) atomicAdd(cnt + lead, (uint32_t)__popcll(same));
                    if (leaf) {
                        if (mi != lead) atomicAdd(cnt + mi, 1u);
                        if (stage != nullptr && nf < (uint32_t)BB_STAGE) stage[nf * n + t] = mi;
                        ++nf;
                    }
                }
            }
        }
        // descend into the right child when it goes on (the left one, if it
        // also does, waits on the stack), else into the left one, else take
        // the stack's top: the visiting order of pushing both and popping
        if (go0 && go1) {
            if (sp < BB_LDS) lstack[sp++ * BLOCK] = ch0;
            else if (sp < BB_STACK) stack[sp++ - BB_LDS] = ch0;
        }
        node = go1 ? ch1 : ch0;
        if (!go0 && !go1) {
            if (sp == 0) break;
            --sp;
            // an LDS read, and a private one in a branch of its own whose
            // value an empty asm takes (else the compiler merged the reads
            // into one flat load through a selected address)
            uint32_t top = lstack[(sp < BB_LDS ? sp : BB_LDS - 1) * BLOCK];
            if (sp >= BB_LDS) {
                top = stack[sp - BB_LDS];
                asm volatile("" : "+v"(top));
            }
            node = top;
        }
    }
    if (!EMIT && stage != nullptr) scnt[t] = nf;
}

// ---------------------------------------------------------------------------
// Rows of the pair list in order.  Row i holds cnt[i] distinct partners j > i
// at pk[off[i] ...] in no particular order (off = exclusive scan of cnt);
// sorted, they are the pairs (i, j) at slots off[i] + k of the list, which
// keeps the slots below max_pairs (rows that start at or above it are not
// there).  A row of at most ROW_SHORT partners is sorted by one lane in
// registers (Batcher's odd-even merge network); longer rows are listed for
// k_rows_long: a workgroup ranks up to ROW_RANK_MAX partners in LDS (the
// rank of j = the partners below it) when that is cheaper than marking the
// row in an LDS bitmap over the partner range, window by window, whose bits
// in order are the row's partners in order.
constexpr uint32_t ROW_SHORT = 16;
constexpr uint32_t ROW_RANK_MAX = 2048;
constexpr uint32_t ROW_WINDOW_WORDS = 1u << 15;  // 2^20 partner indices (128 KiB of LDS) per window

SVD void cmpx(uint32_t &a, uint32_t &b) {
    const uint32_t lo = min(a, b), hi = max(a, b);
    a = lo;
    b = hi;
}

__device__ void sap_map_next(float *map);

__global__ void __launch_bounds__(BLOCK) k_rows_short(uint32_t n, const uint32_t *__restrict__ cnt,
                                                      const uint32_t *__restrict__ off,
                                                      const uint32_t *__restrict__ pk, uint32_t *__restrict__ pairs,
                                                      uint32_t max_pairs, uint32_t *list, float *map) {
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (map && i == 0) sap_map_next(map);  // the sweep's map for its next launch
    if (i >= n) return;
    const uint32_t c = cnt[i];
    if (c == 0) return;
    const uint32_t o = off[i];
    if (o >= max_pairs) return;
    if (c > ROW_SHORT) {
        list[1 + atomicAdd(list, 1u)] = i;
        return;
    }
    uint32_t v[ROW_SHORT];
#pragma unroll
    for (uint32_t k = 0; k < ROW_SHORT; k++) v[k] = k < c ? pk[o + k] : 0xFFFFFFFFu;
    // Batcher's odd-even merge sort of 16 (Knuth 5.3.4, algorithm M)
#pragma unroll
    for (uint32_t p = 1; p < ROW_SHORT; p <<= 1)
#pragma unroll
        for (uint32_t k = p; k >= 1; k >>= 1)
#pragma unroll
            for (uint32_t jj = k % p; jj + k < ROW_SHORT; jj += 2 * k)
#pragma unroll
                for (uint32_t q = 0; q < k; q++)
                    if (q + jj + k < ROW_SHORT && (q + jj) / (2 * p) == (q + jj + k) / (2 * p)) cmpx(v[q + jj], v[q + jj + k]);
    uint2 *out = reinterpret_cast<uint2 *>(pairs);
#pragma unroll
    for (uint32_t k = 0; k < ROW_SHORT; k++)
        if (k < c && o + k < max_pairs) out[o + k] = make_uint2(i, v[k]);
}

__global__ void __launch_bounds__(BLOCK) k_rows_long(uint32_t n, const uint32_t *__restrict__ cnt,
                                                     const uint32_t *__restrict__ off,
                                                     const uint32_t *__restrict__ pk, uint32_t *__restrict__ pairs,
                                                     uint32_t max_pairs, const uint32_t *__restrict__ list,
                                                     uint32_t window_words) {
    extern __shared__ uint4 lds_raw[];  // max(ROW_RANK_MAX, window_words) words, 16-byte aligned
    uint32_t *lds = reinterpret_cast<uint32_t *>(lds_raw);
    __shared__ uint32_t s_part[BLOCK / 64 + 1];
    uint2 *out = reinterpret_cast<uint2 *>(pairs);
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t m = list[0];
    for (uint32_t q = blockIdx.x; q < m; q += gridDim.x) {
        const uint32_t i = list[1 + q], c = cnt[i], o = off[i];
        __syncthreads();  // the previous row's LDS reads are done
        // ranks cost about c^2 compares over the workgroup, a bitmap about
        // three passes over its (n - i) / 32 words plus c
        if (c <= ROW_RANK_MAX && (uint64_t)c * c <= 3ull * ((n - i) / 32u) * BLOCK / 64u + (uint64_t)c * 64u) {
            for (uint32_t k = tid; k < c; k += BLOCK) lds[k] = pk[o + k];
            for (uint32_t k = c + tid; k < ((c + 3u) & ~3u); k += BLOCK) lds[k] = 0xFFFFFFFFu;  // pad to 4
            __syncthreads();
            const uint4 *l4 = lds_raw;
            for (uint32_t k = tid; k < c; k += BLOCK) {
                const uint32_t x = lds[k];
                uint32_t r = 0;
                for (uint32_t u = 0; u < (c + 3u) / 4u; u++) {  // four partners per LDS read
                    const uint4 v = l4[u];
                    r += (v.x < x ? 1u : 0u) + (v.y < x ? 1u : 0u) + (v.z < x ? 1u : 0u) + (v.w < x ? 1u : 0u);
                }
                if (o + r < max_pairs) out[o + r] = make_uint2(i, x);
            }
            continue;
        }
        // partners lie in [i + 1, n): windows of 32 * window_words indices
        const uint32_t wbits = window_words * 32u;
        uint32_t before = 0;  // partners in earlier windows
        for (uint32_t w0 = i + 1; w0 < n; w0 += wbits) {
            for (uint32_t k = tid; k < window_words; k += BLOCK) lds[k] = 0u;
            __syncthreads();
            for (uint32_t k = tid; k < c; k += BLOCK) {
                const uint32_t d = pk[o + k] - w0;
                if (d < wbits) atomicOr(&lds[d >> 5], 1u << (d & 31u));
            }
            __syncthreads();
            // thread t owns words [t * per, (t + 1) * per): its bits' count,
            // scanned across the workgroup, is the rank of its first bit
            const uint32_t per = window_words / BLOCK;
            uint32_t bitsum = 0;
            for (uint32_t k = 0; k < per; k++) bitsum += (uint32_t)__popc(lds[tid * per + k]);
            uint32_t x = bitsum;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const uint32_t y = __shfl_up(x, d, 64);
                if ((int)lane >= d) x += y;
            }
            if (lane == 63) s_part[wave] = x;
            __syncthreads();
            uint32_t rank = before + x - bitsum, total = 0;
            for (uint32_t w = 0; w < BLOCK / 64; w++) {
                if (w < wave) rank += s_part[w];
                total += s_part[w];
            }
            for (uint32_t k = 0; k < per; k++) {
                uint32_t bitsw = lds[tid * per + k];
                while (bitsw) {
                    const uint32_t b = (uint32_t)__builtin_ctz(bitsw);
                    bitsw &= bitsw - 1u;
                    if (o + rank < max_pairs) out[o + rank] = make_uint2(i, w0 + (tid * per + k) * 32u + b);
                    ++rank;
                }
            }
            before += total;
            __syncthreads();  // the window's LDS reads are done before the next clear
        }
    }
}

// k_bb_post's work (the pair count and the overflow flag), and where the
// kept part of the pair list ends: rows that start below max_pairs are kept.
// The offsets are an exclusive scan of the counts, so row ends off + cnt do
// not decrease with the row, and the kept rows are a prefix [0, r): the kept
// part ends at off[r-1] + cnt[r-1].  r (the first row in [1, n] whose offset
// is >= max_pairs, n if none; off[0] = 0 < max_pairs) is found by one wave:
// each round its lanes probe 64 evenly spaced rows of the interval that
// holds r, and r lies after the last probe below max_pairs and at or before
// the first one at or above it (a few rounds instead of log2(n)).
SVD void bb_rows(uint32_t n, const uint32_t *__restrict__ cnt, const uint32_t *__restrict__ off, uint32_t *misc,
                 uint32_t max_pairs, uint32_t *counters) {
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t lo = 1, hi = n;  // r in [lo, hi]; rows below lo start below max_pairs
    while (lo < hi) {
        const uint32_t step = (hi - lo + 63u) / 64u, q = lo + lane * step;
        const bool valid = q < hi;
        const uint64_t ge = __ballot(valid && off[q] >= max_pairs);
        if (ge) {
            const uint32_t k = (uint32_t)__builtin_ctzll(ge);
            hi = lo + k * step;
            if (k) lo = lo + (k - 1u) * step + 1u;
        } else {
            const uint32_t last = (uint32_t)(63 - __builtin_clzll(__ballot(valid)));
            lo = lo + last * step + 1u;
        }
    }
    if (lane == 0) {
        const uint32_t t = counters[0];
        misc[MISC_NP] = t < max_pairs ? t : max_pairs;
        if (t > max_pairs) counters[2] |= SV_STATUS_PAIR_OVERFLOW;
        misc[MISC_BPE] = off[lo - 1] + cnt[lo - 1];
    }
}

// Thread x of a workgroup appends the rows staged by lanes t = base + x +
// k * BLOCK, k < BB_SCATTER_ITEMS (scnt[t] <= BB_STAGE); lanes with more
// rows are left to the EMIT walk.  One append atomic per workgroup (the
// atomics of one counter serialise, so a workgroup takes many lanes): the
// waves' kept counts are scanned in the workgroup (all rows are kept when
// the total fits max_pairs; otherwise each staged row's offset is checked),
// and each wave writes its entries in rounds across its lanes.
constexpr int BB_SCATTER_ITEMS = 8;

SVD uint32_t bb_staged(const uint32_t *__restrict__ scnt, uint32_t t, uint32_t n) {
    const uint32_t m = t < n ? scnt[t] : 0u;
    return m > (uint32_t)BB_STAGE ? 0u : m;
}

__global__ void __launch_bounds__(BLOCK) k_bb_scatter(uint32_t n, const uint32_t *__restrict__ ids,
                                                      const uint32_t *__restrict__ stage,
                                                      const uint32_t *__restrict__ scnt,
                                                      const uint32_t *__restrict__ off,
                                                      uint32_t *counters, uint32_t *tail,
                                                      uint32_t *__restrict__ pk, uint32_t *__restrict__ pv,
                                                      uint32_t max_pairs, int bits, bool combined,
                                                      const uint32_t *__restrict__ cnt, uint32_t *misc) {
    __shared__ uint32_t s_wave[BLOCK / 64 + 1];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (blockIdx.x == 0 && wave == 0) bb_rows(n, cnt, off, misc, max_pairs, counters);
    const uint32_t base = blockIdx.x * (BLOCK * BB_SCATTER_ITEMS) + threadIdx.x;
    const bool all_kept = counters[0] <= max_pairs;
    uint32_t tot = 0;
    for (int it = 0; it < BB_SCATTER_ITEMS; it++) {
        const uint32_t t = base + (uint32_t)(it * BLOCK), m = bb_staged(scnt, t, n);
        if (all_kept) {
            tot += m;
        } else {
            for (uint32_t k = 0; k < m; k++) tot += off[stage[k * n + t]] < max_pairs ? 1u : 0u;
        }
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) tot += __shfl_xor(tot, d, 64);
    if (lane == 0) s_wave[wave] = tot;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (int w = 0; w < BLOCK / 64; w++) {
            const uint32_t v = s_wave[w];
            s_wave[w] = acc;
            acc += v;
        }
        s_wave[BLOCK / 64] = acc ? atomicAdd(tail, acc) : 0u;
    }
    __syncthreads();
    uint32_t o = s_wave[BLOCK / 64] + s_wave[wave];  // the wave's first slot
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int it = 0; it < BB_SCATTER_ITEMS; it++) {
        const uint32_t t = base + (uint32_t)(it * BLOCK), m = bb_staged(scnt, t, n);
        const uint32_t j = m ? ids[t] : 0u;
        for (uint32_t k0 = 0; __ballot(k0 < m); k0 += 4) {
            uint32_t mi[4];  // four staged rows in flight per lane
#pragma unroll
            for (int q = 0; q < 4; q++) mi[q] = k0 + q < m ? stage[(k0 + q) * n + t] : 0u;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const bool keep = k0 + q < m && (all_kept || off[mi[q]] < max_pairs);
                const uint64_t km = __ballot(keep);
                if (keep) {
                    const uint32_t s = o + (uint32_t)__popcll(km & lt);
                    if (combined) {
                        pk[s] = (mi[q] << bits) | j;
                    } else {
                        pk[s] = j;
                        pv[s] = mi[q];
                    }
                }
                o += (uint32_t)__popcll(km);
            }
        }
    }
}

__global__ void k_bb_write(const uint32_t *__restrict__ misc, const uint32_t *__restrict__ ki,
                           const uint32_t *__restrict__ kj, int bits, bool combined, uint32_t *__restrict__ pairs) {
    const uint32_t np = misc[MISC_NP];
    const uint32_t mask = bits >= 32 ? 0xFFFFFFFFu : ((1u << bits) - 1u);
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < np; p += gridDim.x * blockDim.x) {
        uint32_t i, j;
        if (combined) {
            i = ki[p] >> bits;
            j = ki[p] & mask;
        } else {
            i = ki[p];
            j = kj[p];
        }
        pairs[2 * p] = i;
        pairs[2 * p + 1] = j;
    }
}

__global__ void k_bb_post(uint32_t *counters, uint32_t *misc, uint32_t max_pairs) {
    uint32_t t = counters[0];
    misc[MISC_NP] = t < max_pairs ? t : max_pairs;
    if (t > max_pairs) counters[2] |= SV_STATUS_PAIR_OVERFLOW;
}

// ---------------------------------------------------------------------------
// Broadphase by sweep and prune (extension beside physics.comp:451-472, the
// same pair list).  Boxes are ordered by a 16-bit key of their lower x bound
// (the top half of the order-preserving integer image of the float, -0
// taken as +0), which does not decrease with the bound: a box q after box p
// whose lower x bound is at most p's upper one has a key at most key(p's
// upper bound), and every box between them too.  So the boxes p must be
// tested against are a run after it, which ends where the keys pass
// key(upper x of p); candidates in the run get the exact test (the
// reference's: touching counts, static-static skipped), and each unordered
// pair is tested once, from its earlier box.  Pair (b_p, b_q) goes to row
// min(b_p, b_q); rows are counted, scanned, filled and sorted as in the
// LBVH broadphase (k_rows_short, k_rows_long).
//
// k_sap: one wave per 8 consecutive boxes of the order, 64 candidates a
// round, at most SAP_SPAN candidates past the first; a box whose run goes
// on finds where it ends (binary search over the keys) and is listed (box,
// next candidate, first item) for k_sap_long, which continues the listed
// runs in items of SAP_ITEM candidates, one wave per item.  One 64-bit
// atomic counts the entries (high word) and their items (low word), so the
// entries' first items increase with the entry.
constexpr uint32_t SAP_ROWS = 8;      // boxes per wave of k_sap
constexpr uint32_t SAP_SPAN = 4096;   // candidates k_sap takes after a workgroup's first box
constexpr uint32_t SAP_ITEM = 512;    // candidates per item of k_sap_long

// Keys come from a map set by the previous launch's lower x bounds: with
// scale > 0, key(x) = floor((x - x0) * scale) clamped to [0, 2^16 - 1]
// (rounded subtraction and multiplication by a positive constant do not
// decrease with x, nor does the floor; -0 and +0 give the same key), fine
// enough that boxes rarely share a key; else (the first launch) the top
// half of the order-preserving integer image of x (-0 taken as +0).  One map
// serves every kernel of a launch.
struct SapMap {
    float x0, scale;
};

SVD uint32_t sap_ord(float x) {
    const uint32_t u = __float_as_uint(x + 0.0f);  // -0 -> +0
    return u & 0x80000000u ? ~u : u | 0x80000000u;
}
SVD float sap_unord(uint32_t u) { return __uint_as_float(u & 0x80000000u ? u & 0x7FFFFFFFu : ~u); }

SVD uint32_t sap_key(float x, SapMap m) {
    if (m.scale > 0.0f) {
        const float t = (x - m.x0) * m.scale;
        return t <= 0.0f ? 0u : (t >= 65535.0f ? 65535u : (uint32_t)t);
    }
    return sap_ord(x) >> 16;
}

// The map for the next launch from this launch's lower-bound extremes
// (order-preserving images, reset for the next launch).  One thread, after
// the launch's last use of its map (the row sort's first kernel).
__device__ void sap_map_next(float *map) {
    uint32_t *ext = reinterpret_cast<uint32_t *>(map + 2);
    const uint32_t lo = ext[0], hi = ext[1];
    float x0 = 0.0f, scale = 0.0f;
    if (lo <= hi) {
        x0 = sap_unord(lo);
        const float span = sap_unord(hi) - x0;
        if (span > 0.0f && span < 3.0e38f) {
            scale = 65535.0f / span;
            if (!(scale < 3.0e38f)) scale = 0.0f;
        }
    }
    map[0] = x0;
    map[1] = scale;
    ext[0] = 0xFFFFFFFFu;
    ext[1] = 0u;
}

// The boxes are put in key order by counting: k_sap_keys counts the boxes
// of each of the 2^16 keys, a scan gives each key its first slot, and
// k_sap_order puts each box at its key's next slot, taken by counting the
// key's count back down (which leaves the counts at zero for the next
// launch).  The order within one key is of no account: the runs only need
// keys that do not decrease.
constexpr uint32_t SAP_KEYS = 1u << 16;

// Key counts, and the zeroed row counts, fill counts, list heads (long rows
// of the row sort, continued runs) and the flat list's tail.
__global__ void __launch_bounds__(BLOCK) k_sap_keys(uint32_t n, const float4 *__restrict__ box,
                                                    uint32_t *__restrict__ kcount, uint32_t *__restrict__ cnt,
                                                    uint32_t *__restrict__ fill, uint32_t *__restrict__ rows_list,
                                                    uint32_t *__restrict__ cont, uint32_t *__restrict__ flat_tail,
                                                    float *map) {
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i == 0) {
        rows_list[0] = 0u;
        cont[0] = 0u;  // the entry and item counter (64-bit)
        cont[1] = 0u;
        flat_tail[0] = 0u;
    }
    const SapMap m{map[0], map[1]};
    // this launch's lower-bound extremes (the next launch's map), one
    // atomic per wave each
    uint32_t omin = 0xFFFFFFFFu, omax = 0u;
    if (i < n) {
        const float x = box[2 * i].x;
        atomicAdd(kcount + sap_key(x, m), 1u);
        cnt[i] = 0u;
        fill[i] = 0u;
        omin = omax = sap_ord(x);
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        omin = min(omin, (uint32_t)__shfl_xor(omin, d, 64));
        omax = max(omax, (uint32_t)__shfl_xor(omax, d, 64));
    }
    __shared__ uint32_t s_ext[2][BLOCK / 64];
    if ((threadIdx.x & 63u) == 0) {
        s_ext[0][threadIdx.x >> 6] = omin;
        s_ext[1][threadIdx.x >> 6] = omax;
    }
    __syncthreads();
    if (threadIdx.x == 0) {  // one pair of atomics per workgroup
#pragma unroll
        for (int w = 1; w < BLOCK / 64; w++) {
            omin = min(omin, s_ext[0][w]);
            omax = max(omax, s_ext[1][w]);
        }
        if (omin <= omax) {
            uint32_t *ext = reinterpret_cast<uint32_t *>(map + 2);
            atomicMin(ext, omin);
            atomicMax(ext + 1, omax);
        }
    }
}

// The boxes in key order: lo.w = 1 if static, hi.w = the body index; and
// their keys.
__global__ void __launch_bounds__(BLOCK) k_sap_order(uint32_t n, const float4 *__restrict__ box,
                                                     uint32_t *kcount, const uint32_t *__restrict__ kfirst,
                                                     uint32_t *__restrict__ skey, float4 *__restrict__ sbox,
                                                     const float *__restrict__ map) {
    const uint32_t b = blockIdx.x * BLOCK + threadIdx.x;
    if (b >= n) return;
    const float4 lo = box[2 * b], hi = box[2 * b + 1];
    const uint32_t k = sap_key(lo.x, SapMap{map[0], map[1]});
    const uint32_t p = kfirst[k] + atomicSub(kcount + k, 1u) - 1u;
    skey[p] = k;
    sbox[2 * p] = make_float4(lo.x, lo.y, lo.z, lo.w != 0.0f ? 1.0f : 0.0f);
    sbox[2 * p + 1] = make_float4(hi.x, hi.y, hi.z, __uint_as_float(b));
}

// Pairs go out two ways.  The counting passes bump the row counts and stage
// each wave's pairs in LDS (SAP_STAGE at most), writing a run of them to the
// flat list at a slot taken with one atomic when the stage is nearly full
// and at the end (writes stop at the list's capacity; the count goes on).
// k_sap_place then puts the listed pairs in their rows.  Only when the count
// passed the capacity do the emit passes sweep again and place every pair
// in its row themselves (otherwise they leave at once).
constexpr uint32_t SAP_STAGE = 256;

struct SapOut {
    uint32_t *cnt;
    const uint32_t *off;
    uint32_t *fill;
    uint32_t *pk;
    uint32_t max_pairs;
    uint2 *flat;
    uint32_t flat_cap;
    uint32_t *flat_tail;
    const float *map;
};

SVD void sap_flush(const uint2 *stage, uint32_t &k, uint32_t lane, const SapOut &o) {
    __builtin_amdgcn_wave_barrier();  // the wave's stage writes are done
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(o.flat_tail, k);
    base = __shfl(base, 0, 64);
    for (uint32_t e = lane; e < k; e += 64)
        if (base + e < o.flat_cap) o.flat[base + e] = stage[e];
    __builtin_amdgcn_wave_barrier();  // the reads are done before the stage refills
    k = 0;
}

// The pairs of row box rb with the 64 candidates of a round (lane: body qb),
// m = the lanes with a pair.
template <bool EMIT>
SVD void sap_take(uint64_t m, uint32_t rb, uint32_t qb, uint32_t lane, uint64_t lt, uint2 *stage, uint32_t &k,
                  const SapOut &o) {
    const bool found = (m >> lane) & 1ull;
    const uint32_t i = rb < qb ? rb : qb, j = rb < qb ? qb : rb;
    if (EMIT) {
        bb_place(found, i, j, lt, o.off, o.fill, o.pk, o.max_pairs);
        return;
    }
    // the lanes whose pair is in the row box's own row: one atomic
    const uint64_t own = __ballot(found && rb < qb);
    if (own && (own & lt) == 0ull && found && rb < qb) atomicAdd(o.cnt + rb, (uint32_t)__popcll(own));
    if (found && qb < rb) atomicAdd(o.cnt + qb, 1u);
    if (found) stage[k + (uint32_t)__popcll(m & lt)] = make_uint2(i, j);
    k += (uint32_t)__popcll(m);
    if (k > SAP_STAGE - 64) sap_flush(stage, k, lane, o);
}

// k_sap: a workgroup takes SAP_ROWS * 4 consecutive boxes (a wave per
// SAP_ROWS) and streams the candidates after its first box through LDS, 256
// a tile (the next tile loaded while one is tested), until the tile's first
// key is past every row's limit or SAP_SPAN candidates; per (row, 64
// candidates) one lane mask, as in the all-pairs kernel; a row whose run
// ended before a chunk skips it.
template <bool EMIT>
__device__ __forceinline__ void sap_long_items(uint32_t n, const float4 *__restrict__ sbox,
                                               const uint32_t *__restrict__ skey, const SapOut &o,
                                               const uint32_t *__restrict__ cont, uint32_t block, uint32_t nblocks);

// The emit pass also sweeps the continuation items, in its workgroups from
// row_blocks on (emit mode lists no items, so the two parts are independent).
template <bool EMIT>
__global__ void __launch_bounds__(BLOCK) k_sap(uint32_t n, const float4 *__restrict__ sbox,
                                               const uint32_t *__restrict__ skey, SapOut o, uint32_t *cont,
                                               uint32_t row_blocks) {
    __shared__ float4 s_lo[BLOCK], s_hi[BLOCK];
    __shared__ uint32_t s_key[BLOCK], s_kmax[BLOCK / 64];
    __shared__ uint2 s_stage[BLOCK / 64][SAP_STAGE];
    if constexpr (EMIT) {
        if (*o.flat_tail <= o.flat_cap) return;  // the flat list holds every pair
        if (blockIdx.x >= row_blocks) {
            sap_long_items<true>(n, sbox, skey, o, cont, blockIdx.x - row_blocks, gridDim.x - row_blocks);
            return;
        }
    }
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t P0 = blockIdx.x * SAP_ROWS * (BLOCK / 64);
    if (P0 >= n) return;
    const uint32_t p0 = P0 + wave * SAP_ROWS;
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    uint2 *stage = s_stage[wave];
    uint32_t staged = 0;
    const SapMap km{o.map[0], o.map[1]};
    float4 rlo[SAP_ROWS], rhi[SAP_ROWS];
    uint32_t klim[SAP_ROWS];
    uint32_t kmax = 0, rstatic = 0;
#pragma unroll
    for (uint32_t r = 0; r < SAP_ROWS; r++) {
        const uint32_t p = p0 + r < n ? p0 + r : n - 1;
        rlo[r] = sbox[2 * p];
        rhi[r] = sbox[2 * p + 1];
        klim[r] = sap_key(rhi[r].x, km);
        if (p0 + r < n) kmax = klim[r] > kmax ? klim[r] : kmax;
        rstatic |= (rlo[r].w != 0.0f ? 1u : 0u) << r;
        asm volatile("" : "+v"(rlo[r].x), "+v"(rlo[r].y), "+v"(rlo[r].z), "+v"(rhi[r].x), "+v"(rhi[r].y),
                     "+v"(rhi[r].z));
    }
    rstatic = __builtin_amdgcn_readfirstlane(rstatic);  // wave-uniform: scalar masks and branches
    kmax = __builtin_amdgcn_readfirstlane(kmax);
    if (lane == 0) s_kmax[wave] = kmax;
    __syncthreads();
    uint32_t kblk = 0;
#pragma unroll
    for (uint32_t w = 0; w < BLOCK / 64; w++) kblk = s_kmax[w] > kblk ? s_kmax[w] : kblk;
    const uint32_t qend = min(n, P0 + 1 + SAP_SPAN);
    uint32_t t0 = P0 + 1;
    float4 nlo, nhi;
    uint32_t nkey;
    {
        const uint32_t q = t0 + threadIdx.x, qc = q < n ? q : n - 1;
        nlo = sbox[2 * qc];
        nhi = sbox[2 * qc + 1];
        nkey = q < qend ? skey[qc] : 0xFFFFFFFFu;  // past the span: above every limit
    }
    for (; t0 < qend; t0 += BLOCK) {
        __syncthreads();  // the previous tile's reads are done
        s_lo[threadIdx.x] = nlo;
        s_hi[threadIdx.x] = nhi;
        s_key[threadIdx.x] = nkey;
        __syncthreads();
        if (s_key[0] > kblk) break;  // keys do not decrease: every run has ended
        if (t0 + BLOCK < qend) {
            const uint32_t q = t0 + BLOCK + threadIdx.x, qc = q < n ? q : n - 1;
            nlo = sbox[2 * qc];
            nhi = sbox[2 * qc + 1];
            nkey = q < qend ? skey[qc] : 0xFFFFFFFFu;
        }
#pragma unroll
        for (uint32_t ch = 0; ch < BLOCK / 64; ch++) {
            const uint32_t qi = ch * 64 + lane, qq = t0 + qi;
            const uint32_t kfirst = s_key[ch * 64];  // the chunk's smallest key
            if (kfirst > kmax) break;               // this wave's runs have all ended
            const float4 lo = s_lo[qi], hi = s_hi[qi];
            const uint32_t kq = s_key[qi], qb = __float_as_uint(hi.w);
            const uint64_t sq = __builtin_amdgcn_ballot_w64(lo.w != 0.0f);
#pragma unroll
            for (uint32_t r = 0; r < SAP_ROWS; r++) {
                const uint32_t kl = __builtin_amdgcn_readfirstlane(klim[r]);  // wave-uniform: a scalar skip
                if (p0 + r >= n || kfirst > kl) continue;  // row r's run ended before this chunk
                // one predicate, no short circuit (each && was a branch)
                const bool hit = (kq <= kl) & (qq > p0 + r) & !(boxes_gap(rlo[r], rhi[r], lo, hi) > 0.0f);
                uint64_t m = __builtin_amdgcn_ballot_w64(hit);
                if ((rstatic >> r) & 1u) m &= ~sq;
                if (m) sap_take<EMIT>(m, __builtin_amdgcn_readfirstlane(__float_as_uint(rhi[r].w)), qb, lane, lt,
                                      stage, staged, o);
            }
        }
    }
    if (EMIT) return;
    if (staged) sap_flush(stage, staged, lane, o);
    // runs that go past the span: continued by k_sap_long
    if (t0 >= qend && qend < n && lane < SAP_ROWS && p0 + lane < n) {
        uint32_t kl = 0;
#pragma unroll
        for (uint32_t r = 0; r < SAP_ROWS; r++)
            if (r == lane) kl = klim[r];
        if (skey[qend] <= kl) {
            uint32_t lo = qend + 1, hi = n;  // the run ends at the first key above kl in (qend, n]
            while (lo < hi) {
                const uint32_t mid = lo + (hi - lo) / 2;
                if (skey[mid] <= kl) lo = mid + 1;
                else hi = mid;
            }
            const uint32_t items = (lo - qend + SAP_ITEM - 1) / SAP_ITEM;
            const unsigned long long old =
                atomicAdd(reinterpret_cast<unsigned long long *>(cont), (1ull << 32) | items);
            const uint32_t e = (uint32_t)(old >> 32);
            cont[2 + 3 * e] = p0 + lane;
            cont[3 + 3 * e] = qend;
            cont[4 + 3 * e] = (uint32_t)old;  // its first item
        }
    }
}

// The flat list's pairs to their rows (when it held them all).
__global__ void __launch_bounds__(BLOCK) k_sap_place(const uint2 *__restrict__ flat, const uint32_t *flat_tail,
                                                     uint32_t flat_cap, const uint32_t *__restrict__ off,
                                                     uint32_t *fill, uint32_t *__restrict__ pk, uint32_t max_pairs) {
    const uint32_t total = *flat_tail;
    if (total > flat_cap) return;  // the emit passes place them
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    for (uint32_t e0 = blockIdx.x * BLOCK; e0 < total; e0 += gridDim.x * BLOCK) {
        const uint32_t e = e0 + threadIdx.x;
        const uint2 pr = e < total ? flat[e] : make_uint2(0u, 0u);
        bb_place(e < total, pr.x, pr.y, lt, off, fill, pk, max_pairs);
    }
}

// The listed runs, one wave per item of SAP_ITEM candidates: item g belongs
// to the last entry whose first item is <= g (binary search).
// Workgroup `block` of `nblocks` sweeping continuation items (k_sap_long,
// and the emit pass's workgroups past its row blocks).
template <bool EMIT>
__device__ __forceinline__ void sap_long_items(uint32_t n, const float4 *__restrict__ sbox,
                                               const uint32_t *__restrict__ skey, const SapOut &o,
                                               const uint32_t *__restrict__ cont, uint32_t block, uint32_t nblocks) {
    __shared__ uint2 s_stage[BLOCK / 64][SAP_STAGE];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t gw = block * (BLOCK / 64) + wave, nw = nblocks * (BLOCK / 64);
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    uint2 *stage = s_stage[wave];
    uint32_t staged = 0;
    const uint32_t items = cont[0], m = cont[1];
    for (uint32_t g = gw; g < items; g += nw) {
        uint32_t lo = 0, hi = m;  // last entry with first item <= g, in [0, m)
        while (hi - lo > 1) {
            const uint32_t mid = lo + (hi - lo) / 2;
            if (cont[4 + 3 * mid] <= g) lo = mid;
            else hi = mid;
        }
        const uint32_t p = cont[2 + 3 * lo], qbeg = cont[3 + 3 * lo] + (g - cont[4 + 3 * lo]) * SAP_ITEM;
        const uint32_t qend = min(n, qbeg + SAP_ITEM);
        const float4 rlo = sbox[2 * p], rhi = sbox[2 * p + 1];
        const uint32_t klim = sap_key(rhi.x, SapMap{o.map[0], o.map[1]}), rb = __float_as_uint(rhi.w);
        uint32_t qn = min(qbeg + lane, n - 1);
        float4 nlo = sbox[2 * qn], nhi = sbox[2 * qn + 1];
        uint32_t nkq = skey[qn];
        for (uint32_t q0 = qbeg; q0 < qend; q0 += 64) {
            const float4 lo4 = nlo, hi4 = nhi;
            const uint32_t kq = nkq;
            if (q0 + 64 < qend) {
                qn = min(q0 + 64 + lane, n - 1);
                nlo = sbox[2 * qn];
                nhi = sbox[2 * qn + 1];
                nkq = skey[qn];
            }
            if (__builtin_amdgcn_readfirstlane(kq) > klim) break;
            const bool found = q0 + lane < qend && kq <= klim && !(lo4.w != 0.0f && rlo.w != 0.0f) &&
                               !boxes_apart(rlo, rhi, lo4, hi4);
            const uint64_t mm = __ballot(found);
            if (mm) sap_take<EMIT>(mm, rb, __float_as_uint(hi4.w), lane, lt, stage, staged, o);
        }
    }
    if (!EMIT && staged) sap_flush(stage, staged, lane, o);
}

__global__ void __launch_bounds__(BLOCK) k_sap_long(uint32_t n, const float4 *__restrict__ sbox,
                                                    const uint32_t *__restrict__ skey, SapOut o,
                                                    const uint32_t *__restrict__ cont) {
    sap_long_items<false>(n, sbox, skey, o, cont, blockIdx.x, gridDim.x);
}

}  // namespace

namespace {

// Rows to pairs: k_rows_short over every row, then k_rows_long over the
// rows it listed (list[0] must be zero before).
void launch_sort_rows(hipStream_t st, uint32_t n, const uint32_t *cnt, const uint32_t *off, const uint32_t *pk,
                      uint32_t *pairs, uint32_t max_pairs, uint32_t *list, uint32_t window_words = 0,
                      float *map = nullptr) {
    uint32_t ww = div_up(div_up(n, 32u), (uint32_t)BLOCK) * BLOCK;  // window words, a multiple of BLOCK
    if (ww > ROW_WINDOW_WORDS) ww = ROW_WINDOW_WORDS;
    if (window_words) ww = window_words;
    const uint32_t lds_bytes = (ww > ROW_RANK_MAX ? ww : ROW_RANK_MAX) * (uint32_t)sizeof(uint32_t);
    static bool lds_set = false;
    if (!lds_set) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(k_rows_long),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, ROW_WINDOW_WORDS * sizeof(uint32_t));
        lds_set = true;
    }
    hipLaunchKernelGGL(k_rows_short, dim3(div_up(n, BLOCK)), dim3(BLOCK), 0, st, n, cnt, off, pk, pairs, max_pairs,
                       list, map);
    hipLaunchKernelGGL(k_rows_long, dim3(256), dim3(BLOCK), lds_bytes, st, n, cnt, off, pk, pairs, max_pairs,
                       (const uint32_t *)list, ww);
}

template <typename T>
hipError_t lazy_alloc(sv_ctx *c, T *&p, size_t count) {
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, (count ? count : 1) * sizeof(T));
    if (e == hipSuccess) {
        c->s.lazy.push_back(q);
        p = static_cast<T *>(q);
    }
    return e;
}

bool bb_reserve(sv_ctx *c) {
    if (c->s.bb_ready) return true;
    Scratch &s = c->s;
    size_t N = c->caps.max_bodies, P = c->caps.max_pairs;
    size_t nodes = 2 * N + 1, parts = div_up(N, BLOCK * BOUNDS_ITEMS) + 1;
    hipError_t e = lazy_alloc(c, s.bb_lo, nodes);
    if (e == hipSuccess) e = lazy_alloc(c, s.bb_hi, nodes);
    if (e == hipSuccess) e = lazy_alloc(c, s.bb_rec, 4 * N + 4);
    if (e == hipSuccess) e = lazy_alloc(c, s.bb_left, nodes);
    if (e == hipSuccess) e = lazy_alloc(c, s.bb_right, nodes);
    if (e == hipSuccess) e = lazy_alloc(c, s.bb_parent, nodes);
    if (e == hipSuccess) e = lazy_alloc(c, s.bb_count, nodes);
    if (e == hipSuccess) e = lazy_alloc(c, s.bb_flags, N + 1);
    if (e == hipSuccess) e = lazy_alloc(c, s.bb_cent, 3 * N + 3);
    if (e == hipSuccess) e = lazy_alloc(c, s.bb_bounds, 6 + 6 * parts);
    for (int q = 0; q < 2 && e == hipSuccess; q++) {
        e = lazy_alloc(c, s.bb_mk[q], N + 1);
        if (e == hipSuccess) e = lazy_alloc(c, s.bb_mv[q], N + 1);
        if (e == hipSuccess) e = lazy_alloc(c, s.bb_pk[q], P + N + 1);
        if (e == hipSuccess) e = lazy_alloc(c, s.bb_pv[q], P + N + 1);
    }
    if (e == hipSuccess) e = lazy_alloc(c, s.bb_cnt, N + 1);
    if (e == hipSuccess) e = lazy_alloc(c, s.bb_off, N + 1);
    if (e == hipSuccess) e = lazy_alloc(c, s.bb_fill, N + 1);
    if (e == hipSuccess) e = lazy_alloc(c, s.bb_stage, (size_t)BB_STAGE * N + 1);
    if (e == hipSuccess) e = lazy_alloc(c, s.bb_scnt, N + 1);
    if (e == hipSuccess) e = lazy_alloc(c, s.bb_rng, N + 1);
    if (e == hipSuccess) e = lazy_alloc(c, s.bb_tops, refit_list_words(N));
    s.bb_ready = e == hipSuccess;
    return s.bb_ready;
}

bool sap_reserve(sv_ctx *c) {
    if (c->s.sap_ready) return true;
    Scratch &s = c->s;
    const size_t N = c->caps.max_bodies, P = c->caps.max_pairs;
    hipError_t e = hipSuccess;
    e = lazy_alloc(c, s.sap_kcount, SAP_KEYS);
    if (e == hipSuccess) e = zero_device(s.sap_kcount, SAP_KEYS * sizeof(uint32_t));  // zero between launches
    if (e == hipSuccess) e = lazy_alloc(c, s.sap_kfirst, SAP_KEYS);
    if (e == hipSuccess) e = lazy_alloc(c, s.sap_map, 4);
    if (e == hipSuccess) {  // no map yet (x0 = scale = 0), extremes empty
        const uint32_t init[4] = {0u, 0u, 0xFFFFFFFFu, 0u};
        e = hipMemcpy(s.sap_map, init, sizeof(init), hipMemcpyHostToDevice);
    }
    if (e == hipSuccess) e = lazy_alloc(c, s.sap_skey, N + 1);
    if (e == hipSuccess) e = lazy_alloc(c, s.sap_box, 2 * N + 2);
    if (e == hipSuccess) e = lazy_alloc(c, s.sap_cnt, N + 1);
    if (e == hipSuccess) e = lazy_alloc(c, s.sap_off, N + 1);
    if (e == hipSuccess) e = lazy_alloc(c, s.sap_fill, N + 1);
    if (e == hipSuccess) e = lazy_alloc(c, s.sap_part, P + N + 1);
    if (e == hipSuccess) e = lazy_alloc(c, s.sap_rows, N + 2);
    if (e == hipSuccess) e = lazy_alloc(c, s.sap_cont, 3 * N + 4);
    if (e == hipSuccess) e = lazy_alloc(c, s.sap_flat, P + N + 1);
    if (e == hipSuccess) e = lazy_alloc(c, s.sap_tail, 4);
    s.sap_ready = e == hipSuccess;
    return s.sap_ready;
}

}  // namespace

bool launch_broadphase_sap(sv_ctx *c, uint32_t n) {
    uint32_t *counters = B<uint32_t>(c, SV_ATOMIC_COUNTERS);
    uint32_t *pairs = B<uint32_t>(c, SV_COLLISION_PAIRS);
    const uint32_t mp = c->caps.max_pairs;
    hipStream_t st = c->stream;
    if (n > 0x7FFFFFFFu || !sap_reserve(c)) return false;
    if (n < 2) {
        fill_u32(c, counters, 0u, 1);
        hipLaunchKernelGGL(k_bb_post, dim3(1), dim3(1), 0, st, counters, c->s.misc, mp);
        return true;
    }
    Scratch &s = c->s;
    const uint32_t g = div_up(n, BLOCK);
    hipLaunchKernelGGL(k_sap_keys, dim3(g), dim3(BLOCK), 0, st, n, s.bpbox, s.sap_kcount, s.sap_cnt, s.sap_fill,
                       s.sap_rows, s.sap_cont, s.sap_tail, s.sap_map);
    scan_exclusive(c, s.sap_kcount, s.sap_kfirst, nullptr, SAP_KEYS, nullptr);
    hipLaunchKernelGGL(k_sap_order, dim3(g), dim3(BLOCK), 0, st, n, s.bpbox, s.sap_kcount,
                       (const uint32_t *)s.sap_kfirst, s.sap_skey, s.sap_box, (const float *)s.sap_map);
    const uint32_t *skey = s.sap_skey;
    const uint32_t gs = div_up(n, SAP_ROWS * (uint32_t)(BLOCK / 64));  // SAP_ROWS boxes per wave
    const uint32_t gl = 128;                                          // k_sap_long workgroups
    const uint32_t flat_cap = mp + n;
    const SapOut out{s.sap_cnt, s.sap_off, s.sap_fill, s.sap_part, mp, s.sap_flat, flat_cap, s.sap_tail, s.sap_map};
    // count the rows and list the pairs, scan the rows (with the pair count
    // and the overflow flag), place the listed pairs in their rows (or sweep
    // again to place them when the list was too short), sort the rows
    hipLaunchKernelGGL(k_sap<false>, dim3(gs), dim3(BLOCK), 0, st, n, s.sap_box, skey, out, s.sap_cont, gs);
    hipLaunchKernelGGL(k_sap_long, dim3(gl), dim3(BLOCK), 0, st, n, s.sap_box, skey, out,
                       (const uint32_t *)s.sap_cont);
    scan_exclusive(c, s.sap_cnt, s.sap_off, nullptr, n, counters, mp, s.misc + MISC_NP, counters + 2);
    hipLaunchKernelGGL(k_sap_place, dim3(256), dim3(BLOCK), 0, st, (const uint2 *)s.sap_flat,
                       (const uint32_t *)s.sap_tail, flat_cap, s.sap_off, s.sap_fill, s.sap_part, mp);
    hipLaunchKernelGGL(k_sap<true>, dim3(gs + gl), dim3(BLOCK), 0, st, n, s.sap_box, skey, out, s.sap_cont, gs);
    // (the row sort's first kernel also sets the map for the next launch)
    launch_sort_rows(st, n, s.sap_cnt, s.sap_off, s.sap_part, pairs, mp, s.sap_rows, 0, s.sap_map);
    return true;
}

void debug_sort_rows(sv_ctx *c, uint32_t n, const uint32_t *cnt, const uint32_t *off, const uint32_t *pk,
                     uint32_t *pairs, uint32_t max_pairs, uint32_t *list, uint32_t window_words) {
    launch_sort_rows(c->stream, n, cnt, off, pk, pairs, max_pairs, list, window_words);
}

bool launch_broadphase_lbvh(sv_ctx *c, uint32_t n) {
    uint32_t *counters = B<uint32_t>(c, SV_ATOMIC_COUNTERS);
    uint32_t *pairs = B<uint32_t>(c, SV_COLLISION_PAIRS);
    const uint32_t mp = c->caps.max_pairs;
    hipStream_t st = c->stream;
    if (n > 0x7FFFFFFFu || !bb_reserve(c)) return false;  // the walk's records keep body indices in 31 bits
    if (n < 2) {
        fill_u32(c, counters, 0u, 1);
        hipLaunchKernelGGL(k_bb_post, dim3(1), dim3(1), 0, st, counters, c->s.misc, mp);
        return true;
    }
    Scratch &s = c->s;
    const uint32_t g = div_up(n, BLOCK);
    // tree over the broadphase boxes
    const uint32_t parts = div_up(n, BOUNDS_TILE);
    hipLaunchKernelGGL(k_bb_centroids, dim3(parts), dim3(BLOCK), 0, st, n, s.bpbox, s.bb_cent, s.bb_mv[0],
                       s.bb_bounds + 6, s.bb_bounds, s.misc + MISC_BOUNDS_DONE, s.bb_cnt, s.bb_fill);
    hipLaunchKernelGGL(k_morton, dim3(g), dim3(BLOCK), 0, st, n, s.bb_bounds, s.bb_cent, s.bb_mk[0]);
    int w = radix_sort(c, s.bb_mk[0], s.bb_mv[0], s.bb_mk[1], s.bb_mv[1], nullptr, n, 30);
    const uint32_t *ids = s.bb_mv[w];
    // with the leaf boxes and the refit's cleared arrival counters
    hipLaunchKernelGGL(k_hierarchy, dim3(g), dim3(BLOCK), 0, st, n, s.bb_mk[w], s.bb_left, s.bb_right, s.bb_count,
                       s.bb_parent, s.bb_rng, ids, s.bpbox, s.bb_lo, s.bb_hi, s.bb_flags);
    // the refit also writes each internal node's record of its children's boxes (the walk's)
    launch_refit<float4, Box4>(st, n, s.bb_parent, s.bb_left, s.bb_right, (const uint2 *)s.bb_rng, s.bb_flags,
                               reinterpret_cast<float4 *>(s.bb_lo), reinterpret_cast<float4 *>(s.bb_hi), s.bb_tops,
                               s.bb_rec);
    // count, scan, fill, sort
    const int bits = 32 - __builtin_clz(n - 1);  // body indices < 2^bits
    const bool combined = 2 * bits <= 32;
    hipLaunchKernelGGL(k_bb_pairs<false>, dim3(g), dim3(BLOCK), 0, st, n, ids, s.bpbox, s.bb_rec, s.bb_cnt, (const uint32_t *)nullptr, (uint32_t *)nullptr, (uint32_t *)nullptr,
                       (uint32_t *)nullptr, mp, bits, combined, s.bb_stage, s.bb_scnt);
    scan_exclusive(c, s.bb_cnt, s.bb_off, nullptr, n, counters);
    hipLaunchKernelGGL(k_bb_scatter, dim3(div_up(n, BLOCK * BB_SCATTER_ITEMS)), dim3(BLOCK), 0, st, n, ids, s.bb_stage, s.bb_scnt, s.bb_off, counters,
                       s.bb_fill, s.bb_pk[0], s.bb_pv[0], mp, bits, combined, s.bb_cnt, s.misc);
    hipLaunchKernelGGL(k_bb_pairs<true>, dim3(g), dim3(BLOCK), 0, st, n, ids, s.bpbox, s.bb_rec, (uint32_t *)nullptr, s.bb_off, s.bb_fill, s.bb_pk[0], s.bb_pv[0], mp, bits,
                       combined, (uint32_t *)nullptr, s.bb_scnt);
    const uint32_t cap = mp + n;
    const uint32_t *ki, *kj;
    if (combined) {
        int r = radix_sort(c, s.bb_pk[0], s.bb_pv[0], s.bb_pk[1], s.bb_pv[1], s.misc + MISC_BPE, cap, 2 * bits);
        ki = s.bb_pk[r];
        kj = nullptr;
    } else {
        // stable by j, then stable by i: lexicographic (i, j)
        int r = radix_sort(c, s.bb_pk[0], s.bb_pv[0], s.bb_pk[1], s.bb_pv[1], s.misc + MISC_BPE, cap, bits);
        int r2 = radix_sort(c, s.bb_pv[r], s.bb_pk[r], s.bb_pk[1 - r], s.bb_pv[1 - r], s.misc + MISC_BPE, cap, bits);
        ki = r2 == 0 ? s.bb_pv[r] : s.bb_pk[1 - r];
        kj = r2 == 0 ? s.bb_pk[r] : s.bb_pv[1 - r];
    }
    uint32_t wg = div_up(mp, BLOCK);
    if (wg > 4096) wg = 4096;
    if (wg == 0) wg = 1;
    hipLaunchKernelGGL(k_bb_write, dim3(wg), dim3(BLOCK), 0, st, s.misc, ki, kj, bits, combined, pairs);
    return true;
}

void bvh_stage(sv_ctx *c, uint32_t step, uint32_t n, bool bodies) {
    if (n == 0) return;
    uint32_t g = div_up(n, BLOCK);
    hipStream_t s = c->stream;
    float *cent = B<float>(c, SV_BVH_CENTROIDS);
    uint32_t *morton = B<uint32_t>(c, SV_BVH_MORTON), *sidx = B<uint32_t>(c, SV_BVH_SORT_INDICES);
    uint32_t *prim = B<uint32_t>(c, SV_BVH_PRIM_INDICES);
    float *bounds = B<float>(c, SV_BVH_SCENE_BOUNDS);
    switch (step) {
    case 0:
        if (bodies)
            hipLaunchKernelGGL(k_centroids_bodies, dim3(g), dim3(BLOCK), 0, s, n, B<float>(c, SV_BODY_AABB_MIN),
                               B<float>(c, SV_BODY_AABB_MAX), cent, sidx, prim);
        else
            hipLaunchKernelGGL(k_centroids, dim3(g), dim3(BLOCK), 0, s, n, B<float>(c, SV_VERTICES),
                               B<uint32_t>(c, SV_INDICES), cent, sidx, prim);
        break;
    case 1: {
        uint32_t parts = div_up(n, BLOCK * BOUNDS_ITEMS);
        float *part = B<float>(c, SV_BVH_TEMP);
        hipLaunchKernelGGL(k_bounds_partial, dim3(parts), dim3(BLOCK), 0, s, n, cent, part);
        hipLaunchKernelGGL(k_bounds_final, dim3(1), dim3(BLOCK), 0, s, parts, part, bounds);
        break;
    }
    case 2:
        hipLaunchKernelGGL(k_morton, dim3(g), dim3(BLOCK), 0, s, n, bounds, cent, morton);
        break;
    case 3: {
        uint32_t *mt = B<uint32_t>(c, SV_BVH_MORTON_TEMP), *st = B<uint32_t>(c, SV_BVH_SORT_TEMP);
        int where = radix_sort(c, morton, sidx, mt, st, nullptr, n, 32);
        if (where == 1) {
            hipLaunchKernelGGL(k_copy_u32, dim3(g), dim3(BLOCK), 0, s, n, mt, morton);
            hipLaunchKernelGGL(k_copy_u32, dim3(g), dim3(BLOCK), 0, s, n, st, sidx);
        }
        break;
    }
    case 4: {
        // leaf ranges for the block-local refit (none with fixed scratch:
        // step 7 then climbs with device-coherent hand-offs only)
        Scratch &sc = c->s;
        sc.bvh_rng_n = 0;
        if (!c->scratch_fixed && sc.bvh_rng_cap < n) {
            void *q = nullptr, *r = nullptr;
            if (hipMalloc(&q, (size_t)n * sizeof(uint2)) == hipSuccess) sc.lazy.push_back(q);
            else q = nullptr;
            if (q && hipMalloc(&r, refit_list_words(n) * sizeof(uint32_t)) == hipSuccess) sc.lazy.push_back(r);
            else r = nullptr;
            if (q && r) {
                sc.bvh_rng = static_cast<uint2 *>(q);
                sc.bvh_tops = static_cast<uint32_t *>(r);
                sc.bvh_rng_cap = n;
            }
        }
        uint2 *rng = (!c->scratch_fixed && sc.bvh_rng_cap >= n) ? sc.bvh_rng : nullptr;
        hipLaunchKernelGGL(k_hierarchy, dim3(g), dim3(BLOCK), 0, s, n, morton, B<uint32_t>(c, SV_BVH_LEFT),
                           B<uint32_t>(c, SV_BVH_RIGHT), B<uint32_t>(c, SV_BVH_COUNT), B<uint32_t>(c, SV_BVH_PARENT),
                           rng);
        if (rng) sc.bvh_rng_n = n;
        break;
    }
    case 5:
        if (bodies)
            hipLaunchKernelGGL(k_leaf_bodies, dim3(g), dim3(BLOCK), 0, s, n, sidx, B<float>(c, SV_BODY_AABB_MIN),
                               B<float>(c, SV_BODY_AABB_MAX), prim, B<float>(c, SV_BVH_AABB_MIN),
                               B<float>(c, SV_BVH_AABB_MAX));
        else
            hipLaunchKernelGGL(k_leaf_tris, dim3(g), dim3(BLOCK), 0, s, n, sidx, B<float>(c, SV_VERTICES),
                               B<uint32_t>(c, SV_INDICES), prim, B<float>(c, SV_BVH_AABB_MIN),
                               B<float>(c, SV_BVH_AABB_MAX));
        break;
    case 6:
        hipLaunchKernelGGL(k_clear_flags, dim3(g), dim3(BLOCK), 0, s, n, B<uint32_t>(c, SV_BVH_FLAGS));
        break;
    case 7:  // needs step 6's cleared arrival counters
        if (c->s.bvh_rng && c->s.bvh_rng_n == n)  // ranges of this tree (step 4 on this context)
            launch_refit<float, Box3>(s, n, B<uint32_t>(c, SV_BVH_PARENT), B<uint32_t>(c, SV_BVH_LEFT),
                                      B<uint32_t>(c, SV_BVH_RIGHT), (const uint2 *)c->s.bvh_rng,
                                      B<uint32_t>(c, SV_BVH_FLAGS), B<float>(c, SV_BVH_AABB_MIN),
                                      B<float>(c, SV_BVH_AABB_MAX), c->s.bvh_tops);
        else
            hipLaunchKernelGGL(k_propagate, dim3(g), dim3(BLOCK), 0, s, n, B<uint32_t>(c, SV_BVH_PARENT),
                               B<uint32_t>(c, SV_BVH_LEFT), B<uint32_t>(c, SV_BVH_RIGHT),
                               B<uint32_t>(c, SV_BVH_FLAGS), B<float>(c, SV_BVH_AABB_MIN),
                               B<float>(c, SV_BVH_AABB_MAX));
        c->s.bvh_rng_n = 0;  // ranges serve the refit right after their step 4 only
        break;
    default:
        break;  // 8..11: folded into the stages above
    }
}

}  // namespace sv
