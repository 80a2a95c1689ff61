// sv_pgs.hip - projected Gauss-Seidel of physics.comp:500-583 in the
// reference's sequential contact order, bit for bit, on many lanes.
//
// The serial sweep visits items (iteration t, contact c) in order.  Item
// (t, c) reads and writes the velocities of its two bodies only; a body with
// zero inverse mass and zero inverse inertia is never changed by an update
// (v - 0*imp == v for finite imp), so it is read-only and orders nothing.
// Two items therefore commute unless they share a dynamic body, and running
// every item after its predecessors on each of its dynamic bodies yields
// the exact serial result.  Each body keeps a ticket counter of completed
// updates; item (t, c) may run when ticket[a] == t*deg(a) + rank_a(c) (and
// the same for b), where deg(a) counts a's non-skipped contacts and
// rank_a(c) is c's position among them.  rank/deg come from a stable radix
// sort of the (body, contact) incidences.
//
// Scheduling: every lane holds one item at a time and, as soon as it has
// applied it, takes the next unassigned item in sweep order from a shared
// counter (one atomic per wave per refill).  An item's constants (contact
// frame, effective mass, bias, both bodies' rotations and inverse inertia)
// are loaded when it is taken, so the dependent part of an item is the
// ticket check, the two bodies' velocities, the update and the release.
// The earliest unfinished item is always held by a running lane and has all
// its predecessors done, so the schedule cannot deadlock.  A wave that makes
// no progress for 2 s (a broken invariant) raises an abort flag that ends
// the kernel, and the host reports SV_E_TIMEOUT.
//
// Forms: one workgroup of 1024 lanes with tickets in LDS (the hand-off stays
// inside one CU: workgroup-scope release/acquire), the same with tickets in
// global memory when they do not fit in LDS, and multi-workgroup forms.  In
// the default multi-workgroup form (FORM_COH) the velocities, angular
// velocities and impulses are read and written with device-scope relaxed
// atomics, which go to the device's point of coherence, and a hand-off is a
// wait for the item's stores followed by the ticket bump; the other
// (FORM_MULTI) uses plain accesses with device-scope release/acquire fences,
// which write back and invalidate the XCD's L2 on every hand-off (1.65x
// slower on a 32k-body pile).
#include <type_traits>

#include <vector>

#include "sv_internal.h"
#include "sv_math.h"
#include "sv_dev.h"

namespace sv {

namespace {

constexpr int BLOCK = 256;
constexpr int SOLVE_THREADS = 1024;
constexpr int MC_THREADS = 256;
constexpr int SEQ_THREADS = 64;  // sequence form: one wave per workgroup
constexpr uint32_t SEQ_BACKOFF_MAX = 8;  // pipelined form: most turns a lane skips polling
constexpr uint64_t STALL_TICKS = 200000000ull;  // 2 s of the 100 MHz wall clock
constexpr uint64_t ABORT_CHECK_TICKS = 2000ull;  // 20 us

template <typename T>
inline T *B(sv_ctx *c, int id) { return reinterpret_cast<T *>(c->buf[id]); }

// The sweep's setup in one launch: contacts to solve, min(counters[1],
// max_contacts) (physics.comp:502), the incidence count, the cleared count
// of swept contacts, per-body degrees and the sweep's control words.
__global__ void __launch_bounds__(256) k_sweep_setup(const uint32_t *counters, uint32_t *misc, uint32_t max_contacts,
                                                     uint32_t *deg, uint32_t n_bodies, uint32_t *ctl) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_bodies) deg[i] = 0u;
    if (i == 0) {
        const uint32_t t = counters[1];
        const uint32_t nc = t < max_contacts ? t : max_contacts;
        misc[MISC_NC] = nc;
        misc[MISC_NINC] = 2u * nc;
        misc[MISC_NACT] = 0u;
    }
    if (blockIdx.x == 0)
        for (uint32_t k = threadIdx.x; k < SOLVE_CTL_WORDS; k += blockDim.x) ctl[k] = 0u;
}

// Per-contact constants of the sweep (contact_prep), incidence keys and
// per-body degrees.
__global__ void __launch_bounds__(BLOCK) k_prep(
    uint32_t *__restrict__ misc_w, uint32_t n_bodies, float dt, float baumgarte, float slop,
    const uint32_t *__restrict__ cga, const uint32_t *__restrict__ cgb, const uint32_t *__restrict__ gbody,
    const float *__restrict__ cpos, const float *__restrict__ cnrm, const float *__restrict__ cpen,
    const float *__restrict__ pos, const float *__restrict__ vel, const float *__restrict__ omega,
    const float *__restrict__ inv_mass, const float *__restrict__ inv_i, const float *__restrict__ rot,
    const float *__restrict__ fric, const float *__restrict__ rest, float *__restrict__ prep,
    uint32_t *__restrict__ key, uint32_t *__restrict__ val, uint32_t *__restrict__ deg,
    const float *__restrict__ lambda, uint32_t warm) {
    uint32_t nc = misc_w[MISC_NC];
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += gridDim.x * blockDim.x) {
        uint32_t ga = cga[c], gb = cgb[c];
        uint32_t a = gbody[ga], b = gbody[gb];
        uint32_t flags = contact_prep(a, b, ga, gb, ld3(cnrm, c), ld3(cpos, c), cpen[c], dt, baumgarte, slop, pos,
                                      vel, omega, inv_mass, inv_i, rot, fric, rest, prep + (size_t)c * CPREP_WORDS);
        bool skip = flags & F_SKIP, dyn_a = flags & F_DYN_A, dyn_b = flags & F_DYN_B;
        if (warm && !skip && lambda[c] > 0.0f)
            prep[(size_t)c * CPREP_WORDS + P_FLAGS] = __uint_as_float(flags | F_WARM);
        {  // one atomic per wave for the count of swept contacts
            const uint64_t m = __ballot(!skip);
            if (m && (threadIdx.x & 63) == (uint32_t)__builtin_ctzll(m))
                atomicAdd(&misc_w[MISC_NACT], (uint32_t)__popcll(m));
        }
        bool ia = !skip && dyn_a, ib = !skip && dyn_b;
        key[2 * c] = ia ? a : n_bodies;
        val[2 * c] = 2 * c;
        key[2 * c + 1] = ib ? b : n_bodies;
        val[2 * c + 1] = 2 * c + 1;
        if (ia) atomicAdd(&deg[a], 1u);
        if (ib) atomicAdd(&deg[b], 1u);
    }
}

// rank of each incidence inside its body's segment of the sorted list.
__global__ void __launch_bounds__(BLOCK) k_rank(const uint32_t *__restrict__ misc, uint32_t n_bodies,
                                                const uint32_t *__restrict__ key, const uint32_t *__restrict__ val,
                                                const uint32_t *__restrict__ deg_off, uint32_t *__restrict__ rank) {
    uint32_t ninc = misc[MISC_NINC];
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < ninc; p += gridDim.x * blockDim.x) {
        uint32_t k = key[p];
        if (k < n_bodies) rank[val[p]] = p - deg_off[k];
    }
}

// Loads item q's constants; returns false for a skipped contact
// (both inverse masses zero: physics.comp:531).
SVD bool load_item(Item &it, uint32_t q, uint32_t nc, const float *__restrict__ prep,
                   const uint32_t *__restrict__ rank, const uint32_t *__restrict__ deg,
                   const float *__restrict__ rot, const float *__restrict__ inv_i) {
    uint32_t t = q / nc;
    uint32_t c = q - t * nc;
    const float *P = prep + (size_t)c * CPREP_WORDS;
    uint32_t flags = __float_as_uint(P[P_FLAGS]);
    if (flags & F_SKIP) return false;
    it.c = c;
    it.a = __float_as_uint(P[P_A]);
    it.b = __float_as_uint(P[P_B]);
    it.dyn_a = flags & F_DYN_A;
    it.dyn_b = flags & F_DYN_B;
    it.ta = it.dyn_a ? t * deg[it.a] + rank[2 * c] : 0u;
    it.tb = it.dyn_b ? t * deg[it.b] + rank[2 * c + 1] : 0u;
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
    return true;
}

// Ticket storage and hand-off scope of each solver form.  FORM_COH is the
// multi-workgroup form with the items' mutable words accessed at the
// device's point of coherence (CohMem): a hand-off is then the stores, a
// wait for their completion and the ticket bump, and the taker's ticket
// read and loads, with no L2 write-back (release) or L2 invalidation
// (acquire), which FORM_MULTI pays on every hand-off and which serialise
// on each XCD's L2.
enum : int { FORM_LDS = 0, FORM_GLOBAL = 1, FORM_MULTI = 2, FORM_COH = 3 };

template <int FORM>
SVD uint32_t ticket_load(const uint32_t *t) {
    if (FORM == FORM_LDS) return __hip_atomic_load(t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    return __hip_atomic_load(t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <int FORM>
SVD void ticket_bump(uint32_t *t) {
    if (FORM == FORM_LDS) __hip_atomic_fetch_add(t, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    else __hip_atomic_fetch_add(t, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <int FORM>
SVD void fence_acquire() {
    if (FORM == FORM_MULTI) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    else __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}
// FORM_COH: the item's loads are issued only after the ticket reads they
// depend on returned (the branch on them); the workgroup fence keeps the
// compiler from moving them, and the release waits for every store of the
// item to complete at the point of coherence before the ticket bump.
template <int FORM>
SVD void fence_release() {
    if (FORM == FORM_MULTI) release_agent();
    else if (FORM == FORM_COH) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_s_waitcnt(0);
    } else __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
}

template <int FORM, int THREADS>
__global__ void __launch_bounds__(THREADS) k_solve(
    const uint32_t *__restrict__ misc, uint32_t iters, uint32_t n_bodies, const float *__restrict__ prep,
    const uint32_t *__restrict__ rank, const uint32_t *__restrict__ deg, uint32_t *gticket, float *vel,
    float *omega, const float *__restrict__ rot, const float *__restrict__ inv_i, float *lambda, uint32_t *ctl,
    uint32_t *counters) {
    extern __shared__ uint32_t s_ticket[];
    __shared__ uint32_t s_next;
    const uint32_t nc = misc[MISC_NC];
    const uint32_t total = iters * nc;
    const int lane = threadIdx.x & 63;
    constexpr bool MULTI = FORM == FORM_MULTI || FORM == FORM_COH;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    uint32_t *ticket = gticket;
    uint32_t *next = ctl + 1;
    if (!MULTI) {
        if (FORM == FORM_LDS) {
            for (uint32_t i = threadIdx.x; i < n_bodies; i += THREADS) s_ticket[i] = 0;
            ticket = s_ticket;
        }
        if (threadIdx.x == 0) s_next = 0;
        next = &s_next;
        __syncthreads();
    }
    if (total == 0) return;

    Item it;
    bool has = false, fin = false;
    uint64_t last = wall_clock64();
    for (;;) {
        // refill every lane that has no item, in sweep order
        uint64_t need = __ballot(!has && !fin);
        while (need) {
            uint32_t cnt = (uint32_t)__popcll(need);
            int leader = __builtin_ctzll(need);
            uint32_t base = 0;
            if (lane == leader) {
                if (MULTI)
                    base = __hip_atomic_fetch_add(next, cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                else
                    base = __hip_atomic_fetch_add(next, cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
            base = __shfl(base, leader, 64);
            if (!has && !fin) {
                uint32_t q = base + (uint32_t)__popcll(need & lt);
                if (q >= total) fin = true;
                else has = load_item(it, q, nc, prep, rank, deg, rot, inv_i);
            }
            need = __ballot(!has && !fin);  // lanes whose item was a skipped contact
        }
        if (!__any(has)) break;
        bool progressed = false;
        if (has) {
            bool ready = (!it.dyn_a || ticket_load<FORM>(ticket + it.a) == it.ta) &&
                         (!it.dyn_b || ticket_load<FORM>(ticket + it.b) == it.tb);
            if (ready) {
                fence_acquire<FORM>();
                if (FORM == FORM_COH) apply_item<CohMem>(it, vel, omega, lambda);
                else apply_item<PlainMem>(it, vel, omega, lambda);
                fence_release<FORM>();
                if (it.dyn_a) ticket_bump<FORM>(ticket + it.a);
                if (it.dyn_b) ticket_bump<FORM>(ticket + it.b);
                has = false;
                progressed = true;
            }
        }
        if (__any(progressed)) {
            last = wall_clock64();
        } else {
            // the abort flag (another wave's stall) is read only after 20 us
            // without progress: each read is a round trip to the point of
            // coherence that would otherwise lengthen every poll
            const uint64_t idle = wall_clock64() - last;
            if (idle > ABORT_CHECK_TICKS) {
                if (__hip_atomic_load(ctl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) return;
                if (idle > STALL_TICKS) {
                    __hip_atomic_store(ctl, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (lane == 0) atomicOr(counters + 2, (uint32_t)SV_STATUS_SOLVER_TIMEOUT);
                    return;
                }
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
}

// ---------------------------------------------------------------------------
// Sequence form: the multi-workgroup ordered sweep with each body's state
// and its update count in the same 16-byte words, so a hand-off is a single
// round trip.  Body x has two words at the device's point of coherence,
// (v, k) and (w, k), k = updates of x completed; contact c has (lambda, t),
// t = iterations of c completed.  Item (t, c) polls its bodies' four words
// and its contact's word with device-coherent loads and runs when every
// count it depends on matches (k_a = t deg(a) + rank_a(c), the same for b,
// and t): the values it read are then those the predecessor wrote, because
// a 16-byte (8-byte) aligned word is written and read whole, and no later
// update of those words can happen before this item's.  It writes the new
// words with device-coherent stores and goes on: there is no separate
// ticket and no wait for store completion.  The sweep is identical to the
// serial one, item for item, like the ticket forms.

// raw buffer resource over a device array (offsets in bytes, out-of-range
// loads return 0 and stores are dropped); device-coherent cache policy SC1
constexpr int CPOL_SC1 = 16;
SVD __amdgpu_buffer_rsrc_t seq_rsrc(const void *p) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), 0, 0x7FFFFFFF, 0x00020000);
}

// Record of contact c for the pipelined form: rotation and inverse inertia of
// bodies a and b, and (rank, degree) of both incidences.
constexpr int SEQX_WORDS = 28;  // 7 float4
SVD void seq_record(uint32_t c, const float *__restrict__ prep, const uint32_t *__restrict__ rank,
                    const uint32_t *__restrict__ deg, const float *__restrict__ rot,
                    const float *__restrict__ inv_i, float *__restrict__ rec) {
    const float *P = prep + (size_t)c * CPREP_WORDS;
    const uint32_t flags = __float_as_uint(P[P_FLAGS]);
    if (flags & F_SKIP) return;
    const uint32_t a = __float_as_uint(P[P_A]), b = __float_as_uint(P[P_B]);
    float *R = rec + (size_t)c * SEQX_WORDS;
    for (int k = 0; k < 9; k++) R[k] = rot[(size_t)a * 9 + k];
    for (int k = 0; k < 3; k++) R[9 + k] = inv_i[(size_t)a * 3 + k];
    for (int k = 0; k < 9; k++) R[12 + k] = rot[(size_t)b * 9 + k];
    for (int k = 0; k < 3; k++) R[21 + k] = inv_i[(size_t)b * 3 + k];
    // a static side's rank is never written and its degree is zero:
    // neither is read for a static body
    R[24] = __uint_as_float((flags & F_DYN_A) ? rank[2 * c] : 0u);
    R[25] = __uint_as_float(deg[a]);
    R[26] = __uint_as_float((flags & F_DYN_B) ? rank[2 * c + 1] : 0u);
    R[27] = __uint_as_float(deg[b]);
}

// words of the sequence form from the reference buffers, and (rec non-null)
// the pipelined form's contact records
__global__ void __launch_bounds__(BLOCK) k_seq_pack(const uint32_t *__restrict__ misc, uint32_t n_bodies,
                                                    const float *__restrict__ vel, const float *__restrict__ omega,
                                                    const float *__restrict__ lambda, float4 *__restrict__ sv,
                                                    uint2 *__restrict__ sc, const float *__restrict__ prep,
                                                    const uint32_t *__restrict__ rank,
                                                    const uint32_t *__restrict__ deg, const float *__restrict__ rot,
                                                    const float *__restrict__ inv_i, float *__restrict__ rec) {
    const uint32_t nc = misc[MISC_NC];
    const uint32_t m = n_bodies > nc ? n_bodies : nc;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) {
        if (i < n_bodies) {
            sv[2 * i] = make_float4(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2], 0.0f);
            sv[2 * i + 1] = make_float4(omega[3 * i], omega[3 * i + 1], omega[3 * i + 2], 0.0f);
        }
        if (i < nc) {
            sc[i] = make_uint2(__float_as_uint(lambda[i]), 0u);
            if (rec) seq_record(i, prep, rank, deg, rot, inv_i, rec);
        }
    }
}

// With pos set, each body is also integrated from the velocities just
// written (k_integrate's update, one launch fewer in the full step).
__global__ void __launch_bounds__(BLOCK) k_seq_unpack(const uint32_t *__restrict__ misc, uint32_t n_bodies,
                                                      const float4 *__restrict__ sv, const uint2 *__restrict__ sc,
                                                      float *__restrict__ vel, float *__restrict__ omega,
                                                      float *__restrict__ lambda, float dt,
                                                      const float *__restrict__ inv_mass, float *__restrict__ pos,
                                                      float *__restrict__ quat) {
    const uint32_t nc = misc[MISC_NC];
    const uint32_t m = n_bodies > nc ? n_bodies : nc;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) {
        if (i < n_bodies) {
            const float4 v = sv[2 * i], w = sv[2 * i + 1];
            vel[3 * i] = v.x;
            vel[3 * i + 1] = v.y;
            vel[3 * i + 2] = v.z;
            omega[3 * i] = w.x;
            omega[3 * i + 1] = w.y;
            omega[3 * i + 2] = w.z;
            if (pos) integrate_body(i, dt, inv_mass, pos, quat, vel, omega);
        }
        if (i < nc) lambda[i] = __uint_as_float(sc[i].x);
    }
}

template <int THREADS>
__global__ void __launch_bounds__(THREADS) k_solve_seq(
    const uint32_t *__restrict__ misc, uint32_t iters, const float *__restrict__ prep,
    const uint32_t *__restrict__ rank, const uint32_t *__restrict__ deg, float4 *sv, uint2 *sc,
    const float *__restrict__ rot, const float *__restrict__ inv_i, uint32_t *ctl, uint32_t *counters,
    uint32_t streams) {
    const uint32_t nc = misc[MISC_NC];
    const uint32_t total = iters * nc;
    const int lane = threadIdx.x & 63;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    if (total == 0) return;
    const __amdgpu_buffer_rsrc_t rv = seq_rsrc(sv), rc = seq_rsrc(sc);

    typedef unsigned int u4 __attribute__((ext_vector_type(4)));
    typedef unsigned int u2 __attribute__((ext_vector_type(2)));
    Item it;
    uint32_t t = 0;
    // words of the held item and which of them already show the count it
    // waits for: a word that does cannot change before this item writes it,
    // so only the others are polled again
    u4 a0, a1, b0, b1;
    u2 lc;
    uint32_t got = 0;
    bool has = false, fin = false;
    // ticket streams, as in the pipelined kernel (seq_ticket)
    const uint32_t lg = streams > 1u ? 31u - __builtin_clz(streams) : 0u;
    const uint32_t stream = blockIdx.x & (streams - 1u);
    uint32_t *tk = ctl + 64 + 16 * stream;
    uint64_t last = wall_clock64();
    for (;;) {
        uint64_t need = __ballot(!has && !fin);
        while (need) {
            uint32_t cnt = (uint32_t)__popcll(need);
            int leader = __builtin_ctzll(need);
            uint32_t base = 0;
            if (lane == leader)
                base = __hip_atomic_fetch_add(tk, cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            base = __shfl(base, leader, 64);
            if (!has && !fin) {
                const uint32_t r = base + (uint32_t)__popcll(need & lt);
                const uint32_t q = ((r >> 6) << (6 + lg)) + (stream << 6) + (r & 63u);
                if (q >= total) fin = true;
                else {
                    has = load_item(it, q, nc, prep, rank, deg, rot, inv_i);
                    t = q / nc;
                    got = 0;
                }
            }
            need = __ballot(!has && !fin);
        }
        if (!__any(has)) break;
        bool progressed = false;
        if (has) {
            const int oa = (int)(it.a * 32u), ob = (int)(it.b * 32u), oc = (int)(it.c * 8u);
            if (!(got & 1u)) a0 = __builtin_amdgcn_raw_buffer_load_b128(rv, oa, 0, CPOL_SC1);
            if (!(got & 2u)) a1 = __builtin_amdgcn_raw_buffer_load_b128(rv, oa + 16, 0, CPOL_SC1);
            if (!(got & 4u)) b0 = __builtin_amdgcn_raw_buffer_load_b128(rv, ob, 0, CPOL_SC1);
            if (!(got & 8u)) b1 = __builtin_amdgcn_raw_buffer_load_b128(rv, ob + 16, 0, CPOL_SC1);
            if (!(got & 16u)) lc = __builtin_amdgcn_raw_buffer_load_b64(rc, oc, 0, CPOL_SC1);
            // a static body's words never change: any count will do
            got |= ((!it.dyn_a || a0[3] == it.ta) ? 1u : 0u) | ((!it.dyn_a || a1[3] == it.ta) ? 2u : 0u) |
                   ((!it.dyn_b || b0[3] == it.tb) ? 4u : 0u) | ((!it.dyn_b || b1[3] == it.tb) ? 8u : 0u) |
                   (lc[1] == t ? 16u : 0u);
            const bool ready = got == 31u;
            if (ready) {
                f3 la = mk(__uint_as_float(a0[0]), __uint_as_float(a0[1]), __uint_as_float(a0[2]));
                f3 wa = mk(__uint_as_float(a1[0]), __uint_as_float(a1[1]), __uint_as_float(a1[2]));
                f3 lb = mk(__uint_as_float(b0[0]), __uint_as_float(b0[1]), __uint_as_float(b0[2]));
                f3 wb = mk(__uint_as_float(b1[0]), __uint_as_float(b1[1]), __uint_as_float(b1[2]));
                const float nl = item_update(it, la, wa, lb, wb, __uint_as_float(lc[0]));
                if (it.dyn_a) {
                    const uint32_t k = it.ta + 1u;
                    __builtin_amdgcn_raw_buffer_store_b128(
                        u4{__float_as_uint(la.x), __float_as_uint(la.y), __float_as_uint(la.z), k}, rv, oa, 0,
                        CPOL_SC1);
                    __builtin_amdgcn_raw_buffer_store_b128(
                        u4{__float_as_uint(wa.x), __float_as_uint(wa.y), __float_as_uint(wa.z), k}, rv, oa + 16, 0,
                        CPOL_SC1);
                }
                if (it.dyn_b) {
                    const uint32_t k = it.tb + 1u;
                    __builtin_amdgcn_raw_buffer_store_b128(
                        u4{__float_as_uint(lb.x), __float_as_uint(lb.y), __float_as_uint(lb.z), k}, rv, ob, 0,
                        CPOL_SC1);
                    __builtin_amdgcn_raw_buffer_store_b128(
                        u4{__float_as_uint(wb.x), __float_as_uint(wb.y), __float_as_uint(wb.z), k}, rv, ob + 16, 0,
                        CPOL_SC1);
                }
                __builtin_amdgcn_raw_buffer_store_b64(u2{__float_as_uint(nl), t + 1u}, rc, oc, 0, CPOL_SC1);
                has = false;
                progressed = true;
            }
        }
        if (__any(progressed)) {
            last = wall_clock64();
        } else {
            // the abort flag (another wave's stall) is read only after 20 us
            // without progress: each read is a round trip to the point of
            // coherence that would otherwise lengthen every poll
            const uint64_t idle = wall_clock64() - last;
            if (idle > ABORT_CHECK_TICKS) {
                if (__hip_atomic_load(ctl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) return;
                if (idle > STALL_TICKS) {
                    __hip_atomic_store(ctl, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (lane == 0) atomicOr(counters + 2, (uint32_t)SV_STATUS_SOLVER_TIMEOUT);
                    return;
                }
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
}

// ---------------------------------------------------------------------------
// Pipelined sequence form: the sequence-form hand-off with each lane's next
// item fetched while it works on the current one.  A lane holds up to three
// items: the current one (polled and applied), the next one (its constants
// in flight or loaded) and a ticket (its position in sweep order, returned
// by the wave's atomic on the shared counter).  In one turn of the loop a
// lane polls and applies its current item, promotes the next item when the
// current one is done, turns last turn's ticket into loads of the next
// item, and the wave asks for tickets for lanes without one.  Every load a
// turn issues is consumed one turn later, after the next poll's wait, so a
// hand-off costs one round trip however often lanes refill; an item's
// constants come from the contact's sweep constants and its record below,
// whose addresses depend on the contact only.  Tickets are taken in
// increasing order per lane, so a lane's next item and ticket always follow
// its current item: the earliest unfinished item reaches a lane's current
// slot within two turns and then has all its predecessors done.

typedef unsigned int u4v __attribute__((ext_vector_type(4)));

// The loaded words of one item: its constants row and its record.  They are
// decoded only a turn after the loads are issued, so issuing them waits for
// nothing.
struct RawItem {
    u4v p[4];      // constants words 0-15
    uint32_t fl;   // constants word 16 (flags)
    u4v x[SEQX_WORDS / 4];
};
SVD void load_raw_item(RawItem &r, uint32_t c, __amdgpu_buffer_rsrc_t rp, __amdgpu_buffer_rsrc_t rx) {
    static_assert(CPREP_WORDS % 4 == 0 && SEQX_WORDS % 4 == 0 && P_FLAGS == 16 && P_IMB < 16, "row layout");
    const int op = (int)(c * (CPREP_WORDS * 4u)), ox = (int)(c * (SEQX_WORDS * 4u));
#pragma unroll
    for (int k = 0; k < 4; k++) r.p[k] = __builtin_amdgcn_raw_buffer_load_b128(rp, op + 16 * k, 0, 0);
    r.fl = __builtin_amdgcn_raw_buffer_load_b32(rp, op + 4 * P_FLAGS, 0, 0);
#pragma unroll
    for (int k = 0; k < SEQX_WORDS / 4; k++) r.x[k] = __builtin_amdgcn_raw_buffer_load_b128(rx, ox + 16 * k, 0, 0);
}

// Item (t, c) from its loaded words: the same values as load_item.  Returns
// the contact's flags (F_SKIP: not an item).
SVD uint32_t decode_item(Item &it, const RawItem &r, uint32_t c, uint32_t t) {
    uint32_t w[16], x[SEQX_WORDS];
#pragma unroll
    for (int k = 0; k < 16; k++) w[k] = r.p[k / 4][k % 4];
#pragma unroll
    for (int k = 0; k < SEQX_WORDS; k++) x[k] = r.x[k / 4][k % 4];
    const uint32_t flags = r.fl;
    auto f = [](uint32_t u) { return __uint_as_float(u); };
    it.c = c;
    it.a = w[P_A];
    it.b = w[P_B];
    it.dyn_a = flags & F_DYN_A;
    it.dyn_b = flags & F_DYN_B;
    it.ta = it.dyn_a ? t * x[25] + x[24] : 0u;
    it.tb = it.dyn_b ? t * x[27] + x[26] : 0u;
    it.n = mk(f(w[P_N]), f(w[P_N + 1]), f(w[P_N + 2]));
    it.ra = mk(f(w[P_RA]), f(w[P_RA + 1]), f(w[P_RA + 2]));
    it.rb = mk(f(w[P_RB]), f(w[P_RB + 1]), f(w[P_RB + 2]));
    it.kn = f(w[P_KN]);
    it.bias = f(w[P_BIAS]);
    it.mu = f(w[P_MU]);
    it.ima = f(w[P_IMA]);
    it.imb = f(w[P_IMB]);
    it.Ra.c0 = mk(f(x[0]), f(x[1]), f(x[2]));
    it.Ra.c1 = mk(f(x[3]), f(x[4]), f(x[5]));
    it.Ra.c2 = mk(f(x[6]), f(x[7]), f(x[8]));
    it.iia = mk(f(x[9]), f(x[10]), f(x[11]));
    it.Rb.c0 = mk(f(x[12]), f(x[13]), f(x[14]));
    it.Rb.c1 = mk(f(x[15]), f(x[16]), f(x[17]));
    it.Rb.c2 = mk(f(x[18]), f(x[19]), f(x[20]));
    it.iib = mk(f(x[21]), f(x[22]), f(x[23]));
    return flags;
}

template <int THREADS>
__global__ void __launch_bounds__(THREADS) k_solve_seqp(
    const uint32_t *__restrict__ misc, uint32_t iters, const float *__restrict__ prep,
    const float4 *__restrict__ rec, float4 *sv, uint2 *sc, uint32_t *ctl, uint32_t *counters, uint32_t streams) {
    const uint32_t nc = misc[MISC_NC];
    const uint32_t total = iters * nc;
    const int lane = threadIdx.x & 63;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    if (total == 0) return;
    const __amdgpu_buffer_rsrc_t rv = seq_rsrc(sv), rc = seq_rsrc(sc);
    const __amdgpu_buffer_rsrc_t rp = seq_rsrc(prep), rx = seq_rsrc(rec);

    typedef unsigned int u4 __attribute__((ext_vector_type(4)));
    typedef unsigned int u2 __attribute__((ext_vector_type(2)));
    Item it;
    RawItem nx;
    uint32_t t = 0, tn = 0, cn = 0, qk = 0;
    u4 a0, a1, b0, b1;
    u2 lc;
    uint32_t got = 0;
    bool has = false, has_n = false, has_k = false, fin = false;
    // the wave's ticket request of the last turn: its lanes, and the lane
    // that holds the counter value the atomic returned
    uint64_t pend = 0;
    int pend_leader = 0;
    uint32_t pend_base = 0;
    // ticket streams: with 8, stream s (workgroups s, s+8, ...; about one
    // XCD each) hands out the items of blocks s, s+8, s+16, ... of 64
    // consecutive items, in order, from its own counter (one counter served
    // every wave's refills one atomic at a time); the earliest unfinished
    // item is then always held, or next, in its stream
    const uint32_t lg = streams > 1u ? 31u - __builtin_clz(streams) : 0u;  // streams: a power of two
    const uint32_t stream = blockIdx.x & (streams - 1u);
    uint32_t *tk = ctl + 64 + 16 * stream;
    // turns this lane skips polling: the counts it last read leave at least
    // `backoff` + 2 updates of a body (or iterations of the contact) before
    // its item can run, and a body takes at most about one update per turn
    // (each waits for the previous one's words), so earlier polls would only
    // load the memory system every hand-off goes through
    uint32_t backoff = 0;
    uint64_t last = wall_clock64();
    for (;;) {
        bool progressed = false;
        bool polled = false;
        if (has && backoff) backoff--;
        else if (has) {
            polled = true;
            const int oa = (int)(it.a * 32u), ob = (int)(it.b * 32u), oc = (int)(it.c * 8u);
            if (!(got & 1u)) a0 = __builtin_amdgcn_raw_buffer_load_b128(rv, oa, 0, CPOL_SC1);
            if (!(got & 2u)) a1 = __builtin_amdgcn_raw_buffer_load_b128(rv, oa + 16, 0, CPOL_SC1);
            if (!(got & 4u)) b0 = __builtin_amdgcn_raw_buffer_load_b128(rv, ob, 0, CPOL_SC1);
            if (!(got & 8u)) b1 = __builtin_amdgcn_raw_buffer_load_b128(rv, ob + 16, 0, CPOL_SC1);
            if (!(got & 16u)) lc = __builtin_amdgcn_raw_buffer_load_b64(rc, oc, 0, CPOL_SC1);
        }
        // one wait per turn on every path: the poll's words and the loads of
        // the last turn's next item (so promoting it below waits for nothing)
        __builtin_amdgcn_s_waitcnt(0);
        if (has) {
            const int oa = (int)(it.a * 32u), ob = (int)(it.b * 32u), oc = (int)(it.c * 8u);
            got |= ((!it.dyn_a || a0[3] == it.ta) ? 1u : 0u) | ((!it.dyn_a || a1[3] == it.ta) ? 2u : 0u) |
                   ((!it.dyn_b || b0[3] == it.tb) ? 4u : 0u) | ((!it.dyn_b || b1[3] == it.tb) ? 8u : 0u) |
                   (lc[1] == t ? 16u : 0u);
            if (got == 31u) {
                f3 la = mk(__uint_as_float(a0[0]), __uint_as_float(a0[1]), __uint_as_float(a0[2]));
                f3 wa = mk(__uint_as_float(a1[0]), __uint_as_float(a1[1]), __uint_as_float(a1[2]));
                f3 lb = mk(__uint_as_float(b0[0]), __uint_as_float(b0[1]), __uint_as_float(b0[2]));
                f3 wb = mk(__uint_as_float(b1[0]), __uint_as_float(b1[1]), __uint_as_float(b1[2]));
                const float nl = item_update(it, la, wa, lb, wb, __uint_as_float(lc[0]));
                if (it.dyn_a) {
                    const uint32_t k = it.ta + 1u;
                    __builtin_amdgcn_raw_buffer_store_b128(
                        u4{__float_as_uint(la.x), __float_as_uint(la.y), __float_as_uint(la.z), k}, rv, oa, 0,
                        CPOL_SC1);
                    __builtin_amdgcn_raw_buffer_store_b128(
                        u4{__float_as_uint(wa.x), __float_as_uint(wa.y), __float_as_uint(wa.z), k}, rv, oa + 16, 0,
                        CPOL_SC1);
                }
                if (it.dyn_b) {
                    const uint32_t k = it.tb + 1u;
                    __builtin_amdgcn_raw_buffer_store_b128(
                        u4{__float_as_uint(lb.x), __float_as_uint(lb.y), __float_as_uint(lb.z), k}, rv, ob, 0,
                        CPOL_SC1);
                    __builtin_amdgcn_raw_buffer_store_b128(
                        u4{__float_as_uint(wb.x), __float_as_uint(wb.y), __float_as_uint(wb.z), k}, rv, ob + 16, 0,
                        CPOL_SC1);
                }
                __builtin_amdgcn_raw_buffer_store_b64(u2{__float_as_uint(nl), t + 1u}, rc, oc, 0, CPOL_SC1);
                has = false;
                progressed = true;
            } else if (polled) {
                uint32_t d = 0;
                if (!(got & 1u)) d = max(d, it.ta - a0[3]);
                if (!(got & 2u)) d = max(d, it.ta - a1[3]);
                if (!(got & 4u)) d = max(d, it.tb - b0[3]);
                if (!(got & 8u)) d = max(d, it.tb - b1[3]);
                if (!(got & 16u)) d = max(d, t - lc[1]);
                backoff = d >= 3u ? min(d - 2u, SEQ_BACKOFF_MAX) : 0u;
            }
        }
        // promote the next item (a skipped contact is dropped)
        if (!has && has_n) {
            has_n = false;
            if (!(decode_item(it, nx, cn, tn) & F_SKIP)) {
                t = tn;
                has = true;
                got = 0;
                backoff = 0;
            }
        }
        // last turn's tickets
        if (pend) {
            const uint32_t base = __shfl(pend_base, pend_leader, 64);
            if ((pend >> lane) & 1ull) {
                const uint32_t r = base + (uint32_t)__popcll(pend & lt);
                const uint32_t q = ((r >> 6) << (6 + lg)) + (stream << 6) + (r & 63u);
                if (q >= total) fin = true;
                else {
                    qk = q;
                    has_k = true;
                }
            }
            pend = 0;
        }
        // a ticket becomes the next item: its loads are consumed next turn
        if (has_k && !has_n) {
            tn = qk / nc;
            cn = qk - tn * nc;
            load_raw_item(nx, cn, rp, rx);
            has_n = true;
            has_k = false;
        }
        // tickets for lanes that hold nothing (a lane that reserved an item
        // while its current one waits could hold up an item that is ready):
        // the counter value is read next turn
        const uint64_t need = __ballot(!has && !has_n && !has_k && !fin);
        if (need) {
            pend = need;
            pend_leader = __builtin_ctzll(need);
            if (lane == pend_leader) {
                // an offset the compiler cannot see as uniform keeps this a
                // plain one-lane atomic whose value is read only next turn
                // (a uniform address is rewritten into a wave-aggregated
                // atomic that waits for the value at once)
                uint32_t z = 0;
                asm volatile("" : "+v"(z));
                pend_base = __hip_atomic_fetch_add(tk + z, (uint32_t)__popcll(need), __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        if (!pend && !__any(has || has_n || has_k)) break;
        if (__any(progressed)) {
            last = wall_clock64();
        } else {
            const uint64_t idle = wall_clock64() - last;
            if (idle > ABORT_CHECK_TICKS) {
                if (__hip_atomic_load(ctl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) return;
                if (idle > STALL_TICKS) {
                    __hip_atomic_store(ctl, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (lane == 0) atomicOr(counters + 2, (uint32_t)SV_STATUS_SOLVER_TIMEOUT);
                    return;
                }
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
}

// ---------------------------------------------------------------------------
// Paired pipelined form: the pipelined sequence kernel with two lanes per
// item, lane 2p for body a and lane 2p+1 for body b.  Each lane polls and
// writes its own body's words (and both poll the contact's word), loads only
// its own body's rotation and inverse inertia, and runs its body's half of
// item_update; the two halves meet through a lane swap (DPP quad permute
// 1,0,3,2) wherever item_update combines both bodies (the normal velocity,
// the tangent, the friction mass), and the terms both need are computed by
// both lanes from the same operands in the same order, so the update is the
// one item_update computes, bit for bit.  A turn's dependent arithmetic is
// then about 300 instructions instead of 500; a wave holds 32 items.

SVD uint32_t pair_swap(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, false);  // quad_perm(1, 0, 3, 2)
}
SVD float pair_swapf(float v) { return __uint_as_float(pair_swap(__float_as_uint(v))); }
SVD f3 pair_swap3(f3 v) { return mk(pair_swapf(v.x), pair_swapf(v.y), pair_swapf(v.z)); }

// One lane's share of an item: the contact's shared constants and its own
// body's.
struct Half {
    f3 n, r;           // contact normal, own lever arm
    float kn, bias, mu, ima, imb, im;
    m3 R;              // own rotation
    f3 ii;             // own inverse inertia
    uint32_t c, body, k;  // contact, own body, own body's awaited count
    bool dyn;
};
struct RawHalf {
    u4v p0, p2, p3;  // constants words 0-3, 8-11, 12-15 (normal, k_n, bias, mu, inverse masses)
    f3 r;            // own lever arm
    uint32_t im;     // own inverse mass
    uint32_t body;   // own body
    uint32_t fl;     // flags
    u4v x[3];        // own body's rotation and inverse inertia
    uint2 rd;        // own incidence's (rank, degree)
};
// Every own-body word is read at an offset that depends on the side, so the
// decode selects nothing by side (a select between two elements of a
// register array, by lane, becomes a per-lane index into the array).
SVD void load_raw_half(RawHalf &r, uint32_t c, uint32_t side, const float *__restrict__ prep,
                       const float *__restrict__ rec) {
    const float *p = prep + (size_t)c * CPREP_WORDS;
    const float *x = rec + (size_t)c * SEQX_WORDS;
    const u4v *p4 = reinterpret_cast<const u4v *>(p);
    r.p0 = p4[0];
    r.p2 = p4[2];
    r.p3 = p4[3];
    const float *rr = p + P_RA + 3 * side;
    r.r = mk(rr[0], rr[1], rr[2]);
    r.im = __float_as_uint(p[P_IMA + side]);
    r.body = __float_as_uint(p[P_A + side]);
    r.fl = __float_as_uint(p[P_FLAGS]);
    const u4v *x4 = reinterpret_cast<const u4v *>(x + 12 * side);
#pragma unroll
    for (int k = 0; k < 3; k++) r.x[k] = x4[k];
    r.rd = reinterpret_cast<const uint2 *>(x + 24)[side];
}
SVD uint32_t decode_half(Half &h, const RawHalf &r, uint32_t c, uint32_t t, uint32_t side) {
    static_assert(F_DYN_A == 2u && F_DYN_B == 4u && P_N == 0 && P_RA == 3 && P_RB == 6 && P_KN == 9 &&
                      P_BIAS == 10 && P_MU == 11 && P_IMA == 12 && P_IMB == 13 && P_A == 14 && P_B == 15,
                  "constants row layout");
    auto f = [](uint32_t u) { return __uint_as_float(u); };
    const uint32_t flags = r.fl;
    h.c = c;
    h.n = mk(f(r.p0[0]), f(r.p0[1]), f(r.p0[2]));
    h.kn = f(r.p2[1]);
    h.bias = f(r.p2[2]);
    h.mu = f(r.p2[3]);
    h.ima = f(r.p3[0]);
    h.imb = f(r.p3[1]);
    h.r = r.r;
    h.im = f(r.im);
    h.body = r.body;
    h.R.c0 = mk(f(r.x[0][0]), f(r.x[0][1]), f(r.x[0][2]));
    h.R.c1 = mk(f(r.x[0][3]), f(r.x[1][0]), f(r.x[1][1]));
    h.R.c2 = mk(f(r.x[1][2]), f(r.x[1][3]), f(r.x[2][0]));
    h.ii = mk(f(r.x[2][1]), f(r.x[2][2]), f(r.x[2][3]));
    h.dyn = (flags >> (1u + side)) & 1u;  // F_DYN_A = 2, F_DYN_B = 4
    h.k = h.dyn ? t * r.rd.y + r.rd.x : 0u;
    return flags;
}

// item_update (sv_dev.h) split over the pair: l, w are this lane's body's
// velocity and angular velocity; side 0 is body a, side 1 body b.  Both
// lanes return the new impulse.
//
// Each lane works with u = (other body's contact-point velocity) - (its
// own): vb - va on body a's lane, va - vb on body b's.  IEEE rounding is
// odd (RN(-x) = -RN(x)) and x - y = x + (-y), so the reference's relative
// velocity, tangent and impulses are u's and this lane's quantities up to a
// sign per lane, and the reference's "a minus, b plus" updates become one
// addition on both lanes.  (Only the sign of an exact zero can differ in
// between; it reaches a stored value only when a velocity component is -0,
// which the step never produces: gravity adds +0 terms and every rounded
// sum that cancels exactly is +0.)
SVD float half_update(const Half &h, bool side, f3 &l, f3 &w, float old) {
    f3 vs = l + cross3(w, h.r);
    f3 u = pair_swap3(vs) - vs;
    const float du = dot3(u, h.n);
    const float vn = side ? -du : du;  // dot(vb - va, n)
    float dln = (-vn + h.bias) / h.kn;
    const float nl = gmax(old + dln, 0.0f);
    dln = nl - old;
    const f3 imp = dln * h.n;
    // a: la - ima imp, wa - I(ra x imp); b: lb + imb imp, wb + I(rb x imp)
    const f3 ims = sel3(side, imp, -imp);
    l = l + h.im * ims;
    w = w + apply_inv_I(h.R, h.ii, cross3(h.r, ims));

    vs = l + cross3(w, h.r);
    u = pair_swap3(vs) - vs;
    // the tangent as this lane sees it: the reference's on body a's lane,
    // its negation on body b's
    const f3 vt = u - dot3(u, h.n) * h.n;
    const float vtl = len3(vt);
    if (vtl > 1e-6f) {
        const f3 td = mk(vt.x / vtl, vt.y / vtl, vt.z / vtl);
        const f3 rt = cross3(h.r, td);
        const float T = dot3(rt, apply_inv_I(h.R, h.ii, rt));
        const float To = pair_swapf(T);
        const float Ta = side ? To : T, Tb = side ? T : To;
        const float kt = h.ima + h.imb + Ta + Tb;
        const float dlt = gmin(vtl / kt, h.mu * nl);
        // a: la + ima impt, wa + I(ra x impt); b: lb - imb impt,
        // wb - I(rb x impt): with this lane's tangent, additions on both
        const f3 impt = dlt * td;
        l = l + h.im * impt;
        w = w + apply_inv_I(h.R, h.ii, cross3(h.r, impt));
    }
    return nl;
}

template <int THREADS>
__global__ void __launch_bounds__(THREADS) k_solve_seqp2(
    const uint32_t *__restrict__ misc, uint32_t iters, const float *__restrict__ prep,
    const float4 *__restrict__ rec, float4 *sv, uint2 *sc, uint32_t *ctl, uint32_t *counters, uint32_t streams) {
    const uint32_t nc = misc[MISC_NC];
    const uint32_t total = iters * nc;
    const int lane = threadIdx.x & 63;
    const uint32_t side = (uint32_t)lane & 1u;
    const int lane0 = lane & ~1;  // the pair's first lane
    const uint64_t ltp = (lane0 == 0) ? 0ull : (~0ull >> (64 - lane0));
    if (total == 0) return;
    const __amdgpu_buffer_rsrc_t rv = seq_rsrc(sv), rc = seq_rsrc(sc);
    const float *recf = reinterpret_cast<const float *>(rec);
    // ticket streams (seq_streams), one ticket per pair
    const uint32_t lg = streams > 1u ? 31u - __builtin_clz(streams) : 0u;
    const uint32_t stream = blockIdx.x & (streams - 1u);
    uint32_t *tk = ctl + 64 + 16 * stream;

    typedef unsigned int u4 __attribute__((ext_vector_type(4)));
    typedef unsigned int u2 __attribute__((ext_vector_type(2)));
    // both lanes of a pair hold the same item and take the same branches
    Half h;
    RawHalf nx;
    uint32_t t = 0, tn = 0, cn = 0, qk = 0;
    u4 v0, v1;  // own body's (v, k), (w, k)
    u2 lc;      // the contact's (lambda, t)
    uint32_t got = 0;
    // lane state as bits of one vector register (as separate booleans they
    // are lane masks in scalar registers, which this kernel runs out of)
    constexpr uint32_t HAS = 1, HAS_N = 2, HAS_K = 4, FIN = 8;
    uint32_t st = 0;
    uint64_t pend = 0;
    int pend_leader = 0;
    uint32_t pend_base = 0;
    uint64_t last = wall_clock64();
    for (;;) {
        bool progressed = false;
        if (st & HAS) {
            const int ob = (int)(h.body * 32u), oc = (int)(h.c * 8u);
            if (!(got & 1u)) v0 = __builtin_amdgcn_raw_buffer_load_b128(rv, ob, 0, CPOL_SC1);
            if (!(got & 2u)) v1 = __builtin_amdgcn_raw_buffer_load_b128(rv, ob + 16, 0, CPOL_SC1);
            if (!(got & 4u)) lc = __builtin_amdgcn_raw_buffer_load_b64(rc, oc, 0, CPOL_SC1);
        }
        // one wait per turn on every path (polls and the next item's loads)
        __builtin_amdgcn_s_waitcnt(0);
        if (st & HAS) {
            // a static body's words never change: any count will do
            got |= ((!h.dyn || v0[3] == h.k) ? 1u : 0u) | ((!h.dyn || v1[3] == h.k) ? 2u : 0u) |
                   (lc[1] == t ? 4u : 0u);
            const uint32_t both = got & pair_swap(got);
            if (both == 7u) {
                f3 l = mk(__uint_as_float(v0[0]), __uint_as_float(v0[1]), __uint_as_float(v0[2]));
                f3 w = mk(__uint_as_float(v1[0]), __uint_as_float(v1[1]), __uint_as_float(v1[2]));
                const float nl = half_update(h, side != 0u, l, w, __uint_as_float(lc[0]));
                if (h.dyn) {
                    const int ob = (int)(h.body * 32u);
                    const uint32_t k = h.k + 1u;
                    __builtin_amdgcn_raw_buffer_store_b128(
                        u4{__float_as_uint(l.x), __float_as_uint(l.y), __float_as_uint(l.z), k}, rv, ob, 0, CPOL_SC1);
                    __builtin_amdgcn_raw_buffer_store_b128(
                        u4{__float_as_uint(w.x), __float_as_uint(w.y), __float_as_uint(w.z), k}, rv, ob + 16, 0,
                        CPOL_SC1);
                }
                if (!side)
                    __builtin_amdgcn_raw_buffer_store_b64(u2{__float_as_uint(nl), t + 1u}, rc, (int)(h.c * 8u), 0,
                                                          CPOL_SC1);
                st &= ~HAS;
                progressed = true;
            }
        }
        // promote the next item (a skipped contact is dropped)
        if ((st & (HAS | HAS_N)) == HAS_N) {
            st &= ~HAS_N;
            if (!(decode_half(h, nx, cn, tn, side) & F_SKIP)) {
                t = tn;
                st |= HAS;
                got = 0;
            }
        }
        // last turn's tickets: one per pair, counted on the pairs' first lanes
        if (pend) {
            const uint32_t base = __shfl(pend_base, pend_leader, 64);
            if ((pend >> lane0) & 1ull) {
                const uint32_t r = base + (uint32_t)__popcll(pend & ltp);
                const uint32_t q = ((r >> 6) << (6 + lg)) + (stream << 6) + (r & 63u);
                if (q >= total) st |= FIN;
                else {
                    qk = q;
                    st |= HAS_K;
                }
            }
            pend = 0;
        }
        if ((st & (HAS_K | HAS_N)) == HAS_K) {
            tn = qk / nc;
            cn = qk - tn * nc;
            load_raw_half(nx, cn, side, prep, recf);
            st = (st & ~HAS_K) | HAS_N;
        }
        // tickets for pairs that hold nothing
        const uint64_t need = __ballot(side == 0u && st == 0u);
        if (need) {
            pend = need;
            pend_leader = __builtin_ctzll(need);
            if (lane == pend_leader) {
                // an offset the compiler cannot see as uniform keeps this a
                // plain one-lane atomic whose value is read only next turn
                uint32_t z = 0;
                asm volatile("" : "+v"(z));
                pend_base = __hip_atomic_fetch_add(tk + z, (uint32_t)__popcll(need), __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        if (!pend && !__any((st & (HAS | HAS_N | HAS_K)) != 0u)) break;
        if (__any(progressed)) {
            last = wall_clock64();
        } else {
            const uint64_t idle = wall_clock64() - last;
            if (idle > ABORT_CHECK_TICKS) {
                if (__hip_atomic_load(ctl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) return;
                if (idle > STALL_TICKS) {
                    __hip_atomic_store(ctl, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (lane == 0) atomicOr(counters + 2, (uint32_t)SV_STATUS_SOLVER_TIMEOUT);
                    return;
                }
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
}

// ---------------------------------------------------------------------------
// Flow form.  Item (t, c) has one predecessor per dynamic body (the previous
// update of that body in sweep order: none for the body's first contact in
// iteration 0) and one successor per dynamic body.  A countdown per item
// starts at its predecessor count; a finishing item decrements both
// successors' countdowns, and a successor whose countdown reaches zero is
// ready: the finishing lane runs it next, a second ready successor goes to a
// queue, and lanes without work take items from the queue in push order.
// Ready items share no dynamic body, so any interleaving of them gives the
// serial result.  Every hand-off is a device-scope release (before the
// decrement or the queue store) and acquire (after the decrement that reached
// zero or the load that found the queue slot filled).

// Successor words from the sorted incidences: position p of body x's
// segment is followed by p + 1, the segment's last position wraps to its
// first in the next iteration.
__global__ void __launch_bounds__(BLOCK) k_succ(const uint32_t *__restrict__ misc, uint32_t n_bodies,
                                                const uint32_t *__restrict__ skey, const uint32_t *__restrict__ sval,
                                                const uint32_t *__restrict__ deg_off, const uint32_t *__restrict__ deg,
                                                float *__restrict__ prep) {
    uint32_t ninc = misc[MISC_NINC];
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < ninc; p += gridDim.x * blockDim.x) {
        uint32_t x = skey[p];
        if (x >= n_bodies) continue;
        uint32_t v = sval[p], start = deg_off[x], end = start + deg[x];
        bool last = p + 1 >= end;
        uint32_t np = last ? start : p + 1;
        uint32_t succ = (sval[np] >> 1) | (last ? SUCC_WRAP : 0u) | (p > start ? SUCC_HASPRED : 0u);
        prep[(size_t)(v >> 1) * CPREP_WORDS + P_SUCC_A + (v & 1u)] = __uint_as_float(succ);
    }
}

SVD uint32_t flow_npred(const float *__restrict__ P, uint32_t t) {
    uint32_t flags = __float_as_uint(P[P_FLAGS]);
    bool da = flags & F_DYN_A, db = flags & F_DYN_B;
    if (t > 0) return (uint32_t)da + (uint32_t)db;
    uint32_t sa = __float_as_uint(P[P_SUCC_A]), sb = __float_as_uint(P[P_SUCC_B]);
    return (uint32_t)(da && (sa & SUCC_HASPRED)) + (uint32_t)(db && (sb & SUCC_HASPRED));
}

// countdowns = predecessor counts; queue slots empty (0)
__global__ void __launch_bounds__(BLOCK) k_flow_init(const uint32_t *__restrict__ misc, uint32_t iters,
                                                     const float *__restrict__ prep, uint32_t *__restrict__ count,
                                                     uint32_t *__restrict__ queue) {
    const uint32_t nc = misc[MISC_NC];
    const uint32_t span = iters * nc;
    for (uint32_t id = blockIdx.x * blockDim.x + threadIdx.x; id < span; id += gridDim.x * blockDim.x) {
        uint32_t t = id / nc, c = id - t * nc;
        count[id] = flow_npred(prep + (size_t)c * CPREP_WORDS, t);
        queue[id] = 0u;
    }
}

// items without predecessors (iteration 0, first contact of each dynamic
// body) start the queue; a slot holds item id + 1
__global__ void __launch_bounds__(BLOCK) k_flow_roots(const uint32_t *__restrict__ misc,
                                                      const float *__restrict__ prep, uint32_t *__restrict__ queue,
                                                      uint32_t *ctl) {
    const uint32_t nc = misc[MISC_NC];
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += gridDim.x * blockDim.x) {
        const float *P = prep + (size_t)c * CPREP_WORDS;
        if (__float_as_uint(P[P_FLAGS]) & F_SKIP) continue;
        if (flow_npred(P, 0) != 0) continue;
        uint32_t pos = atomicAdd(ctl + 2, 1u);
        queue[pos] = c + 1u;
    }
}

// Hand-off fences of the flow form: COH = mutable words at the point of
// coherence (CohMem), else device-scope release/acquire.
template <bool COH>
SVD void handoff_acquire() {
    if (COH) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    else __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}
template <bool COH>
SVD void handoff_release() {
    if (COH) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_s_waitcnt(0);
    } else {
        release_agent();
    }
}

// ctl: [0] abort flag, [1] queue head (slots claimed), [2] queue tail
// (slots pushed), [3] items done.  The queue never holds more than the
// iters * nact items of the sweep, and span = iters * nc bounds every slot.
template <bool COH>
__global__ void __launch_bounds__(MC_THREADS) k_flow(
    const uint32_t *__restrict__ misc, uint32_t iters, const float *__restrict__ prep, float *vel, float *omega,
    const float *__restrict__ rot, const float *__restrict__ inv_i, float *lambda, uint32_t *count, uint32_t *queue,
    uint32_t *ctl, uint32_t *counters) {
    const uint32_t nc = misc[MISC_NC];
    const uint32_t span = iters * nc;
    const uint32_t total = iters * misc[MISC_NACT];
    const int lane = threadIdx.x & 63;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    if (total == 0) return;
    Item it;
    bool has = false, fin = false;
    uint32_t id = 0, qpos = SUCC_NONE;
    uint64_t last = wall_clock64();
    for (;;) {
        // lanes with no item and no claimed slot claim the next slots (one atomic per wave)
        const bool want = !has && !fin && qpos == SUCC_NONE;
        uint64_t need = __ballot(want);
        if (need) {
            int leader = __builtin_ctzll(need);
            uint32_t base = 0;
            if (lane == leader)
                base = __hip_atomic_fetch_add(ctl + 1, (uint32_t)__popcll(need), __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT);
            base = __shfl(base, leader, 64);
            if (want) {
                qpos = base + (uint32_t)__popcll(need & lt);
                if (qpos >= span) fin = true;  // beyond the last slot that can be filled
            }
        }
        bool got = false;
        if (!has && !fin) {
            uint32_t v = __hip_atomic_load(queue + qpos, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (v) {
                handoff_acquire<COH>();
                id = v - 1u;
                has = true;
                got = true;
                qpos = SUCC_NONE;
            }
        }
        bool ran = false;
        if (has) {
            uint32_t t = id / nc, c = id - t * nc;
            const float *P = prep + (size_t)c * CPREP_WORDS;
            uint32_t flags = __float_as_uint(P[P_FLAGS]);
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
            uint32_t sa = __float_as_uint(P[P_SUCC_A]), sb = __float_as_uint(P[P_SUCC_B]);
            apply_item<typename std::conditional<COH, CohMem, PlainMem>::type>(it, vel, omega, lambda);
            ran = true;
            uint32_t ida = SUCC_NONE, idb = SUCC_NONE;
            if (it.dyn_a) {
                uint32_t ts = t + ((sa & SUCC_WRAP) ? 1u : 0u);
                if (ts < iters) ida = ts * nc + (sa & SUCC_ID);
            }
            if (it.dyn_b) {
                uint32_t ts = t + ((sb & SUCC_WRAP) ? 1u : 0u);
                if (ts < iters) idb = ts * nc + (sb & SUCC_ID);
            }
            // publish this update, then count it down in both successors
            handoff_release<COH>();
            bool ra = false, rb = false;
            if (ida != SUCC_NONE)
                ra = __hip_atomic_fetch_sub(count + ida, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 1u;
            if (idb != SUCC_NONE)
                rb = __hip_atomic_fetch_sub(count + idb, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 1u;
            if (ra || rb) handoff_acquire<COH>();
            has = ra || rb;
            id = ra ? ida : idb;
            const bool push = ra && rb;  // both ready: run a, queue b
            uint64_t pm = __ballot(push);
            if (pm) {
                int leader = __builtin_ctzll(pm);
                uint32_t base = 0;
                if (lane == leader)
                    base = __hip_atomic_fetch_add(ctl + 2, (uint32_t)__popcll(pm), __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_AGENT);
                base = __shfl(base, leader, 64);
                if (push) {
                    handoff_release<COH>();
                    __hip_atomic_store(queue + base + (uint32_t)__popcll(pm & lt), idb + 1u, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                }
            }
        }
        uint64_t rm = __ballot(ran);
        if (rm && lane == __builtin_ctzll(rm))
            __hip_atomic_fetch_add(ctl + 3, (uint32_t)__popcll(rm), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (rm || __ballot(got)) {
            last = wall_clock64();
        } else {
            if (__hip_atomic_load(ctl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) return;
            if (__hip_atomic_load(ctl + 3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= total) return;
            if (wall_clock64() - last > STALL_TICKS) {
                __hip_atomic_store(ctl, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (lane == 0) atomicOr(counters + 2, (uint32_t)SV_STATUS_SOLVER_TIMEOUT);
                return;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        if (!__ballot(!fin)) return;
    }
}

}  // namespace

// Sequence-form words (allocated on first use, sized from the capacities).
bool seq_reserve(sv_ctx *c) {
    Scratch &s = c->s;
    if (s.seq_v) return true;
    void *v = nullptr, *w = nullptr;
    if (hipMalloc(&v, 2 * (size_t)(c->caps.max_bodies ? c->caps.max_bodies : 1) * sizeof(float4)) != hipSuccess)
        return false;
    if (hipMalloc(&w, (size_t)(c->caps.max_contacts ? c->caps.max_contacts : 1) * sizeof(uint2)) != hipSuccess) {
        (void)hipFree(v);
        return false;
    }
    s.lazy.push_back(v);
    s.lazy.push_back(w);
    s.seq_v = static_cast<float4 *>(v);
    s.seq_c = static_cast<uint2 *>(w);
    return true;
}

// Records of the pipelined sequence form (allocated on first use).
bool seqx_reserve(sv_ctx *c) {
    Scratch &s = c->s;
    if (s.seq_x) return true;
    void *x = nullptr;
    const size_t mc = c->caps.max_contacts ? c->caps.max_contacts : 1;
    if (hipMalloc(&x, mc * SEQX_WORDS * sizeof(float)) != hipSuccess) return false;
    s.lazy.push_back(x);
    s.seq_x = static_cast<float4 *>(x);
    return true;
}

// Largest body count whose tickets fit in the single-workgroup kernel's LDS.
uint32_t solver_lds_bodies(sv_ctx *c) {
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device) != hipSuccess) lds = 65536;
    int avail = lds - 1024;  // static LDS of the kernel and slack
    return avail > 0 ? (uint32_t)(avail / 4) : 0u;
}

// Ticket streams of the sequence-form sweeps: 8 (one counter per XCD's
// share of the workgroups), fewer for fewer workgroups so every stream has
// some.  With one counter for every wave's refills the 32k-pile sweep took
// 1.44 ms at 192 workgroups and 1.90 at 256; with 8 it takes 1.30 and 1.28
// (16 streams: 1.28), and the 1M-body sweep 1.11 ms instead of 1.79.
int sweep_depth(sv_ctx *c, uint32_t *depth, uint64_t *items) {
    uint32_t nc = 0;
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipMemcpy(&nc, c->s.misc + MISC_NC, sizeof(nc), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return set_error(SV_E_DEVICE, "sv_sweep_depth: device read failed");
    if (nc > c->caps.max_contacts) nc = c->caps.max_contacts;
    static_assert(P_B == P_A + 1 && P_FLAGS == P_A + 2, "a, b and flags are adjacent words");
    std::vector<uint32_t> rows(3 * (size_t)nc);
    if (nc)
        e = hipMemcpy2D(rows.data(), 3 * sizeof(uint32_t), c->s.cprep + P_A, CPREP_WORDS * sizeof(float),
                        3 * sizeof(uint32_t), nc, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return set_error(SV_E_DEVICE, "sv_sweep_depth: device read failed");
    std::vector<uint32_t> level(c->caps.max_bodies, 0u);  // level of the last item on each body
    uint32_t d = 0;
    uint64_t cnt = 0;
    for (uint32_t t = 0; t < c->params.pgs_iterations; t++) {
        for (uint32_t k = 0; k < nc; k++) {
            const uint32_t a = rows[3 * k], b = rows[3 * k + 1], fl = rows[3 * k + 2];
            if ((fl & F_SKIP) || a >= c->caps.max_bodies || b >= c->caps.max_bodies) continue;
            const bool da = fl & F_DYN_A, db = fl & F_DYN_B;
            uint32_t l = 0;
            if (da) l = level[a];
            if (db && level[b] > l) l = level[b];
            ++l;
            if (da) level[a] = l;
            if (db) level[b] = l;
            if (l > d) d = l;
            ++cnt;
        }
    }
    *depth = d;
    *items = cnt;
    return SV_OK;
}

uint32_t seq_streams(uint32_t wgs) {
    uint32_t s = 8u;
    while (s > 1u && s > wgs) s >>= 1;
    return s;
}

// Step 5 (physics.comp:500-583) for the contacts counters[1] names.
// Reads lambda_n as the starting impulses (step 4 clears them, or, with
// warm start, carries them over).
static bool launch_sweep(sv_ctx *c, uint32_t n_bodies, float dt, bool integrate) {
    uint32_t *counters = B<uint32_t>(c, SV_ATOMIC_COUNTERS);
    uint32_t mc = c->caps.max_contacts;
    const bool setup_done = c->sweep_setup_done;  // by the full step's contact compaction
    c->sweep_setup_done = false;
    if (mc == 0 || n_bodies == 0) return false;
    if (!setup_done)
        hipLaunchKernelGGL(k_sweep_setup, dim3(div_up(n_bodies, 256)), dim3(256), 0, c->stream, counters, c->s.misc,
                           mc, c->s.deg, n_bodies, c->s.solve_ctl);
    // (the step's AABB stage has just cached every body's rotation)
    if (!c->rot_fresh) launch_rot(c, n_bodies);
    uint32_t iters = c->params.pgs_iterations;
    bool flow = c->solver_flow;
    if (flow) {
        size_t span = (size_t)iters * mc;
        ensure_u32(c->s.flow_count, c->s.flow_count_cap, (uint32_t)span);
        ensure_u32(c->s.flow_queue, c->s.flow_queue_cap, (uint32_t)span);
        // without its buffers the sweep runs in the ordered form (same result)
        if (c->s.flow_count_cap < span || c->s.flow_queue_cap < span) flow = false;
    }
    uint32_t grid = div_up(mc, BLOCK);
    if (grid > 8192) grid = 8192;
    hipLaunchKernelGGL(k_prep, dim3(grid), dim3(BLOCK), 0, c->stream, c->s.misc, n_bodies, dt, c->params.baumgarte,
                       c->params.slop, B<uint32_t>(c, SV_CONTACT_GEOM_A), B<uint32_t>(c, SV_CONTACT_GEOM_B),
                       B<uint32_t>(c, SV_GEOM_BODY_IDX), B<float>(c, SV_CONTACT_POS), B<float>(c, SV_CONTACT_NORMAL),
                       B<float>(c, SV_CONTACT_PENETRATION), B<float>(c, SV_BODY_POS), B<float>(c, SV_BODY_VEL),
                       B<float>(c, SV_BODY_OMEGA), B<float>(c, SV_BODY_INV_MASS), B<float>(c, SV_BODY_INV_INERTIA),
                       c->s.rot, B<float>(c, SV_GEOM_FRICTION), B<float>(c, SV_GEOM_RESTITUTION), c->s.cprep,
                       c->s.inc_key[0], c->s.inc_val[0], c->s.deg, B<float>(c, SV_CONTACT_LAMBDA_N),
                       c->warm_factor > 0.0f ? 1u : 0u);
    const bool warm = c->warm_factor > 0.0f;
    if (c->solver_colored && !warm) {
        if (!launch_solve_colored(c, n_bodies)) c->launch_err = hipErrorLaunchFailure;
        return false;
    }
    int key_bits = 32 - __builtin_clz(n_bodies);  // keys are in [0, n_bodies]
    int where = radix_sort(c, c->s.inc_key[0], c->s.inc_val[0], c->s.inc_key[1], c->s.inc_val[1],
                           c->s.misc + MISC_NINC, 2u * mc, key_bits);
    scan_exclusive(c, c->s.deg, c->s.deg_off, nullptr, n_bodies, nullptr);
    if (warm) launch_warm_apply(c, n_bodies, c->s.inc_val[where]);
    if (c->solver_colored) {
        if (!launch_solve_colored(c, n_bodies)) c->launch_err = hipErrorLaunchFailure;
        return false;
    }
    uint32_t rgrid = div_up(2u * (uint64_t)mc, BLOCK);
    if (rgrid > 8192) rgrid = 8192;
    float *vel = B<float>(c, SV_BODY_VEL), *omega = B<float>(c, SV_BODY_OMEGA);
    float *inv_i = B<float>(c, SV_BODY_INV_INERTIA), *lam = B<float>(c, SV_CONTACT_LAMBDA_N);
    if (flow) {
        hipLaunchKernelGGL(k_succ, dim3(rgrid), dim3(BLOCK), 0, c->stream, c->s.misc, n_bodies,
                           c->s.inc_key[where], c->s.inc_val[where], c->s.deg_off, c->s.deg, c->s.cprep);
        uint32_t igrid = div_up((uint64_t)iters * mc, BLOCK);
        if (igrid > 16384) igrid = 16384;
        hipLaunchKernelGGL(k_flow_init, dim3(igrid), dim3(BLOCK), 0, c->stream, c->s.misc, iters, c->s.cprep,
                           c->s.flow_count, c->s.flow_queue);
        hipLaunchKernelGGL(k_flow_roots, dim3(grid), dim3(BLOCK), 0, c->stream, c->s.misc, c->s.cprep,
                           c->s.flow_queue, c->s.solve_ctl);
        uint32_t wgs = c->solver_workgroups ? c->solver_workgroups : 64u;
        StageTimer sweep_timer(c, "phys/pgs_sweep");
        hipLaunchKernelGGL(c->sweep_variant != SWEEP_FENCED ? k_flow<true> : k_flow<false>, dim3(wgs), dim3(MC_THREADS), 0, c->stream,
                           c->s.misc, iters, c->s.cprep, vel,
                           omega, c->s.rot, inv_i, lam, c->s.flow_count, c->s.flow_queue, c->s.solve_ctl, counters);
        return false;
    }
    hipLaunchKernelGGL(k_rank, dim3(rgrid), dim3(BLOCK), 0, c->stream, c->s.misc, n_bodies, c->s.inc_key[where],
                       c->s.inc_val[where], c->s.deg_off, c->s.rank);
    // Sequence kernel: the paired pipelined one wins on a 32k-body pile
    // (sweep 1.05 ms per substep; 1.28 one lane per item, 1.4 basic), the
    // basic one on the 1M-body scene (1.11 ms; 1.46 pipelined, 2.09 paired:
    // the pipelined lanes idle two turns per refill, which a wide dependency
    // graph cannot hide).
    const bool seq_variant = c->sweep_variant == SWEEP_SEQ || c->sweep_variant == SWEEP_PIPE ||
                             c->sweep_variant == SWEEP_SEQ_BASIC || c->sweep_variant == SWEEP_PAIR;
    const bool small = n_bodies < (1u << 18);
    const bool paired = c->sweep_variant == SWEEP_PAIR || (c->sweep_variant == SWEEP_SEQ && small);
    const bool pipe = c->sweep_variant == SWEEP_PIPE || paired;
    // (buffer offsets of the words and records are 31-bit)
    const bool seq_ok = seq_variant && (uint64_t)n_bodies * 32u < 0x7FFFFFFFull &&
                        (uint64_t)mc * 8u < 0x7FFFFFFFull &&
                        (!pipe || (uint64_t)mc * SEQX_WORDS * 4u < 0x7FFFFFFFull) && seq_reserve(c) &&
                        (!pipe || seqx_reserve(c));
    // Form choice: small scenes keep every hand-off inside one CU; large ones
    // need the wider in-order window of the multi-CU form.  With 8 ticket
    // streams the paired kernel on a 32k-body pile (1.5M sweep items per
    // substep) was fastest with 512 one-wave workgroups (1.05 ms; 1.01 at
    // 384 and 1.03 at 448 at the first timed frame but slower over the timed
    // frames, 1.10 at 576, 1.30 at 768), the one-lane pipelined kernel with
    // 256 (1.28 ms; 1.30 at 192, 1.30 at 320, 1.33 at 384), the basic kernel on the
    // 1M-body scene (7.2M items) with 640 (1.11 ms; 1.39 at 384, 1.17 at
    // 512, 1.14 at 768; the pipelined one 1.46 at 512); the ticket form with
    // coherent accesses at 64 four-wave workgroups (2.96 ms; 3.21 at 32).
    const bool seq_form = seq_variant;
    // (the paired kernel holds 32 items per wave: twice the workgroups)
    const uint32_t seq_wgs = (n_bodies >= (1u << 18) ? 640u : 256u) * (paired ? 2u : 1u);
    uint32_t wgs = c->solver_workgroups ? c->solver_workgroups
                                        : (n_bodies >= 2048 ? (seq_form ? seq_wgs : 64u) : 1u);
    if (wgs > 1 && seq_ok) {
        uint32_t pgrid = div_up(n_bodies > mc ? n_bodies : mc, BLOCK);
        if (pgrid > 8192) pgrid = 8192;
        hipLaunchKernelGGL(k_seq_pack, dim3(pgrid), dim3(BLOCK), 0, c->stream, c->s.misc, n_bodies, vel, omega, lam,
                           c->s.seq_v, c->s.seq_c, c->s.cprep, c->s.rank, c->s.deg, c->s.rot, inv_i,
                           pipe ? reinterpret_cast<float *>(c->s.seq_x) : nullptr);
        if (pipe) {
            StageTimer sweep_timer(c, "phys/pgs_sweep");
            hipLaunchKernelGGL(paired ? k_solve_seqp2<SEQ_THREADS> : k_solve_seqp<SEQ_THREADS>, dim3(wgs),
                               dim3(SEQ_THREADS), 0, c->stream, c->s.misc,
                               iters, c->s.cprep, c->s.seq_x, c->s.seq_v, c->s.seq_c, c->s.solve_ctl, counters,
                               seq_streams(wgs));
        } else {
            StageTimer sweep_timer(c, "phys/pgs_sweep");
            hipLaunchKernelGGL(k_solve_seq<SEQ_THREADS>, dim3(wgs), dim3(SEQ_THREADS), 0, c->stream, c->s.misc,
                               iters, c->s.cprep, c->s.rank, c->s.deg, c->s.seq_v, c->s.seq_c, c->s.rot, inv_i,
                               c->s.solve_ctl, counters, seq_streams(wgs));
        }
        hipLaunchKernelGGL(k_seq_unpack, dim3(pgrid), dim3(BLOCK), 0, c->stream, c->s.misc, n_bodies, c->s.seq_v,
                           c->s.seq_c, vel, omega, lam, dt, B<float>(c, SV_BODY_INV_MASS),
                           integrate ? B<float>(c, SV_BODY_POS) : (float *)nullptr, B<float>(c, SV_BODY_QUAT));
        return integrate;
    }
    StageTimer sweep_timer(c, "phys/pgs_sweep");
    if (wgs > 1) {
        fill_u32(c, c->s.ticket, 0u, n_bodies);
        // (all working workgroups on one XCD was 4.5x slower with L2
        // maintenance hand-offs: the XCD's L2 serialises them)
        if (c->sweep_variant != SWEEP_FENCED)
            hipLaunchKernelGGL((k_solve<FORM_COH, MC_THREADS>), dim3(wgs), dim3(MC_THREADS), 0, c->stream, c->s.misc,
                               iters, n_bodies, c->s.cprep, c->s.rank, c->s.deg, c->s.ticket, vel, omega, c->s.rot,
                               inv_i, lam, c->s.solve_ctl, counters);
        else
            hipLaunchKernelGGL((k_solve<FORM_MULTI, MC_THREADS>), dim3(wgs), dim3(MC_THREADS), 0, c->stream,
                               c->s.misc, iters, n_bodies, c->s.cprep, c->s.rank, c->s.deg, c->s.ticket, vel, omega,
                               c->s.rot, inv_i, lam, c->s.solve_ctl, counters);
    } else if (n_bodies <= c->solver_lds_bodies) {
        hipLaunchKernelGGL((k_solve<FORM_LDS, SOLVE_THREADS>), dim3(1), dim3(SOLVE_THREADS), n_bodies * 4u,
                           c->stream, c->s.misc, iters, n_bodies, c->s.cprep, c->s.rank, c->s.deg, c->s.ticket, vel,
                           omega, c->s.rot, inv_i, lam, c->s.solve_ctl, counters);
    } else {
        fill_u32(c, c->s.ticket, 0u, n_bodies);
        hipLaunchKernelGGL((k_solve<FORM_GLOBAL, SOLVE_THREADS>), dim3(1), dim3(SOLVE_THREADS), 0, c->stream,
                           c->s.misc, iters, n_bodies, c->s.cprep, c->s.rank, c->s.deg, c->s.ticket, vel, omega,
                           c->s.rot, inv_i, lam, c->s.solve_ctl, counters);
    }
    return false;
}

bool launch_solve(sv_ctx *c, uint32_t n_bodies, float dt, bool integrate) {
    const bool done = launch_sweep(c, n_bodies, dt, integrate);
    // (the save reads impulses and contact ids only: after the integration is fine)
    if (c->warm_factor > 0.0f) launch_warm_save(c);
    return done;
}

}  // namespace sv
