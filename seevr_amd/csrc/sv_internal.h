// sv_internal.h - context layout and the launchers each kernel file exports
// to the C ABI layer (sv_capi.cpp).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>

#include "../../include/seevr.h"

namespace sv {

// Internal scratch buffers (not part of the reference binding table).
struct Scratch {
    uint32_t *hrange = nullptr;   // uint[2N] hull (start, count) read by the kernels
    float *rot = nullptr;         // float[9N] per-body rotation columns of this substep
    float4 *bpbox = nullptr;      // float4[2N] broadphase boxes: (min, static), (max, 0)
    // broadphase cells: nseg = bp_nseg_max(N), the most segments any n <= N uses
    uint32_t *row_count = nullptr;  // uint[N * nseg + 1] broadphase hits per (row, segment)
    uint32_t *row_off = nullptr;    // uint[N * nseg + 1] their exclusive scan
    uint32_t *bp_stage = nullptr;   // uint[N * nseg * BP_STAGE] first partners per (row, segment)
    uint32_t *bp_ovf = nullptr;     // uint[1 + ceil(N / 32) * nseg]: [0] = number of listed tiles, then the
                                    // (row tile, segment) tiles holding a cell with more than BP_STAGE hits
    uint32_t *np_flag = nullptr;    // uint[P] narrowphase hit flags
    float *sup_xyz = nullptr;       // float[3L] candidate vertex of each support-lookup entry (first use)
    uint32_t *np_idx = nullptr;     // uint[P] contact slot of each hit
    float *np_res = nullptr;        // float[8P] pos, normal, depth per pair
    float *cprep = nullptr;         // float[CPREP_WORDS * C] solver constants
    uint32_t *inc_key[2] = {nullptr, nullptr};  // uint[2C] incidence sort keys
    uint32_t *inc_val[2] = {nullptr, nullptr};  // uint[2C] incidence sort values
    uint32_t *deg = nullptr;        // uint[N + 1] dynamic contacts per body
    uint32_t *deg_off = nullptr;    // uint[N + 1]
    uint32_t *rank = nullptr;       // uint[2C] rank of each incidence in its body
    uint32_t *ticket = nullptr;     // uint[N] completed updates per body (solver)
    uint32_t *solve_ctl = nullptr;  // uint[SOLVE_CTL_WORDS]: [0] abort flag, [1..3] solver counters,
                                    // [64 + 16 s] ticket counter of sweep stream s
    uint32_t *flow_count = nullptr; // uint[iters * C] flow-form countdowns (allocated on first use)
    uint32_t flow_count_cap = 0;
    uint32_t *flow_queue = nullptr; // uint[iters * C] flow-form ready queue
    uint32_t flow_queue_cap = 0;
    uint32_t *scan_part = nullptr;  // scan partials
    uint32_t scan_part_cap = 0;
    uint32_t *sort_hist = nullptr;  // radix histograms
    uint32_t sort_hist_cap = 0;
    uint32_t *misc = nullptr;       // uint[64] small device scalars
    // LBVH broadphase (allocated on first use, sized from the capacities)
    float4 *bb_lo = nullptr;        // float4[2N-1] node min; w = smallest body index below (bits)
    float4 *bb_rec = nullptr;       // float4[4(N-1)]: per internal node its children's (lo, hi), as in bb_lo/bb_hi
    float4 *bb_hi = nullptr;        // float4[2N-1] node max; w = 1 if every body below is static (bits)
    uint32_t *bb_left = nullptr, *bb_right = nullptr, *bb_parent = nullptr, *bb_count = nullptr;  // [2N-1]
    uint32_t *bb_flags = nullptr;   // uint[N] refit arrival flags
    float *bb_cent = nullptr;       // float[3N] box centres
    float *bb_bounds = nullptr;     // float[6 + 6 * parts] centre bounds and partials
    uint32_t *bb_mk[2] = {nullptr, nullptr}, *bb_mv[2] = {nullptr, nullptr};  // [N] Morton sort
    uint32_t *bb_cnt = nullptr, *bb_off = nullptr, *bb_fill = nullptr;         // [N + 1] per-row pairs
    uint32_t *bb_pk[2] = {nullptr, nullptr}, *bb_pv[2] = {nullptr, nullptr};  // [P + N] pair sort
    // sweep-and-prune broadphase (lazily allocated)
    bool sap_ready = false;
    uint32_t *sap_kcount = nullptr, *sap_kfirst = nullptr;  // [2^16] boxes per key, first slot per key
    uint32_t *sap_skey = nullptr;                             // [N] keys in key order
    float *sap_map = nullptr;  // [4] key map (x0, scale), then the lower-bound extremes for the next map
    float4 *sap_box = nullptr;                                                      // [2N] boxes in key order
    uint32_t *sap_cnt = nullptr, *sap_off = nullptr, *sap_fill = nullptr;          // [N] per-row pairs
    uint32_t *sap_part = nullptr;  // [P + N] each kept row's partners
    uint32_t *sap_rows = nullptr;  // [N + 2] rows sorted by workgroups
    uint32_t *sap_cont = nullptr;  // [3N + 4] continued runs: items, entries, then (box, next candidate, first item)
    uint2 *sap_flat = nullptr;     // [P + N] pairs as the sweep finds them
    uint32_t *sap_tail = nullptr;  // [4] sap_flat's fill
    uint32_t *bb_stage = nullptr;   // uint[BB_STAGE * N] rows staged by each counting lane
    uint32_t *bb_scnt = nullptr;    // uint[N] rows found by each counting lane
    uint2 *bb_rng = nullptr;        // uint2[N] leaf range (first, last) of each internal node
    uint32_t *bb_tops = nullptr;    // refit list (sv_bvh.hip): per workgroup a count and its nodes
    bool bb_ready = false;
    // leaf ranges of the BVH build's internal nodes (allocated on first use;
    // bvh_rng_n = the leaf count they were computed for, 0 = none valid)
    uint2 *bvh_rng = nullptr;
    uint32_t *bvh_tops = nullptr;  // refit list (sv_bvh.hip): per workgroup a count and its nodes
    uint32_t bvh_rng_cap = 0, bvh_rng_n = 0;
    // graph-coloured sweep and residual (allocated on first use)
    unsigned long long *col_claim = nullptr;  // [N] colour-round claims per body
    uint32_t *col_color = nullptr;            // [C] colour of each contact
    uint32_t *col_cnt = nullptr, *col_off = nullptr, *col_bar = nullptr, *res_acc = nullptr;
    // sequence-form sweep words: (v, count), (w, count) per body and
    // (lambda, iteration) per contact
    float4 *seq_v = nullptr;  // [2N]
    uint2 *seq_c = nullptr;   // [C]
    // pipelined sequence form: per contact, both bodies' rotation and inverse
    // inertia and both incidences' rank and degree (7 float4)
    float4 *seq_x = nullptr;  // [C][7]
    // warm start: (geom a, geom b) keys and impulses of the last sweep
    unsigned long long *warm_key = nullptr;  // [C]
    float *warm_lam = nullptr;               // [C]
    // every lazily allocated block, freed with the context
    std::vector<void *> lazy;
};

// Hand-off variants of the multi-CU sweeps (sv_debug_set_sweep_variant).
// SWEEP_SEQ picks the pipelined sequence kernel below 2^18 bodies and the
// basic one above; SWEEP_PIPE and SWEEP_SEQ_BASIC force one of them.
enum { SWEEP_SEQ = 0, SWEEP_COHERENT = 1, SWEEP_FENCED = 2, SWEEP_PIPE = 3, SWEEP_SEQ_BASIC = 4, SWEEP_PAIR = 5 };

// Keys per radix-sort tile and values per scan tile (sv_prims.hip), and the
// scratch words the sort and the scan (one partial per tile) take for n
// keys or values.  Sort scratch, in this order: per pass the digit counts
// being accumulated, per pass the digit offsets, control words (done
// counter, tile counter per pass, abort word), and per pass, tile and digit
// one look-back status word.
constexpr uint32_t SORT_TILE_SMALL = 2048, SORT_TILE_LARGE = 8192;  // keys per tile, by sort size
constexpr uint32_t SORT_SMALL_MAX = 1u << 20;                       // largest sort on small tiles
constexpr uint32_t SCAN_TILE_VALUES = 1024;
constexpr uint32_t SORT_MAX_PASSES = 4;             // 32-bit keys, 8-bit digits
constexpr uint32_t SORT_MAX_KEYS = (1u << 30) - 1;  // counts share a word with 2 flag bits
constexpr size_t SORT_ACC_OFF = 0, SORT_GOFF_OFF = SORT_MAX_PASSES * 256, SORT_CTL_OFF = 2 * SORT_MAX_PASSES * 256;
constexpr size_t SORT_CTL_WORDS = 64;
constexpr size_t SORT_STATUS_OFF = SORT_CTL_OFF + SORT_CTL_WORDS;
enum : uint32_t { SORT_CTL_DONE = 0, SORT_CTL_TILE = 1, SORT_CTL_ABORT = 8, SORT_CTL_BAR = 9, SORT_CTL_EXIT = 10 };
constexpr uint32_t SORT_ONE_MAX = 1u << 19;  // largest sort done in one launch (256 resident tiles)
// Status words are sized for the most tiles any sort of at most n keys
// uses (small tiles up to SORT_SMALL_MAX keys, large above), so that
// scratch sized for a capacity serves every smaller sort.
constexpr size_t sort_tiles(size_t n) {
    const size_t small = ((n < SORT_SMALL_MAX ? n : SORT_SMALL_MAX) + SORT_TILE_SMALL - 1) / SORT_TILE_SMALL;
    const size_t large = (n + SORT_TILE_LARGE - 1) / SORT_TILE_LARGE;
    return small > large ? small : large;
}
constexpr size_t sort_hist_words(size_t n) { return SORT_STATUS_OFF + (size_t)SORT_MAX_PASSES * 256 * sort_tiles(n); }
// Single-pass scan scratch: a control word, a spare word, then two halves of
// 64-bit status words, one per tile.  Sized for the most tiles a scan of at
// most n values uses; a capacity of w words serves scan_tiles_cap(w) tiles.
constexpr size_t scan_part_words(size_t n) { return 2 + 4 * ((n + SCAN_TILE_VALUES - 1) / SCAN_TILE_VALUES); }
constexpr uint32_t scan_tiles_cap(uint32_t words) { return words < 2 ? 0u : (words - 2) / 4; }
// Centroids per partial of the LBVH's bounds reduction (sv_bvh.hip): the
// partials take 6 floats per tile.
constexpr uint32_t BOUNDS_TILE = 2048;

// Control words of the sweeps (Scratch::solve_ctl).
constexpr uint32_t SOLVE_CTL_WORDS = 1024;

// Words per contact in Scratch::cprep (see sv_pgs.hip).
constexpr int CPREP_WORDS = 20;
// Broadphase tiling (sv_physics.hip): rows per tile, partners staged per
// (row, column segment), and the column segment length for n bodies (at
// least 2048, and at most 64 segments).
constexpr uint32_t BP_ROWS = 32;
constexpr uint32_t BP_STAGE = 4;
inline uint32_t bp_seg_len(uint32_t n) {
    uint32_t s = (n + 63) / 64;
    s = (s + 255) / 256 * 256;
    return s < 2048u ? 2048u : s;
}
inline uint32_t bp_nseg(uint32_t n) { return n ? (n + bp_seg_len(n) - 1) / bp_seg_len(n) : 1u; }
// bp_nseg is not monotonic in n (at n = 131072 it is 64, at n = 131073 it
// is 57), so scratch sized for a capacity uses the largest value over every
// n <= cap: up to 2^17 bodies the segment length is 2048 and bp_nseg(n) =
// ceil(n / 2048); above, the segment length is at least n / 64.
inline uint32_t bp_nseg_max(uint32_t cap) {
    const uint32_t s = cap ? (cap + 2047u) / 2048u : 1u;
    return s < 64u ? s : 64u;
}
// (row, segment) cells and overflow-list words (one count word followed by
// at most one entry per (row tile, segment)) for any n <= cap bodies.
// The all-pairs kernels index cells with 32-bit words: they take at most
// 2^25 bodies (larger scenes use the LBVH broadphase, the same pair list).
constexpr uint32_t BP_ALL_PAIRS_MAX_BODIES = 1u << 25;
inline uint32_t bp_rows_cap(uint32_t cap) { return cap < BP_ALL_PAIRS_MAX_BODIES ? cap : BP_ALL_PAIRS_MAX_BODIES; }
inline size_t bp_cells_cap(uint32_t cap) { return (size_t)bp_rows_cap(cap) * bp_nseg_max(cap) + 1; }
inline size_t bp_ovf_words_cap(uint32_t cap) {
    return ((size_t)bp_rows_cap(cap) + BP_ROWS - 1) / BP_ROWS * bp_nseg_max(cap) + 1;
}

}  // namespace sv

struct sv_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    sv_caps caps{};
    sv_params params{};
    void *buf[SV_BUF_COUNT] = {};
    size_t bytes[SV_BUF_COUNT] = {};
    // false for a buffer the caller bound (sv_bind_buffer): not freed here
    bool owned[SV_BUF_COUNT] = {};
    // bumped by every sv_bind_buffer: captured graphs hold buffer addresses
    // and are rebuilt when it changed
    uint64_t bind_epoch = 0;
    // scan and sort scratch is caller-provided and sized up front: never
    // grown or freed (a view context of an allocation-free entry point)
    bool scratch_fixed = false;
    sv::Scratch s;
    // workgroups of the ordered PGS kernel: 0 = chosen by body count,
    // 1 = single-CU form, > 1 = multi-CU form with that many workgroups
    uint32_t solver_workgroups = 0;
    // largest body count whose solver tickets fit in one workgroup's LDS
    uint32_t solver_lds_bodies = 0;
    // sweep schedule: false = ordered (tickets, items taken in sweep order),
    // true = flow (countdowns, items run as their predecessors finish)
    bool solver_flow = false;
    // graph-coloured sweep (a different order than the reference's: judged by
    // the constraint residual, not by parity)
    bool solver_colored = false;
    // hand-off variant of the multi-CU sweeps (SV_SWEEP_*): sequence words
    // (state and update count in one word), coherent tickets, or tickets with
    // L2 write-back / invalidation fences
    int sweep_variant = sv::SWEEP_SEQ;
    // set by sv_physics_step between its AABB stage (which caches every
    // body's rotation) and its sweep: the sweep then skips recomputing them
    bool rot_fresh = false;
    bool sweep_setup_done = false;  // the contact compaction has done k_sweep_setup's work
    // launch failure to report at the next synchronization
    hipError_t launch_err = hipSuccess;
    // broadphase: 0 = chosen by body count, 1 = all pairs (tiled N^2), 2 = LBVH traversal
    uint32_t broadphase_mode = 0;
    // Scratch::sup_xyz is rebuilt when the hulls, lookups or hull ranges were
    // uploaded (dirty), when their device pointers were handed out (external
    // writes possible: rebuilt every step), or for another body count
    bool sup_dirty = true, sup_external = false;
    uint32_t sup_bodies = 0;
    // warm start (extension beside the reference path): 0 = off; otherwise a
    // contact whose body pair was in contact at the previous sweep starts
    // from warm_factor times that sweep's impulse
    float warm_factor = 0.0f;
    // stage timing
    bool timing = false;
    std::string timing_only;  // non-empty: the one stage timed
    std::vector<std::string> stage_names;
    std::vector<double> stage_ms;
    std::vector<uint64_t> stage_calls;
    std::vector<hipEvent_t> ev_pool;
    struct Pending { int stage; hipEvent_t a, b; };
    std::vector<Pending> pending;
    size_t ev_next = 0;
};

namespace sv {

// misc[] slots
enum : int {
    MISC_NP = 0,       // pairs to narrow (min(counters[0], max_pairs))
    MISC_NC = 1,       // contacts to solve (min(counters[1], max_contacts))
    MISC_NINC = 2,     // incidences (2 * NC)
    MISC_TMP = 3,
    MISC_SCAN_TOTAL = 4,
    MISC_DEG_TOTAL = 5,
    MISC_NACT = 6,     // contacts the sweep updates (not both inverse masses zero)
    MISC_BPE = 7,      // pairs the LBVH broadphase emits before sorting (rows that start below the cap)
    MISC_WARM_N = 8,   // contacts in the warm-start cache
    MISC_BOUNDS_DONE = 9,  // workgroups of the broadphase centre bounds that have arrived (zero between launches)
};

// Zero device memory and wait for it: the context streams are non-blocking,
// so a hipMemset on the null stream is not ordered before their kernels (a
// fresh allocation often reuses freed memory; its old contents must not be
// seen).  For allocation-time and host-side resets only (it synchronizes the
// device).
inline hipError_t zero_device(void *p, size_t bytes) {
    hipError_t e = hipMemset(p, 0, bytes);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    return e;
}

// ---- sv_prims.hip ----
// Exclusive scan of n = min(*n_dev, n_max) (or n_max when n_dev == null)
// uint32 values, one launch; writes the total to *total_dev when non-null,
// and with post_np also min(total, post_cap) to *post_np and
// SV_STATUS_PAIR_OVERFLOW into *post_status when the total exceeds post_cap.
void scan_exclusive(sv_ctx *c, const uint32_t *in, uint32_t *out, const uint32_t *n_dev,
                    uint32_t n_max, uint32_t *total_dev, uint32_t post_cap = 0, uint32_t *post_np = nullptr,
                    uint32_t *post_status = nullptr);
// Stable LSD radix sort of (key, val) by the low `key_bits` bits of key, for
// n = min(*n_dev, n_max) elements (n_max when n_dev == null).  Ping-pongs
// between (k0, v0) and (k1, v1); returns 0 when the result ends in (k0, v0)
// and 1 when in (k1, v1).
int radix_sort(sv_ctx *c, uint32_t *k0, uint32_t *v0, uint32_t *k1, uint32_t *v1,
               const uint32_t *n_dev, uint32_t n_max, int key_bits);
// After a synchronisation: whether a sort's look-back wait expired since the
// last call (the sort's abort word, cleared by the call).
bool sort_take_abort(sv_ctx *c);
// ---- sv_pgs.hip ----
// Ticket streams of the multi-workgroup sequence sweeps for `wgs`
// workgroups: a power of two, at most 8 and at most wgs.
uint32_t seq_streams(uint32_t wgs);
// Depth of the last sweep's dependency DAG (items in sweep order; an item
// depends on the previous item of each of its dynamic bodies) and its item
// count, from the prepared contacts on the device (synchronous).
int sweep_depth(sv_ctx *c, uint32_t *depth, uint64_t *items);
void fill_u32(sv_ctx *c, uint32_t *p, uint32_t value, size_t n);
void ensure_u32(uint32_t *&p, uint32_t &cap, uint32_t need);
// Allocate scan/sort scratch for the largest sizes a context can use.
void prims_reserve(sv_ctx *c, uint32_t max_scan_n, uint32_t max_sort_n);

// ---- sv_physics.hip ----
// exact: read SV_GEOM_HULL_RANGE whatever sv_params.hull_range_u32 says
void launch_hull_ranges(sv_ctx *c, uint32_t n, bool exact = false);
void ensure_support_vertices(sv_ctx *c, uint32_t n);
// clear_counters: also clear counters[0..1] (the full step's first dispatch)
void launch_gravity(sv_ctx *c, uint32_t n, float dt, const float g[3], bool clear_counters = false);
void launch_aabb(sv_ctx *c, uint32_t n);
// launch_gravity (counters cleared) and launch_aabb in one launch
void launch_gravity_aabb(sv_ctx *c, uint32_t n, float dt, const float g[3]);
void launch_broadphase(sv_ctx *c, uint32_t n);
// after_broadphase: the same substep's broadphase has set the pair count;
// clear_lambda: also clear the contact impulses (the next step's work);
// sweep_bodies > 0: also do the sweep's setup for that many bodies.
void launch_narrowphase(sv_ctx *c, bool after_broadphase = false, bool clear_lambda = false,
                        uint32_t sweep_bodies = 0);
void launch_clear_lambda(sv_ctx *c);
void launch_integrate(sv_ctx *c, uint32_t n, float dt);
void launch_update_verts(sv_ctx *c, uint32_t nv, uint32_t nb);
void launch_clear_counters(sv_ctx *c);
void launch_debug_atan2(sv_ctx *c, const float *y, const float *x, float *out, uint32_t n);
void launch_rot(sv_ctx *c, uint32_t n);
void launch_copy_calib(sv_ctx *c, const void *src, void *dst, size_t nbytes);
void launch_xcc_ids(sv_ctx *c, uint32_t *out, uint32_t n);
void launch_handoff(sv_ctx *c, uint32_t *flag, uint32_t rounds, uint32_t partner, uint64_t *out);
void launch_pack_boxes(sv_ctx *c, uint32_t n);

// ---- sv_pgs.hip ----
// integrate = true lets the sweep's write-back launch also integrate the
// bodies (launch_integrate's update); returns whether it did.
bool launch_solve(sv_ctx *c, uint32_t n_bodies, float dt, bool integrate = false);

// ---- sv_warm.hip ----
// Allocates the warm-start cache (false on failure) and empties it.
bool warm_reserve(sv_ctx *c);
// Step 4 with warm start: impulses of contacts whose (geom a, geom b) key is
// in the cache start from warm_factor times the cached impulse, the rest 0.
void launch_warm_lookup(sv_ctx *c);
// After the sweep: the cache becomes this sweep's keys and impulses.
void launch_warm_save(sv_ctx *c);
// Before the sweep: every warm-started contact's impulse applied to its
// dynamic bodies, body by body in contact order (sval = incidences sorted
// by body, Scratch::deg / deg_off their segments).
void launch_warm_apply(sv_ctx *c, uint32_t n_bodies, const uint32_t *sval);
uint32_t solver_lds_bodies(sv_ctx *c);

// ---- sv_color.hip ----
// Coloured sweep of the contacts prepared by the sweep's prep stage; false
// when it cannot run here (no cooperative launch, too many contacts).
bool launch_solve_colored(sv_ctx *c, uint32_t n_bodies);
// Normal-constraint residual of the current contact list and velocities.
int solver_residual(sv_ctx *c, double *rms, double *max_abs, uint32_t *count);
// Contacts per colour of the last coloured sweep (colours, then the
// remainder class); returns the number written or a negative status.
int color_counts(sv_ctx *c, uint32_t *out, uint32_t cap);

// ---- sv_bvh.hip ----
void bvh_stage(sv_ctx *c, uint32_t step, uint32_t n, bool bodies);
// Broadphase by LBVH traversal over Scratch::bpbox: the same pair list,
// order, truncation and counters as the all-pairs kernel.  Returns false
// (and launches nothing) when its scratch cannot be allocated.
bool launch_broadphase_lbvh(sv_ctx *c, uint32_t n);
bool launch_broadphase_sap(sv_ctx *c, uint32_t n);
// The broadphase's row sort alone: rows i < n of cnt[i] partners at
// pk[off[i]...] (any order) to sorted pairs below max_pairs; list = n + 2
// words with list[0] = 0; window_words = 0 picks the bitmap window, else a
// multiple of 256 up to 2^15.
void debug_sort_rows(sv_ctx *c, uint32_t n, const uint32_t *cnt, const uint32_t *off, const uint32_t *pk,
                     uint32_t *pairs, uint32_t max_pairs, uint32_t *list, uint32_t window_words);
// The automatic broadphase choice by body count: all-pairs below
// BP_SWEEP_MIN_BODIES, sweep and prune below BP_LBVH_MIN_BODIES, the LBVH
// from there (measured on the walled pile, broadphase per substep: 16k bodies
// all-pairs 129 us, sweep 133; 64k all-pairs 799, sweep 299, LBVH 624; 128k
// sweep 1636, LBVH 735 - one axis's runs grow with the pile's height).
constexpr uint32_t BP_SWEEP_MIN_BODIES = 1u << 14;
constexpr uint32_t BP_LBVH_MIN_BODIES = 3u << 15;
// The method (SV_BROADPHASE_ALL_PAIRS, _SWEEP or _LBVH) a launch over n
// bodies takes under the context's setting (without the scratch a method
// needs, the launch falls back to all-pairs).
int broadphase_method(const sv_ctx *c, uint32_t n);

// ---- sv_capi.cpp ----
// BVH.build's stage sequence (bvh.zig:63-172) over n triangles or bodies,
// then the context's synchronization and error check.
int bvh_build_run(sv_ctx *c, uint32_t n, bool bodies);

// ---- errors (sv_capi.cpp): sets sv_last_error() and returns code ----
int set_error(int code, const char *msg);

// ---- stage timing (sv_capi.cpp) ----
int timing_begin(sv_ctx *c, const char *name);
void timing_end(sv_ctx *c, int token);

struct StageTimer {
    sv_ctx *c;
    int tok;
    StageTimer(sv_ctx *ctx, const char *name) : c(ctx), tok(timing_begin(ctx, name)) {}
    ~StageTimer() { timing_end(c, tok); }
};

inline uint32_t div_up(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

}  // namespace sv

#define SV_LAUNCH_CHECK() (void)hipGetLastError()
