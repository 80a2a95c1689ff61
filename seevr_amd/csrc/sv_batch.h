// sv_batch.h - shared pieces of the batched engine (sv_batch.hip) and its
// differentiable form (sv_diff.hip): launch arguments, the host-side batch
// object, a workgroup scan and the per-scene incidence ranking.
#pragma once
#include "sv_internal.h"
#include "sv_math.h"
#include "sv_dev.h"

namespace sv {

constexpr uint32_t NB_MAX = SV_BATCH_MAX_SCENE_BODIES;
constexpr uint32_t C_MAX = SV_BATCH_MAX_SCENE_CONTACTS;
constexpr uint32_t INC_MAX = 2 * C_MAX;  // incidences, a power of two
constexpr uint32_t MAX_SUBSTEPS = 4096;
constexpr uint64_t STALL_TICKS = 200000000ull;  // 2 s of the 100 MHz wall clock
// floats per recorded sweep item: v_a, w_a, v_b, w_b before the item, old impulse
constexpr uint32_t TAPE_ITEM = 13;
static_assert((INC_MAX & (INC_MAX - 1)) == 0, "incidence sort size must be a power of two");
static_assert(NB_MAX <= 512 && 2 * C_MAX <= 4096, "incidence key packing: 9-bit body, 12-bit slot");

struct BatchArgs {
    const uint32_t *sc_body;   // [S + 1] first body of each scene
    const float *sc_grav;      // [3S]
    const uint32_t *sc_order;  // [S] processing order
    uint32_t *sc_cnt;          // [3S] pairs, contacts, status bits
    uint32_t num_scenes;
    uint32_t *queue;           // this substep's work counter
    uint32_t *abort_flag;
    float *pos, *quat, *vel, *omega;
    const float *inv_mass, *inv_i, *hull, *fric, *rest;
    const uint32_t *hrange, *sl_off, *sl, *gbody;
    float *amin, *amax, *rot;
    uint32_t *pairs;
    uint32_t ps;  // pair slots per scene
    uint32_t pair_cols;  // broadphase pair tile: lanes per row (1: a lane per row; 4, 16, 64: a lane group)
    float *cpos, *cnrm, *cpen, *lambda;
    uint32_t *cga, *cgb;
    uint32_t cs;  // contact slots per scene
    float *cprep;
    float dt, ld, ad, baumgarte, slop;
    uint32_t iters;
    // differentiable form's tape (record launches only)
    float *t_items;   // [S][iters * cs][TAPE_ITEM] inputs of every sweep item
    float *t_w2;      // [3N] angular velocities after the sweep
    uint32_t *t_nc;   // [S] contacts solved
    float *t_state;   // [13N] state at the substep start: pos | quat | vel | omega
    uint32_t n_all;   // N, the bodies of all scenes (offsets into t_state)
    // phase timing (sv_batch_set_phase_timing): per phase, the 100 MHz wall
    // clock ticks every workgroup spent in it, and the scenes processed;
    // null when off
    unsigned long long *phase_ticks;  // [BATCH_PHASES + 1]
};
// Phases of k_batch_substep, in order: taking the next scene; gravity and
// AABB; broadphase; narrowphase; sweep setup (constants, incidence ranks);
// the ordered sweep; write-back and integration.
constexpr int BATCH_PHASES = 7;

// Exclusive scan of one value per thread over the workgroup; *total gets
// the sum.  Every thread of the block must call it.
template <int T>
SVD uint32_t block_scan(uint32_t v, uint32_t *s_w, uint32_t *total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) s_w[w] = x;
    __syncthreads();
    uint32_t base = 0, tot = 0;
#pragma unroll
    for (int q = 0; q < T / 64; q++) {
        uint32_t t = s_w[q];
        base += q < w ? t : 0u;
        tot += t;
    }
    __syncthreads();
    *total = tot;
    return base + x - v;
}

// Ranks of the (body, contact) incidences of one scene, for the ordered
// sweep and its reverse.  s_key must hold, for c < nc, key[2c + side] =
// (local body << 12) | (2c + side) for a live dynamic side and 0xFFFFFFFF
// otherwise; L is the power of two >= 2 nc the keys are padded to.  Sorts the
// keys (bitonic, in LDS), then writes s_rank[2c + side] = position of the
// incidence among its body's and s_deg[body] = the body's incidence count.
template <int T>
SVD void rank_incidences(uint32_t *s_key, uint16_t *s_rank, uint16_t *s_start, uint16_t *s_deg, uint32_t L) {
    const int tid = threadIdx.x;
    for (uint32_t kk = 2; kk <= L; kk <<= 1) {
        for (uint32_t jj = kk >> 1; jj > 0; jj >>= 1) {
            for (uint32_t i = tid; i < L; i += T) {
                const uint32_t ixj = i ^ jj;
                if (ixj > i) {
                    const uint32_t x = s_key[i], y = s_key[ixj];
                    const bool up = (i & kk) == 0;
                    if ((x > y) == up) {
                        s_key[i] = y;
                        s_key[ixj] = x;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (uint32_t p = tid; p < L; p += T) {
        const uint32_t key = s_key[p];
        if (key == 0xFFFFFFFFu) continue;
        const uint32_t x = key >> 12;
        if (p == 0 || (s_key[p - 1] >> 12) != x) s_start[x] = (uint16_t)p;
    }
    __syncthreads();
    for (uint32_t p = tid; p < L; p += T) {
        const uint32_t key = s_key[p];
        if (key == 0xFFFFFFFFu) continue;
        const uint32_t x = key >> 12;
        const uint32_t r = p - s_start[x];
        s_rank[key & 0xFFFu] = (uint16_t)r;
        if (p + 1 == L || (s_key[p + 1] >> 12) != x) s_deg[x] = (uint16_t)(r + 1);
    }
    __syncthreads();
}

// Launches of the batched substep (sv_batch.hip): record = false for the
// plain batched step; record = true also writes the differentiable form's
// tape (A.t_items, A.t_w2, A.t_nc; A.cprep points into the tape).
void batch_launch_substep(const BatchArgs &A, uint32_t threads, uint32_t grid, bool record, hipStream_t stream);

}  // namespace sv

struct sv_batch {
    sv_ctx *ctx = nullptr;
    sv_batch_caps caps{};
    uint32_t num_scenes = 0, num_bodies = 0;
    uint32_t *sc_body = nullptr, *sc_order = nullptr, *sc_cnt = nullptr, *queue = nullptr, *abort_flag = nullptr;
    float *sc_grav = nullptr;
    float *snap = nullptr;  // pos, quat, vel, omega, counters (autotune)
    unsigned long long *phase_ticks = nullptr;  // [BATCH_PHASES + 1], allocated when phase timing is on
    bool phase_timing = false;
    int cus = 0;
    uint32_t threads = 256, wgs_per_cu = 2;
    uint32_t pair_cols = 1;  // broadphase lanes per row (sv_batch_set_pair_tile)
    bool use_graph = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool graph_valid = false;
    uint32_t g_nv = 0, g_substeps = 0;
    float g_dt = 0.0f;
    sv_params g_params{};
    uint64_t g_epoch = 0;  // the context's bind_epoch when the graph was captured
    // differentiable form (sv_diff.hip), sized by sv_batch_diff_reserve
    uint32_t d_frames = 0, d_substeps = 0, d_iters = 0;
    float *d_ckpt = nullptr;     // [(frames + 1)][13N] state at each frame start (pos, quat, vel, omega)
    float *d_state = nullptr;    // [2][substeps][13N] state at each substep start (two tape sets)
    float *d_w2 = nullptr;       // [substeps][3N] angular velocity after each sweep
    float *d_prep = nullptr;     // [substeps][S * cs][CPREP_WORDS] sweep constants
    float *d_items = nullptr;    // [substeps][S][iters * cs][TAPE_ITEM] sweep item inputs
    uint32_t *d_nc = nullptr;    // [substeps][S] contacts solved
    double *d_adj = nullptr;     // [13N] adjoint of the state (the reverse pass runs in FP64)
    float *d_adj32 = nullptr;    // [13N] the caller's FP32 adjoint in and out
    double *d_contrib = nullptr; // [S * 2cs][20] per-incidence adjoint contributions
    double *d_amat = nullptr;    // [2N] friction | restitution adjoints per geom, FP64 sums
    float *d_amat32 = nullptr;   // [2N] the same rounded to float for the output
    hipStream_t d_stream = nullptr;  // adjoint kernels (default priority), beside the context stream
    hipEvent_t d_ev[5] = {};
};

namespace sv {
BatchArgs batch_args(sv_batch *b, float dt, uint32_t *queue);
// Substeps of every scene on the context stream (no render vertices).
void enqueue_substeps(sv_batch *b, uint32_t substeps, float sub_dt, uint32_t threads, uint32_t wgs_per_cu);
// Synchronizes the batch stream and turns launch errors and the sweep's
// abort flag into status codes.
int batch_finish(sv_batch *b);
int batch_dev_fail(hipError_t e, const char *what);
}  // namespace sv
