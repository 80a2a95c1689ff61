/*
 * seevr.h - C ABI of libseevr_hip.so, a gfx950 (MI355X) implementation of
 * seeVR's per-substep rigid-body physics (physics.comp) and LBVH build
 * (bvh.comp).
 *
 * The library owns the device buffers that seeVR's `Data` struct owns
 * (src/data.zig:36-121) and keeps their tightly packed std430 layouts:
 * float arrays with vec3 stride 3 and quaternion stride 4 (x, y, z, w),
 * uint arrays for indices.  A host replaces its Vulkan dispatches with the
 * calls below and its staging copies with sv_upload / sv_download.
 *
 * Conventions (every entry point):
 *   - return 0 on success, a negative SV_E_* code on failure; the message of
 *     the last failure on the calling thread is sv_last_error();
 *   - calls are synchronous (they return after the device work finished),
 *     like the reference's vkQueueWaitIdle (src/physics.zig:200);
 *   - one context is not thread-safe, like one Vulkan queue;
 *   - list overflow keeps the first `cap` entries in canonical order and the
 *     counters keep the raw totals, as the reference's counters do
 *     (src/shaders/physics.comp:404-411,465-469).
 *
 * Canonical order.  The reference appends pairs and contacts with atomicAdd,
 * so its list order is whatever order threads won the atomic.  This library
 * fixes the one schedule in which invocations run in index order: pairs in
 * lexicographic (i, j) order, contacts in pair order, and the Gauss-Seidel
 * sweep in contact order.
 */
#ifndef SEEVR_H
#define SEEVR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SV_ABI_VERSION 1

/* ---- status codes ---------------------------------------------------- */
enum {
    SV_OK = 0,
    SV_E_INVALID = -1,   /* bad argument (null pointer, unknown id, ...)     */
    SV_E_CAPACITY = -2,  /* a count exceeds the capacity fixed at creation   */
    SV_E_DEVICE = -3,    /* a HIP runtime call failed                        */
    SV_E_NOMEM = -4,     /* device or host allocation failed                 */
    SV_E_TIMEOUT = -5,   /* the ordered solver's bounded wait expired        */
    SV_E_RANGE = -6      /* offset/size outside a buffer                     */
};

/* Bits of counters[2] (the reference leaves counters[2..3] unused,
 * src/data.zig:101). */
enum {
    SV_STATUS_PAIR_OVERFLOW = 1u << 0,    /* pairs found > max_pairs        */
    SV_STATUS_CONTACT_OVERFLOW = 1u << 1, /* contacts found > max_contacts  */
    SV_STATUS_SOLVER_TIMEOUT = 1u << 2,   /* ordered PGS wait expired       */
    SV_STATUS_SORT_TIMEOUT = 1u << 3,     /* radix-sort look-back expired   */
    SV_STATUS_SCAN_TIMEOUT = 1u << 4      /* scan look-back expired         */
};

/* ---- buffer ids --------------------------------------------------------
 * 0..29 are the physics binding table, in the order of
 * src/physics.zig:37-68 (== src/shaders/physics.comp:16-57).
 * 30..45 are the BVH bindings of src/bvh.zig:32-51 (binding 0 `vertices`
 * is SV_VERTICES; bindings 16/17 are the float/uint views of one temp
 * buffer, SV_BVH_TEMP).  46+ are extensions of this library.            */
typedef enum sv_buf {
    SV_BODY_POS = 0,           /* float[3N]                                  */
    SV_BODY_QUAT = 1,          /* float[4N] (x,y,z,w)                        */
    SV_BODY_VEL = 2,           /* float[3N]                                  */
    SV_BODY_OMEGA = 3,         /* float[3N] world frame                      */
    SV_BODY_INV_MASS = 4,      /* float[N], 0 = static                       */
    SV_BODY_INV_INERTIA = 5,   /* float[3N] body-frame diagonal              */
    SV_BODY_VERT_START = 6,    /* uint[N] first render vertex of the body    */
    SV_BODY_VERT_COUNT = 7,    /* uint[N]                                    */
    SV_GEOM_BODY_IDX = 8,      /* uint[N] (geoms are 1:1 with bodies)        */
    SV_GEOM_DATA = 9,          /* float[8N]: [0]=hull start, [1]=hull count  */
    SV_GEOM_FRICTION = 10,     /* float[N]                                   */
    SV_GEOM_RESTITUTION = 11,  /* float[N]                                   */
    SV_HULL_VERTS = 12,        /* float[3H] local space                      */
    SV_VERTICES = 13,          /* float[3V] world-space render vertices      */
    SV_ORIGINAL_VERTICES = 14, /* float[3V] local-space render vertices      */
    SV_CONTACT_POS = 15,       /* float[3C]                                  */
    SV_CONTACT_NORMAL = 16,    /* float[3C], points from body A toward B     */
    SV_CONTACT_PENETRATION = 17, /* float[C]                                 */
    SV_CONTACT_GEOM_A = 18,    /* uint[C]                                    */
    SV_CONTACT_GEOM_B = 19,    /* uint[C]                                    */
    SV_CONTACT_LAMBDA_N = 20,  /* float[C] accumulated normal impulse        */
    SV_ATOMIC_COUNTERS = 21,   /* uint[4]: [0]=pairs found, [1]=contacts found,
                                  [2]=SV_STATUS_* bits                       */
    SV_BODY_AABB_MIN = 22,     /* float[3N]                                  */
    SV_BODY_AABB_MAX = 23,     /* float[3N]                                  */
    SV_COLLISION_PAIRS = 24,   /* uint[2P], (i, j) with i < j                */
    SV_AABB_TEMP_MIN = 25,     /* unused here (size 0): the per-body AABB is */
    SV_AABB_TEMP_MAX = 26,     /*   reduced without the reference's cascade  */
    SV_AABB_TEMP_BODY = 27,    /*   partials                                 */
    SV_SUPPORT_LOOKUP_OFFSET = 28, /* uint[N], 0 = brute-force support       */
    SV_SUPPORT_LOOKUP = 29,    /* uint[1 + 192 * lookups]                    */

    SV_INDICES = 30,           /* uint[3T] triangle vertex indices           */
    SV_BVH_CENTROIDS = 31,     /* float[3n]                                  */
    SV_BVH_MORTON = 32,        /* uint[n] sorted Morton codes after a build  */
    SV_BVH_MORTON_TEMP = 33,   /* uint[n] sort scratch                       */
    SV_BVH_SORT_INDICES = 34,  /* uint[n] primitive of each sorted position  */
    SV_BVH_SORT_TEMP = 35,     /* uint[n] sort scratch                       */
    SV_BVH_SCENE_BOUNDS = 36,  /* float[6] centroid bounds min xyz, max xyz  */
    SV_BVH_AABB_MIN = 37,      /* float[3(2n-1)] node boxes                  */
    SV_BVH_AABB_MAX = 38,      /* float[3(2n-1)]                             */
    SV_BVH_LEFT = 39,          /* uint[2n-1] left child / leaf first prim    */
    SV_BVH_RIGHT = 40,         /* uint[2n-1] right child (internal nodes)    */
    SV_BVH_COUNT = 41,         /* uint[2n-1] 0 = internal, 1 = leaf          */
    SV_BVH_PARENT = 42,        /* uint[2n-1], root = 0xFFFFFFFF              */
    SV_BVH_PRIM_INDICES = 43,  /* uint[n]                                    */
    SV_BVH_FLAGS = 44,         /* uint[n-1] refit arrival flags              */
    SV_BVH_TEMP = 45,          /* uint/float scratch (histograms, partials)  */

    SV_GEOM_HULL_RANGE = 46,   /* extension: uint[2N] (start, count) of each
                                  geom's hull, exact beyond 2^24 hull verts.
                                  Read instead of geom_data[0..1] when
                                  sv_params.hull_range_u32 is set.           */
    SV_BUF_COUNT = 47
} sv_buf;

/* ---- capacities and parameters ----------------------------------------
 * The reference fixes these at compile time (src/data.zig:11-29,
 * src/shaders/physics.comp:66-70); here they are chosen at creation.     */
typedef struct sv_caps {
    uint32_t max_bodies;      /* MAX_BODIES == MAX_GEOMS (data.zig:13-14)    */
    uint32_t max_hull_verts;  /* MAX_HULL_VERTS (data.zig:17)                */
    uint32_t max_vertices;    /* MAX_VERTICES (data.zig:11)                  */
    uint32_t max_triangles;   /* MAX_TRIANGLES (data.zig:12)                 */
    uint32_t max_pairs;       /* MAX_COLLISION_PAIRS (data.zig:16)           */
    uint32_t max_contacts;    /* MAX_CONTACTS (data.zig:15)                  */
    uint32_t max_support_entries; /* MAX_GEOMS * 192 (data.zig:208)          */
} sv_caps;

typedef struct sv_params {
    uint32_t pgs_iterations;  /* PGS_ITERATIONS, 4 in physics.comp:68        */
    float baumgarte;          /* BAUMGARTE 0.2 (physics.comp:66)             */
    float slop;               /* SLOP 0.005 (physics.comp:67)                */
    float linear_damping;     /* 0.999 per substep (physics.comp:422)        */
    float angular_damping;    /* 0.998 per substep (physics.comp:425)        */
    uint32_t hull_range_u32;  /* 1: read SV_GEOM_HULL_RANGE, 0: geom_data    */
} sv_params;

/* Defaults equal the reference constants, with the capacities of
 * src/data.zig:11-18. */
void sv_default_caps(sv_caps *caps);
void sv_default_params(sv_params *params);

typedef struct sv_ctx sv_ctx;

/* Physics.init + BVH.init + Data.init buffer creation
 * (src/physics.zig:34-78, src/bvh.zig:29-61, src/data.zig:179-250).
 * Uses the current HIP device. */
int sv_create(const sv_caps *caps, const sv_params *params, sv_ctx **out);
/* Physics.deinit / BVH.deinit / Data.deinit (physics.zig:219, bvh.zig:208,
 * data.zig:667). Null is accepted. */
int sv_destroy(sv_ctx *ctx);

int sv_get_params(sv_ctx *ctx, sv_params *out);
int sv_set_params(sv_ctx *ctx, const sv_params *params);

/* Device address and size of a buffer (descriptor binding, vulkan.zig:465). */
int sv_buffer_ptr(sv_ctx *ctx, sv_buf id, void **dev, size_t *bytes);
/* Bind caller-owned device memory as buffer `id`, the way Physics.init and
 * BVH.init bind the buffers that Data owns through descriptor sets
 * (src/physics.zig:37-68, src/bvh.zig:32-51, src/vulkan.zig:465-506).  The
 * memory must hold at least the bytes the capacities need (sv_buffer_ptr
 * reports them before binding), be 16-byte aligned, stay valid until the
 * context is destroyed or the id is bound again, and keep the std430 layout.
 * The library frees its own allocation of that id and never frees bound
 * memory.  Every later stage reads and writes the bound memory in place, so
 * a renderer can read SV_VERTICES without a copy.  Waits for the context's
 * queued work first. */
int sv_bind_buffer(sv_ctx *ctx, sv_buf id, void *dev, size_t nbytes);
/* Staging upload / readback (vulkan.zig:304-328). Offsets in bytes. */
int sv_upload(sv_ctx *ctx, sv_buf id, const void *src, size_t offset, size_t nbytes);
int sv_download(sv_ctx *ctx, sv_buf id, void *dst, size_t offset, size_t nbytes);

/* Physics.step (src/physics.zig:80-165): `substeps` substeps of
 * clear counters, gravity, AABB, broadphase, narrowphase, clear impulses,
 * PGS, integrate, then render vertices once.  No-op when num_bodies == 0
 * or substeps == 0.  sub_dt = dt / substeps. */
int sv_physics_step(sv_ctx *ctx, uint32_t num_bodies, uint32_t num_vertices,
                    uint32_t num_hull_verts, uint32_t substeps, float dt,
                    const float gravity[3]);

/* Physics.readCounters (src/physics.zig:213-217): raw pair and contact
 * totals of the last substep. */
int sv_read_counters(sv_ctx *ctx, uint32_t out[2]);

/* One physics stage, addressed like the reference's push constants
 * {step, count, dt, gravity, num_bodies} (src/physics.zig:7-15,
 * src/shaders/physics.comp:846-861).  Steps:
 *   0 gravity (count = bodies)            1, 9 per-body AABB (count = bodies)
 *   2 broadphase (count = bodies)         3 narrowphase over counters[0] pairs
 *   4 clear impulses                      5 PGS over counters[1] contacts
 *   6 integrate (count = bodies)          7 render vertices (count = verts)
 *   8 clear counters                      10, 11 accepted, no work (the
 *                                            AABB needs no cascade here)
 * Used to test each stage from identical inputs. */
int sv_physics_dispatch(sv_ctx *ctx, uint32_t step, uint32_t count, float dt,
                        const float gravity[3], uint32_t num_bodies);

/* BVH.build (src/bvh.zig:63-172) over `num_tris` triangles of
 * SV_VERTICES / SV_INDICES.  No-op when num_tris == 0. */
int sv_bvh_build(sv_ctx *ctx, uint32_t num_tris);
/* The same LBVH over body AABBs (primitive i = body i, centroid =
 * (min + max) / 2, leaf box = body AABB); reads SV_BODY_AABB_MIN/MAX. */
int sv_bvh_build_bodies(sv_ctx *ctx, uint32_t num_bodies);
/* Output arrays of a body LBVH built into caller memory, in the layout of
 * the SV_BVH_* buffers above (n = bodies, 2n - 1 nodes). */
typedef struct sv_bvh_arrays {
    float *centroids;        /* float[3n]                                    */
    uint32_t *morton;        /* uint[n] sorted Morton codes                  */
    uint32_t *sort_indices;  /* uint[n] body of each sorted position         */
    uint32_t *prim_indices;  /* uint[n]                                      */
    float *scene_bounds;     /* float[6]                                     */
    float *node_min;         /* float[3(2n-1)]                               */
    float *node_max;         /* float[3(2n-1)]                               */
    uint32_t *left, *right, *count, *parent;  /* uint[2n-1]                  */
} sv_bvh_arrays;

/* Largest body count of sv_bvh_build_bodies_noalloc: its scratch is a fixed
 * array of the library's code object (about 14 MB per GPU, held until the
 * process ends). */
#define SV_NOALLOC_MAX_BODIES (1u << 20)

/* sv_bvh_build_bodies without any allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * caller passes the body boxes (float[3n] min and max, device memory) and
 * the output arrays; everything else the build needs (sort buffers and
 * digit histograms, scan partials, bounds partials, arrival counters) lives
 * in one static device array of the library.  Same results as
 * sv_bvh_build_bodies, bit for bit.  Runs on `stream` (a hipStream_t; NULL =
 * the null stream) of the current device and returns after it finished.
 * num_bodies above SV_NOALLOC_MAX_BODIES is SV_E_CAPACITY.  Every call on a
 * GPU shares that GPU's scratch array: calls that overlap on one GPU (other
 * host threads, other streams) race and must not be issued. */
int sv_bvh_build_bodies_noalloc(const float *aabb_min, const float *aabb_max, uint32_t num_bodies,
                                const sv_bvh_arrays *out, void *stream);

/* One BVH stage, addressed like bvh.comp's push constants {step, num_tris}
 * (src/shaders/bvh.comp:401-416): 0 centroids, 1 scene bounds, 2 Morton,
 * 3 whole radix sort, 4 hierarchy, 5 leaf boxes, 6 clear flags,
 * 7 bottom-up refit; 8-11 accepted, no work. */
int sv_bvh_dispatch(sv_ctx *ctx, uint32_t step, uint32_t num_tris);

/* Evaluate the library's device atan2f on n host values (a numerics
 * check of the function the support lookup bins with). */
int sv_debug_atan2(sv_ctx *ctx, const float *y, const float *x, float *out, uint32_t n);

/* The library's stable LSD radix sort (the one the LBVH, the broadphase
 * pair order and the sweep's body incidences use) on host arrays: sorts the
 * first n of n_max (key, value) pairs by the low key_bits bits of the key,
 * with n read on the device as the pipeline passes it (n <= n_max <=
 * 2^30 - 1, 1 <= key_bits <= 32).  A test entry point. */
int sv_debug_radix_sort(sv_ctx *ctx, const uint32_t *keys, const uint32_t *vals, uint32_t n, uint32_t n_max,
                        uint32_t key_bits, uint32_t *keys_out, uint32_t *vals_out);

/* The library's exclusive scan (the one the broadphase rows, the contact
 * compaction and the sweep's body incidences use) on a host array: scans
 * the first n of n_max values, with n read on the device as the pipeline
 * passes it (n <= n_max <= 2^30); the sum of the n values (mod 2^32) goes to
 * *total.  A test entry point. */
int sv_debug_scan(sv_ctx *ctx, const uint32_t *in, uint32_t n, uint32_t n_max, uint32_t *out, uint32_t *total);

/* The broadphase's row sort on host arrays: row i < n holds cnt[i] distinct
 * partners j in (i, n), concatenated in row order in `entries` in any order
 * within a row; writes the sorted pairs (i, j) that fall below max_pairs to
 * pairs_out (2 * min(total, max_pairs) values).  window_words = 0 picks the
 * bitmap window of long rows as the pipeline does, else a multiple of 256 up
 * to 2^15 (several windows for small values).  A test entry point. */
int sv_debug_sort_rows(sv_ctx *ctx, const uint32_t *cnt, const uint32_t *entries, uint32_t n, uint32_t max_pairs,
                       uint32_t window_words, uint32_t *pairs_out);

/* Copy `nbytes` device to device `iters` times with a plain copy kernel
 * (the calibration kernel of the HBM byte counters) and report the average
 * read + write bandwidth in GB/s. */
int sv_debug_copy_bandwidth(sv_ctx *ctx, size_t nbytes, uint32_t iters, double *gbps);

/* Host only (no device work): words the all-pairs broadphase needs at n
 * bodies and what a context with capacity max_bodies allocates for them:
 * out[0] (row, segment) cells needed, out[1] allocated, out[2] overflow-list
 * words needed, out[3] allocated.  A sizing check for tests. */
int sv_debug_broadphase_scratch(uint32_t max_bodies, uint32_t n, uint64_t out[4]);

/* Depth of the last contact sweep's dependency graph and its item count:
 * items (iteration t, contact c) in the serial sweep order of
 * physics.comp:500-583, each depending on the previous item of each of its
 * dynamic bodies; depth = the longest chain.  Every schedule that keeps the
 * serial result needs at least `depth` dependent hand-offs per sweep.
 * Synchronous; reads the prepared contacts of the last substep. */
int sv_sweep_depth(sv_ctx *ctx, uint32_t *depth, uint64_t *items);

/* One-way hand-off latency between two waves through a device-coherent word
 * (the floor of one dependent step of the sweep): `rounds` round trips
 * between workgroup 0 and workgroup 1 (another XCD) or 8 (same_xcd != 0,
 * the same XCD under round-robin placement); xcc[] = the XCC ids the two
 * waves ran on.  A test and measurement entry point. */
int sv_debug_handoff(sv_ctx *ctx, uint32_t rounds, int same_xcd, double *us_one_way, uint32_t xcc[2]);

/* Host only: the ticket streams the multi-workgroup sequence sweeps use
 * with `wgs` workgroups (workgroup w serves stream w mod streams; stream s
 * hands out the items of 64-item blocks s, s + streams, ... in order). */
uint32_t sv_debug_seq_streams(uint32_t wgs);

/* XCC (XCD) hardware register read by the first wave of each of `nblocks`
 * one-wave blocks (a placement probe: which XCD runs which workgroup). */
int sv_debug_xcc_ids(sv_ctx *ctx, uint32_t nblocks, uint32_t *out);

/* Hand-off variant of the multi-workgroup sweeps (all give the same result):
 * 0 = sequence words (default: each body's velocities and update count in
 * one 16-byte word, one round trip per hand-off; the paired pipelined
 * kernel below 2^18 bodies, the basic one above), 1 = tickets with
 * device-coherent accesses, 2 = tickets with L2 write-back / invalidation
 * fences, 3 = sequence words, pipelined kernel (each lane's next item loaded
 * while it works on the current one), 4 = sequence words, basic kernel, 5 =
 * sequence words, pipelined kernel with two lanes per item (one per body).
 * For tests and measurements. */
int sv_debug_set_sweep_variant(sv_ctx *ctx, int variant);

/* Thread-local message of the last failure ("" if none). */
const char *sv_last_error(void);
/* Device name and gfx arch of the context's device. */
int sv_device_info(sv_ctx *ctx, char *name, size_t name_len, char *arch, size_t arch_len);
/* Average duration in microseconds of each stage over the calls since the
 * last reset, measured with HIP events on the library stream.  `names`
 * receives up to `cap` entries; returns the number of stages. */
int sv_stage_times(sv_ctx *ctx, int reset, const char **names, double *avg_us,
                   uint64_t *calls, int cap);
int sv_enable_stage_timing(sv_ctx *ctx, int enable);
/* Time only the stage named `stage` while timing is enabled (NULL or "" =
 * every stage).  Each timed stage adds two events to the stream, which cost
 * the step time; a measurement of one kernel then perturbs the step less. */
int sv_stage_timing_only(sv_ctx *ctx, const char *stage);

/* Schedule of the contact sweep (step 5).  No counterpart in the
 * reference, which runs the sweep on one invocation (physics.comp:523-581).
 * The first three give the reference's serial loop bit for bit and differ
 * only in speed.  SV_SOLVER_AUTO: ordered, one workgroup below 2048 bodies,
 * 32 workgroups from there.  SV_SOLVER_ORDERED: tickets per body, items
 * taken in sweep order by `workgroups` workgroups (1 = one CU).
 * SV_SOLVER_FLOW: per-item countdowns, items run as their predecessors
 * finish, `workgroups` workgroups (0 = 64).  SV_SOLVER_COLORED: the same
 * Gauss-Seidel item over the contacts sorted by a graph colouring (no two
 * contacts of a colour share a dynamic body), all contacts of a colour at
 * once; a different order, so a different (deterministic) result, judged by
 * sv_solver_residual.  Needs cooperative launch and < 2^24 contacts. */
enum { SV_SOLVER_AUTO = 0, SV_SOLVER_ORDERED = 1, SV_SOLVER_FLOW = 2, SV_SOLVER_COLORED = 3 };
int sv_set_solver(sv_ctx *ctx, int schedule, uint32_t workgroups);

/* Residual of the normal constraints for the contact list and velocities in
 * the context, to call after step 5 (sv_physics_dispatch) and before step 6:
 * per live contact, |v_n - bias| if its impulse is positive, else
 * max(bias - v_n, 0) (m/s).  Writes the RMS, the maximum and the count.  No
 * counterpart in the reference. */
int sv_solver_residual(sv_ctx *ctx, double *rms, double *max_abs, uint32_t *count);
/* Contacts per colour of the last coloured sweep, then the remainder class
 * swept by one lane; returns the number of entries written (<= cap). */
int sv_debug_color_counts(sv_ctx *ctx, uint32_t *out, uint32_t cap);

/* Broadphase method (step 2).  Both give the pair list of the reference's
 * all-pairs loop (physics.comp:451-472): the same set, in lexicographic
 * (i, j) order, truncated to max_pairs with counters[0] = the total.  No
 * counterpart in the reference.  SV_BROADPHASE_ALL_PAIRS: the LDS-tiled
 * N^2 test.  SV_BROADPHASE_LBVH: each body walks an LBVH built over the
 * body boxes of this substep.  SV_BROADPHASE_SWEEP: sweep and prune along x
 * (boxes ordered by their lower x bound, each tested against the run of
 * boxes that start before its upper x bound).  SV_BROADPHASE_AUTO picks by
 * body count: all-pairs below 2^14 bodies, sweep below 3 * 2^15, the LBVH
 * from there. */
enum { SV_BROADPHASE_AUTO = 0, SV_BROADPHASE_ALL_PAIRS = 1, SV_BROADPHASE_LBVH = 2, SV_BROADPHASE_SWEEP = 3 };
int sv_set_broadphase(sv_ctx *ctx, int method);
/* The method a broadphase over n bodies takes under the context's setting
 * (SV_BROADPHASE_ALL_PAIRS, _LBVH or _SWEEP), or a negative error code. */
int sv_broadphase_method(sv_ctx *ctx, uint32_t n);

/* Warm start (library extension beside the reference path; the reference
 * clears every impulse each substep, physics.comp:494-497, and lists warm
 * start as a to-do, todo warmup.md:1-3).  factor in (0, 1]: after each
 * sweep the contacts' (geom a, geom b) keys and impulses are kept; at the
 * next step 4 a contact of the same pair starts from factor times that
 * impulse, applied to both bodies at its first update in the sweep.
 * factor 0 (the default) is the reference's path.  Any call empties the
 * cache.  Applies to sv_physics_step and sv_physics_dispatch (not the
 * batched engine). */
int sv_set_warm_start(sv_ctx *ctx, float factor);

/* ---- host scene construction (src/data.zig) ------------------------- */
typedef struct sv_scene sv_scene;

typedef struct sv_scene_counts {
    uint32_t num_vertices, num_triangles, num_bodies, num_geoms;
    uint32_t num_hull_verts, num_materials, num_support_entries;
} sv_scene_counts;

int sv_scene_create(const sv_caps *caps, sv_scene **out);
int sv_scene_destroy(sv_scene *scene);
/* Data.addMaterial (data.zig:356-370); returns the material index. */
int sv_scene_add_material(sv_scene *scene, const float albedo[3], float roughness,
                          float metallic, const float emission[3], float ior,
                          float density, uint32_t *material_out);
/* Data.addBox (data.zig:392-485). */
int sv_scene_add_box(sv_scene *scene, const float center[3], const float half[3],
                     const float color[3], float friction, float restitution,
                     uint32_t material, uint32_t *body_out);
/* Data.addSphere (data.zig:487-597). */
int sv_scene_add_sphere(sv_scene *scene, const float center[3], float radius,
                        const float color[3], uint32_t segments, float friction,
                        float restitution, uint32_t material, uint32_t *body_out);
/* Overwrite a body's orientation and velocities (seeded scenes start with
 * random orientations; the reference's addBox/addSphere set identity). */
int sv_scene_set_body_state(sv_scene *scene, uint32_t body, const float quat[4],
                            const float vel[3], const float omega[3]);
int sv_scene_get_counts(const sv_scene *scene, sv_scene_counts *out);
/* Host staging array of a buffer id (the s_* slices of data.zig:123-170),
 * valid until the next add_* call. */
int sv_scene_array(const sv_scene *scene, sv_buf id, const void **host, size_t *bytes);
/* Data.upload (data.zig:599-655), including the derived body_vert_start /
 * body_vert_count and the exact SV_GEOM_HULL_RANGE extension. */
int sv_scene_upload(sv_scene *scene, sv_ctx *ctx);

/* Scene files (library extension; the reference lists scene save/load as
 * a to-do, README.md:99).  The file holds the scene's source arrays (the
 * s_* staging slices of data.zig:123-170) with a section table and an
 * FNV-1a checksum per section; loading checks every count, index and range
 * against the scene's capacities and replaces the scene's contents only when
 * all checks pass.  serialize with dst = NULL returns the size only. */
int sv_scene_serialize(const sv_scene *scene, void *dst, size_t capacity, size_t *bytes_out);
int sv_scene_deserialize(sv_scene *scene, const void *src, size_t bytes);
int sv_scene_save(const sv_scene *scene, const char *path);
int sv_scene_load(sv_scene *scene, const char *path);
/* Copy a context's body state (pos, quat, vel, omega) and render vertices
 * into the scene, so that a saved scene resumes the simulation there. */
int sv_scene_read_state(sv_scene *scene, sv_ctx *ctx);

/* Seeded benchmark scenes (library extension, no reference counterpart):
 *   1: 1,000 bodies in a 10x10 m column over a 50x1x50 m ground
 *   2: 32,768 bodies in a 30x30 m walled container (ground + 4 walls)
 *   4: `count` bodies (64 for the differentiable workload) over a ground
 *   5: 1,048,576 bodies (70% spheres) over a 220x1x220 m ground
 * `count` = 0 picks the workload's own body count; `sphere_segments` = 0
 * picks 16 (8 for workload 5). */
int sv_scene_build_workload(sv_scene *scene, int workload, uint32_t count,
                            uint32_t sphere_segments, uint64_t seed);

/* Independent scenes for the batched engine (library extension): scenes
 * [first, first + count) of a seeded family, each with its own ground at the
 * origin and floor(exp(U[ln min_bodies, ln max_bodies])) bodies spawned at
 * heights 1 m to `top` m (0 = 20).  Scene k draws from its own stream, so
 * any sub-range builds the same bodies.  Writes body_start[0..count] and
 * gravity_scale[0..count). */
int sv_scene_build_batch(sv_scene *scene, uint32_t first, uint32_t count, uint32_t min_bodies,
                         uint32_t max_bodies, uint32_t sphere_segments, float top, uint64_t seed,
                         uint32_t *body_start, float *gravity_scale);

/* ---- batched engine (library extension, no reference counterpart) -----
 * Steps many independent scenes of uneven body counts.  The bodies of all
 * scenes are concatenated in the buffers of one context (sv_batch_context;
 * fill them with sv_upload / sv_scene_upload); scene s owns bodies
 * [body_start[s], body_start[s+1]) and collides only with itself.  Each
 * scene gets max_scene_pairs pair slots and max_scene_contacts contact slots
 * at offset s times those capacities in SV_COLLISION_PAIRS and the contact
 * buffers.  A substep of every scene is one persistent launch (workgroups
 * take scenes, largest first, from a queue); a frame of `substeps` launches
 * plus the render-vertex update is replayed as one captured hipGraph.  Every
 * scene's results equal sv_physics_step on that scene alone, bit for bit,
 * including its pair and contact order and overflow truncation.  Hull
 * ranges are read from SV_GEOM_HULL_RANGE (sv_scene_upload fills it): the
 * concatenated hulls of a few thousand scenes pass 2^24 vertices, where
 * geom_data's float offsets are not exact. */
#define SV_BATCH_MAX_SCENE_BODIES 512
#define SV_BATCH_MAX_SCENE_CONTACTS 2048
typedef struct sv_batch sv_batch;
typedef struct sv_batch_caps {
    uint32_t max_scenes;
    uint32_t max_bodies;          /* all scenes together                     */
    uint32_t max_hull_verts;
    uint32_t max_vertices;
    uint32_t max_triangles;       /* render triangles (BVH input)            */
    uint32_t max_support_entries;
    uint32_t max_scene_bodies;    /* <= SV_BATCH_MAX_SCENE_BODIES            */
    uint32_t max_scene_pairs;     /* pair list capacity of each scene        */
    uint32_t max_scene_contacts;  /* <= SV_BATCH_MAX_SCENE_CONTACTS          */
} sv_batch_caps;
int sv_batch_create(const sv_batch_caps *caps, const sv_params *params, sv_batch **out);
int sv_batch_destroy(sv_batch *batch);
/* The context that holds the scenes' buffers (owned by the batch). */
sv_ctx *sv_batch_context(sv_batch *batch);
/* Scene table: body_start[0..num_scenes] and per-scene gravity[3 * s]. */
int sv_batch_set_scenes(sv_batch *batch, uint32_t num_scenes, const uint32_t *body_start,
                        const float *gravity);
/* One frame of every scene: `substeps` substeps of dt / substeps, then the
 * render vertices [0, num_vertices). */
int sv_batch_step(sv_batch *batch, uint32_t num_vertices, uint32_t substeps, float dt);
/* Per scene: total pairs found, total contacts found (last substep) and the
 * SV_STATUS_* bits of the last frame, 3 words per scene. */
int sv_batch_read_counters(sv_batch *batch, uint32_t *out);
/* Launch shape: threads per workgroup (64, 128, 256 or 512), workgroups per
 * CU of the persistent grid, and whether frames replay a hipGraph. */
int sv_batch_set_config(sv_batch *batch, uint32_t threads, uint32_t workgroups_per_cu, int use_graph);
/* Broadphase pair tile of the batched substep (a library extension; the
 * reference has no batched engine): lanes_per_row = 1 gives each lane a row
 * of the scene's all-pairs test, 2..64 (a power of two) give a group of that
 * many lanes one row, testing that many candidates a step.  The pair list
 * is the same for every tile.  Default 1. */
int sv_batch_set_pair_tile(sv_batch *batch, uint32_t lanes_per_row);
int sv_batch_pair_tile(sv_batch *batch, uint32_t *lanes_per_row);
/* Times every launch shape on the current scenes (state is restored) and
 * keeps the fastest; then times the pair tiles 1, 4, 16 and 64 at that
 * shape and keeps the fastest (read it with sv_batch_pair_tile); returns
 * the shape and its time per substep. */
int sv_batch_autotune(sv_batch *batch, uint32_t substeps, float dt, uint32_t *threads,
                      uint32_t *workgroups_per_cu, double *us_per_substep);
/* Phase timing of the batched substep kernel (a library extension, for
 * profiling).  While on, each workgroup's first thread reads the device's
 * 100 MHz wall clock after every phase and adds the phase's time to a
 * device total when it leaves.  The phases are: taking the next scene;
 * gravity and AABB; broadphase; narrowphase; sweep setup (constants and
 * incidence ranks); the ordered sweep; write-back and integration.
 * sv_batch_phase_times returns the 7 totals in microseconds summed over
 * workgroups, and the number of scenes processed, then zeroes them. */
int sv_batch_set_phase_timing(sv_batch *batch, int on);
int sv_batch_phase_times(sv_batch *batch, double us[7], uint64_t *scenes);

/* ---- differentiable batched rollout (library extension) ---------------
 * Gradients of `frames` frames of every scene with respect to the initial
 * state and the per-geom friction and restitution, with collision detection
 * held fixed per substep (the contacts of the forward run are constants).
 * State vectors are packed as pos[3N] | quat[4N] | vel[3N] | omega[3N]
 * (13N floats, N = all bodies of the batch); pointers may be host or device
 * memory (sv_upload / sv_download also take either).  sv_batch_diff_forward steps the frames like sv_batch_step
 * (without render vertices) and checkpoints every frame start;
 * sv_batch_diff_backward takes the adjoint of the final state (NULL = 0),
 * recomputes each frame from its checkpoint and writes the adjoint of the
 * initial state and of each geom's friction and restitution.  The batch
 * state is left at the end of the forward run. */
int sv_batch_diff_reserve(sv_batch *batch, uint32_t max_frames, uint32_t substeps);
int sv_batch_diff_forward(sv_batch *batch, uint32_t frames, uint32_t substeps, float dt);
int sv_batch_diff_backward(sv_batch *batch, uint32_t frames, uint32_t substeps, float dt,
                           const float *grad_state, float *out_state, float *out_friction,
                           float *out_restitution);

#ifdef __cplusplus
}
#endif
#endif /* SEEVR_H */
