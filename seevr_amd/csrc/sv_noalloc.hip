// sv_noalloc.hip - sv_bvh_build_bodies_noalloc: the body LBVH build of
// sv_bvh_build_bodies (bvh.comp's stages over body boxes) into caller
// arrays, with every other device buffer it needs carved out of one static
// device array, so that a call allocates nothing.  The build runs on a view
// context: a stack sv_ctx whose buffer table points at the caller's arrays
// and at parts of that array, and whose scan and sort scratch is fixed.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>

#include "sv_internal.h"

namespace {

constexpr size_t align64(size_t words) { return (words + 63) / 64 * 64; }  // 256-byte parts

// Word offsets of the parts of the scratch array for n bodies.  The sort's
// scratch comes first at its size for the maximum, so that its control
// words (a completion counter and tile counters that every sort leaves at
// zero) sit at one address for every n; the other parts grow with n, so the
// layout for any n <= the maximum fits the array.
struct Carve {
    size_t hist, morton_temp, sort_temp, flags, bounds, total;
};
constexpr Carve carve(size_t n) {
    Carve c{};
    size_t o = 0;
    c.hist = o; o += align64(sv::sort_hist_words(SV_NOALLOC_MAX_BODIES));             // sort histograms, control, look-back
    c.morton_temp = o; o += align64(n);                                                // sort ping-pong keys
    c.sort_temp = o; o += align64(n);                                                  // sort ping-pong values
    c.flags = o; o += align64(n);                                                      // refit arrival counters
    c.bounds = o; o += align64(6 * ((n + sv::BOUNDS_TILE - 1) / sv::BOUNDS_TILE));     // bounds partials
    c.total = o;
    return c;
}
constexpr size_t SCRATCH_WORDS = carve(SV_NOALLOC_MAX_BODIES).total;
static_assert(SCRATCH_WORDS * sizeof(uint32_t) <= (16u << 20), "the static scratch stays below 16 MiB");

// The scratch.  In host code `scratch` names a host-side stand-in, not the
// device array: host code uses only the per-device address that
// hipGetSymbolAddress returns.
static __device__ __align__(256) uint32_t scratch[SCRATCH_WORDS];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
uint32_t *g_scratch[MAX_DEVICES] = {};  // device address of `scratch` on each device

int fail(int code, const char *fmt, unsigned a, unsigned b) {
    char msg[256];
    snprintf(msg, sizeof(msg), fmt, a, b);
    return sv::set_error(code, msg);
}

// Address of this device's copy of `scratch`, looked up once per device on
// the first call made with that device current.
int device_scratch(uint32_t **out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return sv::set_error(SV_E_DEVICE, "sv_bvh_build_bodies_noalloc: no current HIP device");
    if (dev < 0 || dev >= MAX_DEVICES)
        return fail(SV_E_DEVICE, "sv_bvh_build_bodies_noalloc: device %u beyond the %u supported", (unsigned)dev,
                    (unsigned)MAX_DEVICES);
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(SV_E_DEVICE, "sv_bvh_build_bodies_noalloc: hipGetSymbolAddress failed (%u) on device %u",
                        (unsigned)e, (unsigned)dev);
        g_scratch[dev] = static_cast<uint32_t *>(p);
    }
    *out = g_scratch[dev];
    return SV_OK;
}

}  // namespace

extern "C" int sv_bvh_build_bodies_noalloc(const float *aabb_min, const float *aabb_max, uint32_t n,
                                           const sv_bvh_arrays *o, void *stream) {
    if (!o) return sv::set_error(SV_E_INVALID, "sv_bvh_build_bodies_noalloc: null output arrays");
    if (n == 0) return SV_OK;
    if (n > SV_NOALLOC_MAX_BODIES)
        return fail(SV_E_CAPACITY, "sv_bvh_build_bodies_noalloc: %u bodies exceed the maximum of %u", n,
                    SV_NOALLOC_MAX_BODIES);
    if (!aabb_min || !aabb_max || !o->centroids || !o->morton || !o->sort_indices || !o->prim_indices ||
        !o->scene_bounds || !o->node_min || !o->node_max || !o->left || !o->right || !o->count || !o->parent)
        return sv::set_error(SV_E_INVALID, "sv_bvh_build_bodies_noalloc: null array");
    uint32_t *base = nullptr;
    int rc = device_scratch(&base);
    if (rc) return rc;
    const Carve cv = carve(n);
    sv_ctx v;  // view context: nothing of it is allocated or freed
    v.stream = static_cast<hipStream_t>(stream);
    v.scratch_fixed = true;
    v.buf[SV_BODY_AABB_MIN] = const_cast<float *>(aabb_min);
    v.buf[SV_BODY_AABB_MAX] = const_cast<float *>(aabb_max);
    v.buf[SV_BVH_CENTROIDS] = o->centroids;
    v.buf[SV_BVH_MORTON] = o->morton;
    v.buf[SV_BVH_SORT_INDICES] = o->sort_indices;
    v.buf[SV_BVH_PRIM_INDICES] = o->prim_indices;
    v.buf[SV_BVH_SCENE_BOUNDS] = o->scene_bounds;
    v.buf[SV_BVH_AABB_MIN] = o->node_min;
    v.buf[SV_BVH_AABB_MAX] = o->node_max;
    v.buf[SV_BVH_LEFT] = o->left;
    v.buf[SV_BVH_RIGHT] = o->right;
    v.buf[SV_BVH_COUNT] = o->count;
    v.buf[SV_BVH_PARENT] = o->parent;
    v.buf[SV_BVH_MORTON_TEMP] = base + cv.morton_temp;
    v.buf[SV_BVH_SORT_TEMP] = base + cv.sort_temp;
    v.buf[SV_BVH_FLAGS] = base + cv.flags;  // cleared by the build's step 6
    v.buf[SV_BVH_TEMP] = base + cv.bounds;
    v.s.sort_hist = base + cv.hist;  // zero at load; every sort leaves its control words at zero
    v.s.sort_hist_cap = (uint32_t)sv::sort_hist_words(SV_NOALLOC_MAX_BODIES);
    return sv::bvh_build_run(&v, n, true);
}
