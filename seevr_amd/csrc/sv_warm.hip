// sv_warm.hip - contact persistence and warm start, an extension beside the
// reference path (the reference clears every impulse each substep,
// physics.comp:494-497, and lists warm start as a to-do, todo warmup.md:1-3).
//
// There is one contact per body pair (physics.comp:475-491) and contacts
// come in the canonical pair order, so the (geom a, geom b) keys of a
// contact list are strictly increasing.  After each sweep the keys and
// impulses are kept; at the next substep's step 4 each new contact finds its
// key by binary search (lower bound: the first cached entry not below the
// key) and starts from warm_factor times the cached impulse, or 0.  Before
// the sweep every such impulse is applied to its two bodies, in contact
// order (the serial form: for c in order, v_a -= m_a^-1 l n, w_a -= I_a^-1
// (r_a x l n), v_b += ..., w_b += ...).  A body's updates depend only on
// constants of its own contacts, so one lane per body walking the body's
// incidences in contact order gives the serial result bit for bit.  (An
// impulse applied at the contact's own first update instead would change
// nothing: that update solves the contact exactly and cancels it.)  With
// warm start off nothing here runs.
#include "sv_internal.h"
#include "sv_math.h"
#include "sv_dev.h"

namespace sv {

namespace {

constexpr int BLOCK = 256;

template <typename T>
inline T *B(sv_ctx *c, int id) { return reinterpret_cast<T *>(c->buf[id]); }

SVD unsigned long long pair_key(uint32_t ga, uint32_t gb) {
    return ((unsigned long long)ga << 32) | (unsigned long long)gb;
}

__global__ void __launch_bounds__(BLOCK) k_warm_lookup(const uint32_t *__restrict__ counters,
                                                       const uint32_t *__restrict__ misc,
                                                       const uint32_t *__restrict__ cga,
                                                       const uint32_t *__restrict__ cgb,
                                                       const unsigned long long *__restrict__ keys,
                                                       const float *__restrict__ lam_prev, float factor,
                                                       uint32_t max_contacts, float *__restrict__ lambda) {
    const uint32_t nc = counters[1] < max_contacts ? counters[1] : max_contacts;
    const uint32_t np = misc[MISC_WARM_N];
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < max_contacts; c += gridDim.x * blockDim.x) {
        float v = 0.0f;
        if (c < nc && np) {
            const unsigned long long k = pair_key(cga[c], cgb[c]);
            uint32_t lo = 0, hi = np;
            while (lo < hi) {
                const uint32_t mid = lo + (hi - lo) / 2;
                if (keys[mid] < k) lo = mid + 1;
                else hi = mid;
            }
            if (lo < np && keys[lo] == k) v = factor * lam_prev[lo];
        }
        lambda[c] = v;
    }
}

__global__ void __launch_bounds__(BLOCK) k_warm_save(const uint32_t *__restrict__ counters,
                                                     const uint32_t *__restrict__ cga,
                                                     const uint32_t *__restrict__ cgb,
                                                     const float *__restrict__ lambda,
                                                     unsigned long long *__restrict__ keys,
                                                     float *__restrict__ lam_out, uint32_t max_contacts,
                                                     uint32_t *misc_w) {
    const uint32_t nc = counters[1] < max_contacts ? counters[1] : max_contacts;
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += gridDim.x * blockDim.x) {
        keys[c] = pair_key(cga[c], cgb[c]);
        lam_out[c] = lambda[c];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) misc_w[MISC_WARM_N] = nc;
}

__global__ void __launch_bounds__(BLOCK) k_warm_apply(uint32_t n_bodies, const uint32_t *__restrict__ sval,
                                                      const uint32_t *__restrict__ deg_off,
                                                      const uint32_t *__restrict__ deg,
                                                      const float *__restrict__ prep,
                                                      const float *__restrict__ lambda, float *vel, float *omega,
                                                      const float *__restrict__ rot,
                                                      const float *__restrict__ inv_i) {
    for (uint32_t x = blockIdx.x * blockDim.x + threadIdx.x; x < n_bodies; x += gridDim.x * blockDim.x) {
        const uint32_t p0 = deg_off[x], p1 = p0 + deg[x];
        if (p0 == p1) continue;
        f3 v = ld3(vel, x), w = ld3(omega, x);
        const m3 R = ld_rot(rot, x);
        const f3 ii = ld3(inv_i, x);
        bool changed = false;
        for (uint32_t p = p0; p < p1; p++) {
            const uint32_t e = sval[p], c = e >> 1;
            const float *P = prep + (size_t)c * CPREP_WORDS;
            if (!(__float_as_uint(P[P_FLAGS]) & F_WARM)) continue;
            const f3 imp = lambda[c] * mk(P[P_N], P[P_N + 1], P[P_N + 2]);
            if (e & 1u) {
                const f3 r = mk(P[P_RB], P[P_RB + 1], P[P_RB + 2]);
                v = v + P[P_IMB] * imp;
                w = w + apply_inv_I(R, ii, cross3(r, imp));
            } else {
                const f3 r = mk(P[P_RA], P[P_RA + 1], P[P_RA + 2]);
                v = v - P[P_IMA] * imp;
                w = w - apply_inv_I(R, ii, cross3(r, imp));
            }
            changed = true;
        }
        if (changed) {
            st3(vel, x, v);
            st3(omega, x, w);
        }
    }
}

uint32_t grid_for(uint32_t n) {
    uint32_t g = (n + BLOCK - 1) / BLOCK;
    return g == 0 ? 1u : (g > 8192u ? 8192u : g);
}

}  // namespace

bool warm_reserve(sv_ctx *c) {
    Scratch &s = c->s;
    if (!s.warm_key) {
        size_t n = c->caps.max_contacts ? c->caps.max_contacts : 1u;
        void *k = nullptr, *l = nullptr;
        if (hipMalloc(&k, n * sizeof(unsigned long long)) != hipSuccess) return false;
        if (hipMalloc(&l, n * sizeof(float)) != hipSuccess) {
            (void)hipFree(k);
            return false;
        }
        s.lazy.push_back(k);
        s.lazy.push_back(l);
        s.warm_key = static_cast<unsigned long long *>(k);
        s.warm_lam = static_cast<float *>(l);
    }
    fill_u32(c, s.misc + MISC_WARM_N, 0u, 1);
    return true;
}

void launch_warm_lookup(sv_ctx *c) {
    const uint32_t mc = c->caps.max_contacts;
    if (!mc) return;
    hipLaunchKernelGGL(k_warm_lookup, dim3(grid_for(mc)), dim3(BLOCK), 0, c->stream,
                       B<uint32_t>(c, SV_ATOMIC_COUNTERS), c->s.misc, B<uint32_t>(c, SV_CONTACT_GEOM_A),
                       B<uint32_t>(c, SV_CONTACT_GEOM_B), c->s.warm_key, c->s.warm_lam, c->warm_factor, mc,
                       B<float>(c, SV_CONTACT_LAMBDA_N));
}

void launch_warm_apply(sv_ctx *c, uint32_t n_bodies, const uint32_t *sval) {
    if (!n_bodies) return;
    hipLaunchKernelGGL(k_warm_apply, dim3(grid_for(n_bodies)), dim3(BLOCK), 0, c->stream, n_bodies, sval, c->s.deg_off,
                       c->s.deg, c->s.cprep, B<float>(c, SV_CONTACT_LAMBDA_N), B<float>(c, SV_BODY_VEL),
                       B<float>(c, SV_BODY_OMEGA), c->s.rot, B<float>(c, SV_BODY_INV_INERTIA));
}

void launch_warm_save(sv_ctx *c) {
    const uint32_t mc = c->caps.max_contacts;
    if (!mc) return;
    hipLaunchKernelGGL(k_warm_save, dim3(grid_for(mc)), dim3(BLOCK), 0, c->stream,
                       B<uint32_t>(c, SV_ATOMIC_COUNTERS), B<uint32_t>(c, SV_CONTACT_GEOM_A),
                       B<uint32_t>(c, SV_CONTACT_GEOM_B), B<float>(c, SV_CONTACT_LAMBDA_N), c->s.warm_key,
                       c->s.warm_lam, mc, c->s.misc);
}

}  // namespace sv
