// sv_diff.hip - differentiable form of the batched engine: gradients of a
// multi-frame rollout with respect to the initial state and the per-geom
// friction and restitution.
//
// The differentiated map is the substep of physics.comp with collision
// detection held fixed: the contacts found by the forward run (points,
// normals, depths, bodies) are constants of each substep.  Forward: every
// frame start is checkpointed.  Backward, per frame from the last: restore
// the frame's checkpoint, recompute its substeps with the batched kernel in
// its recording form (sweep constants, the inputs of every sweep item, the
// angular velocity after the sweep), then run the adjoint kernel over the
// substeps in reverse.  One workgroup per scene reverses a substep:
//   integrate   p3 = p0 + v2 dt, q3 = normalize(q0 + dt/2 (w2, 0) q0)
//   sweep       every item (t, c) in reverse sweep order, with per-body
//               reverse tickets in LDS (the ordered dataflow run backwards);
//               the item adjoint recomputes the item from its recorded
//               inputs and propagates to the bodies' velocity adjoints, the
//               previous impulse of the contact and the contact constants
//               (r_a, r_b, k_n, bias, mu) and the bodies' rotations.  This is
//               the exact derivative of the iters passes run; at a converged
//               sweep it is the implicit-function adjoint of the Gauss-Seidel
//               fixed point, truncated after iters transposed passes
//   constants   bias (restitution, pre-sweep velocities), k_n (positions and
//               rotations), mu (friction), r = p_contact - p_body; per-body
//               sums are gathered in the sorted incidence order (so they are
//               deterministic); material sums use atomics
//   rotations   R = quat_mat(q0) -> quaternion adjoint
//   gravity     v1 = v0 ld + g dt, w1 = w0 ad
// The FP64 CPU restatement of the same adjoint is the checker.
#include "sv_batch.h"

#include <new>

namespace sv {

namespace {

constexpr uint32_t ACC = 9;       // per-contact constant adjoints: ra 3, rb 3, kn, bias, mu
constexpr uint32_t DC_MAX = 512;  // contacts per scene in the differentiable form (LDS tile)
constexpr uint32_t DINC_MAX = 2 * DC_MAX;
constexpr uint32_t CONTRIB = 18;  // per incidence: p 3, v 3, w 3, R 9

struct AdjArgs {
    const uint32_t *sc_body;
    const float *sc_grav;
    uint32_t num_scenes;
    uint32_t *queue;
    uint32_t *abort_flag;
    uint32_t *sc_cnt;
    const float *state;  // [13N] at the substep start: pos 3N | quat 4N | vel 3N | omega 3N
    const float *w2;     // [3N]
    const float *prep;   // [S * cs][CPREP_WORDS]
    const float *items;  // [S][iters * cs][TAPE_ITEM]
    const uint32_t *nc;  // [S]
    const float *inv_mass, *inv_i, *fric, *rest;
    float *adj;          // [13N] in: after the substep, out: before it
    float *afric, *arest;
    float *contrib;
    uint32_t n, cs, iters;
    float dt, ld, ad;
};

SVD f3 inv_I3(const m3 &R, f3 d, f3 v) { return mvec(R, hmul(d, mtvec(R, v))); }
SVD void add_outer(m3 &A, f3 u, f3 w, float s) {
    A.c0 = A.c0 + (s * w.x) * u;
    A.c1 = A.c1 + (s * w.y) * u;
    A.c2 = A.c2 + (s * w.z) * u;
}
// adjoint of R in y = R diag(d) R^T x for the output adjoint yb, times s
SVD void inv_I_rbar(m3 &Rb, const m3 &R, f3 d, f3 x, f3 yb, float s) {
    add_outer(Rb, yb, hmul(d, mtvec(R, x)), s);
    add_outer(Rb, x, hmul(d, mtvec(R, yb)), s);
}
SVD m3 m3_zero() {
    m3 m;
    m.c0 = m.c1 = m.c2 = mk(0.0f, 0.0f, 0.0f);
    return m;
}

struct ItemAdj {
    f3 la, wa, lb, wb, ra, rb;
    float old, kn, bias, mu;
    m3 Ra, Rb;
};

// Adjoint of apply_item (physics.comp:525-580) at recorded inputs.
SVD void item_vjp(const Item &k, f3 la, f3 wa, f3 lb, f3 wb, float old, f3 gla2, f3 gwa2, f3 glb2, f3 gwb2,
                  float gnl, ItemAdj &o) {
    o.ra = o.rb = mk(0.0f, 0.0f, 0.0f);
    o.kn = o.bias = o.mu = 0.0f;
    o.Ra = o.Rb = m3_zero();
    const f3 va = la + cross3(wa, k.ra), vb = lb + cross3(wb, k.rb);
    const float vn = dot3(vb - va, k.n);
    const float dln = (-vn + k.bias) / k.kn;
    const float x = old + dln;
    const float nl = gmax(x, 0.0f);
    const float d2 = nl - old;
    const f3 imp = d2 * k.n;
    const f3 la1 = la - k.ima * imp, lb1 = lb + k.imb * imp;
    const f3 ua = cross3(k.ra, imp), ub = cross3(k.rb, imp);
    const f3 wa1 = wa - inv_I3(k.Ra, k.iia, ua), wb1 = wb + inv_I3(k.Rb, k.iib, ub);
    const f3 va2 = la1 + cross3(wa1, k.ra), vb2 = lb1 + cross3(wb1, k.rb);
    const f3 dv = vb2 - va2;
    const f3 vt = dv - dot3(dv, k.n) * k.n;
    const float vtl = len3(vt);
    f3 gla1 = gla2, gwa1 = gwa2, glb1 = glb2, gwb1 = gwb2, gdv = mk(0.0f, 0.0f, 0.0f);
    float gnl_acc = gnl;
    if (vtl > 1e-6f) {
        const f3 td = mk(vt.x / vtl, vt.y / vtl, vt.z / vtl);
        const f3 rat = cross3(k.ra, td), rbt = cross3(k.rb, td);
        const f3 Irat = inv_I3(k.Ra, k.iia, rat), Irbt = inv_I3(k.Rb, k.iib, rbt);
        const float kt = k.ima + k.imb + dot3(rat, Irat) + dot3(rbt, Irbt);
        const float s1 = vtl / kt, s2 = k.mu * nl;
        const float dlt = gmin(s1, s2);
        const f3 impt = dlt * td;
        const f3 ta = cross3(k.ra, impt), tb = cross3(k.rb, impt);
        const f3 gta = inv_I3(k.Ra, k.iia, gwa2);
        inv_I_rbar(o.Ra, k.Ra, k.iia, ta, gwa2, 1.0f);
        f3 gimpt = cross3(gta, k.ra);
        o.ra = o.ra + cross3(impt, gta);
        const f3 gtb = -inv_I3(k.Rb, k.iib, gwb2);
        inv_I_rbar(o.Rb, k.Rb, k.iib, tb, gwb2, -1.0f);
        gimpt = gimpt + cross3(gtb, k.rb);
        o.rb = o.rb + cross3(impt, gtb);
        gimpt = (gimpt + k.ima * gla2) - k.imb * glb2;
        const float gdlt = dot3(gimpt, td);
        f3 gtd = dlt * gimpt;
        float gs1 = 0.0f, gs2 = 0.0f;
        if (s2 < s1) gs2 = gdlt; else gs1 = gdlt;
        o.mu += gs2 * nl;
        gnl_acc += gs2 * k.mu;
        float gvtl = gs1 / kt;
        const float gkt = -gs1 * s1 / kt;
        const f3 grat = (2.0f * gkt) * Irat, grbt = (2.0f * gkt) * Irbt;
        inv_I_rbar(o.Ra, k.Ra, k.iia, rat, rat, gkt);
        inv_I_rbar(o.Rb, k.Rb, k.iib, rbt, rbt, gkt);
        o.ra = o.ra + cross3(td, grat);
        gtd = gtd + cross3(grat, k.ra);
        o.rb = o.rb + cross3(td, grbt);
        gtd = gtd + cross3(grbt, k.rb);
        f3 gvt = (1.0f / vtl) * gtd;
        gvtl += -dot3(gtd, vt) / (vtl * vtl);
        gvt = gvt + (gvtl / vtl) * vt;
        gdv = gvt - dot3(gvt, k.n) * k.n;
    }
    const f3 gvb2 = gdv, gva2 = -gdv;
    gla1 = gla1 + gva2;
    gwa1 = gwa1 + cross3(k.ra, gva2);
    o.ra = o.ra + cross3(gva2, wa1);
    glb1 = glb1 + gvb2;
    gwb1 = gwb1 + cross3(k.rb, gvb2);
    o.rb = o.rb + cross3(gvb2, wb1);
    f3 gwa = gwa1;
    const f3 gua = -inv_I3(k.Ra, k.iia, gwa1);
    inv_I_rbar(o.Ra, k.Ra, k.iia, ua, gwa1, -1.0f);
    f3 gimp = cross3(gua, k.ra);
    o.ra = o.ra + cross3(imp, gua);
    f3 gwb = gwb1;
    const f3 gub = inv_I3(k.Rb, k.iib, gwb1);
    inv_I_rbar(o.Rb, k.Rb, k.iib, ub, gwb1, 1.0f);
    gimp = gimp + cross3(gub, k.rb);
    o.rb = o.rb + cross3(imp, gub);
    f3 gla = gla1, glb = glb1;
    gimp = (gimp - k.ima * gla1) + k.imb * glb1;
    const float gd2 = dot3(gimp, k.n);
    gnl_acc += gd2;
    float gold = -gd2;
    const float gx = x < 0.0f ? 0.0f : gnl_acc;
    gold += gx;
    const float gdln = gx;
    o.bias += gdln / k.kn;
    const float gvn = -gdln / k.kn;
    o.kn += -gdln * dln / k.kn;
    const f3 gvb = gvn * k.n, gva = -gvn * k.n;
    gla = gla + gva;
    gwa = gwa + cross3(k.ra, gva);
    o.ra = o.ra + cross3(gva, wa);
    glb = glb + gvb;
    gwb = gwb + cross3(k.rb, gvb);
    o.rb = o.rb + cross3(gvb, wb);
    o.la = gla;
    o.wa = gwa;
    o.lb = glb;
    o.wb = gwb;
    o.old = gold;
}

// quaternion adjoint from the adjoint A of R = quat_mat(q) (physics.comp:74-81)
SVD void quat_mat_vjp(const float *q, const m3 &A, float *qb) {
    const float x = q[0], y = q[1], z = q[2], w = q[3];
    qb[0] += A.c0.y * 2.0f * y + A.c0.z * 2.0f * z + A.c1.x * 2.0f * y - A.c1.y * 4.0f * x + A.c1.z * 2.0f * w +
             A.c2.x * 2.0f * z - A.c2.y * 2.0f * w - A.c2.z * 4.0f * x;
    qb[1] += -A.c0.x * 4.0f * y + A.c0.y * 2.0f * x - A.c0.z * 2.0f * w + A.c1.x * 2.0f * x + A.c1.z * 2.0f * z +
             A.c2.x * 2.0f * w + A.c2.y * 2.0f * z - A.c2.z * 4.0f * y;
    qb[2] += -A.c0.x * 4.0f * z + A.c0.y * 2.0f * w + A.c0.z * 2.0f * x - A.c1.x * 2.0f * w - A.c1.y * 4.0f * z +
             A.c1.z * 2.0f * y + A.c2.x * 2.0f * x + A.c2.y * 2.0f * y;
    qb[3] += A.c0.y * 2.0f * z - A.c0.z * 2.0f * y - A.c1.x * 2.0f * z + A.c1.z * 2.0f * x + A.c2.x * 2.0f * y -
             A.c2.y * 2.0f * x;
}

SVD void acc_rot(float *r9, const m3 &A) {
    r9[0] += A.c0.x; r9[1] += A.c0.y; r9[2] += A.c0.z;
    r9[3] += A.c1.x; r9[4] += A.c1.y; r9[5] += A.c1.z;
    r9[6] += A.c2.x; r9[7] += A.c2.y; r9[8] += A.c2.z;
}
SVD m3 ld_m3(const float *r9) {
    m3 m;
    m.c0 = mk(r9[0], r9[1], r9[2]);
    m.c1 = mk(r9[3], r9[4], r9[5]);
    m.c2 = mk(r9[6], r9[7], r9[8]);
    return m;
}

template <int T>
__global__ void __launch_bounds__(T) k_batch_adjoint(AdjArgs A) {
    __shared__ float s_pb[3 * NB_MAX], s_qb[4 * NB_MAX], s_vb[3 * NB_MAX], s_wb[3 * NB_MAX];
    __shared__ float s_Rb[9 * NB_MAX];
    __shared__ float s_v1[3 * NB_MAX], s_w1[3 * NB_MAX];
    __shared__ uint32_t s_key[DINC_MAX];
    __shared__ uint16_t s_rank[DINC_MAX];
    __shared__ uint16_t s_start[NB_MAX], s_deg[NB_MAX];
    __shared__ uint32_t s_rt[NB_MAX];
    __shared__ float s_lb[DC_MAX];
    __shared__ float s_acc[DC_MAX * ACC];
    __shared__ uint32_t s_q, s_next;
    const int tid = threadIdx.x, lane = tid & 63;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const size_t N = A.n;
    const float *pos0 = A.state, *quat0 = A.state + 3 * N, *vel0 = A.state + 7 * N, *omg0 = A.state + 10 * N;
    float *ap = A.adj, *aq = A.adj + 3 * N, *av = A.adj + 7 * N, *aw = A.adj + 10 * N;
    const float dt = A.dt;

    for (;;) {
        if (tid == 0) {
            uint32_t q = atomicAdd(A.queue, 1u);
            if (__hip_atomic_load(A.abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) q = A.num_scenes;
            s_q = q;
        }
        __syncthreads();
        const uint32_t sc = s_q;
        __syncthreads();
        if (sc >= A.num_scenes) return;
        const uint32_t bs = A.sc_body[sc], ns = A.sc_body[sc + 1] - bs;
        const float gx = A.sc_grav[3 * sc], gy = A.sc_grav[3 * sc + 1], gz = A.sc_grav[3 * sc + 2];
        const uint32_t nc = A.nc[sc];
        const size_t cseg = (size_t)sc * A.cs;

        // post-gravity velocities, and the integrate adjoint
        for (uint32_t i = tid; i < ns; i += T) {
            const uint32_t b = bs + i;
            const bool dyn = A.inv_mass[b] > 0.0f;
            f3 v0 = ld3(vel0, b), w0 = ld3(omg0, b);
            if (dyn) {
                v0 = mk(v0.x * A.ld + gx * dt, v0.y * A.ld + gy * dt, v0.z * A.ld + gz * dt);
                w0 = mk(w0.x * A.ad, w0.y * A.ad, w0.z * A.ad);
            }
            st3(s_v1, i, v0);
            st3(s_w1, i, w0);
            f3 gp = ld3(ap, b), gv = ld3(av, b), gw = ld3(aw, b);
            float g[4] = {aq[4 * b], aq[4 * b + 1], aq[4 * b + 2], aq[4 * b + 3]};
            if (dyn) {
                gv = gv + dt * gp;
                const float *q = quat0 + 4 * (size_t)b;
                const float ox = A.w2[3 * b], oy = A.w2[3 * b + 1], oz = A.w2[3 * b + 2];
                const float qx = q[0], qy = q[1], qz = q[2], qw = q[3];
                const float dx = ox * qw + oy * qz - oz * qy;
                const float dy = -ox * qz + oy * qw + oz * qx;
                const float dz = ox * qy - oy * qx + oz * qw;
                const float dw = -ox * qx - oy * qy - oz * qz;
                const float px = qx + 0.5f * dx * dt, py = qy + 0.5f * dy * dt, pz = qz + 0.5f * dz * dt,
                            pw = qw + 0.5f * dw * dt;
                const float l = __builtin_sqrtf(px * px + py * py + pz * pz + pw * pw);
                if (l > 1e-8f) {
                    const float ux = px / l, uy = py / l, uz = pz / l, uw = pw / l;
                    const float d = g[0] * ux + g[1] * uy + g[2] * uz + g[3] * uw;
                    g[0] = (g[0] - d * ux) / l;
                    g[1] = (g[1] - d * uy) / l;
                    g[2] = (g[2] - d * uz) / l;
                    g[3] = (g[3] - d * uw) / l;
                }
                const float hx = 0.5f * dt * g[0], hy = 0.5f * dt * g[1], hz = 0.5f * dt * g[2], hw = 0.5f * dt * g[3];
                gw = gw + mk(hx * qw - hy * qz + hz * qy - hw * qx, hx * qz + hy * qw - hz * qx - hw * qy,
                             -hx * qy + hy * qx + hz * qw - hw * qz);
                const float n0 = g[0] + hy * oz - hz * oy - hw * ox;
                const float n1 = g[1] - hx * oz + hz * ox - hw * oy;
                const float n2 = g[2] + hx * oy - hy * ox - hw * oz;
                const float n3 = g[3] + hx * ox + hy * oy + hz * oz;
                g[0] = n0; g[1] = n1; g[2] = n2; g[3] = n3;
            }
            st3(s_pb, i, gp);
            st3(s_vb, i, gv);
            st3(s_wb, i, gw);
#pragma unroll
            for (int k = 0; k < 4; k++) s_qb[4 * i + k] = g[k];
#pragma unroll
            for (int k = 0; k < 9; k++) s_Rb[9 * i + k] = 0.0f;
            s_rt[i] = 0u;
            s_deg[i] = 0;  // bodies without incidences keep 0
        }
        // incidence ranks, as in the forward sweep
        for (uint32_t c = tid; c < nc; c += T) {
            const float *P = A.prep + (cseg + c) * CPREP_WORDS;
            const uint32_t flags = __float_as_uint(P[P_FLAGS]);
            const uint32_t a = __float_as_uint(P[P_A]), b = __float_as_uint(P[P_B]);
            const bool live = !(flags & F_SKIP);
            s_key[2 * c] = (live && (flags & F_DYN_A)) ? (((a - bs) << 12) | (2 * c)) : 0xFFFFFFFFu;
            s_key[2 * c + 1] = (live && (flags & F_DYN_B)) ? (((b - bs) << 12) | (2 * c + 1)) : 0xFFFFFFFFu;
            s_lb[c] = 0.0f;
#pragma unroll
            for (int k = 0; k < (int)ACC; k++) s_acc[c * ACC + k] = 0.0f;
        }
        uint32_t L = 2;
        while (L < 2 * nc) L <<= 1;
        for (uint32_t k = 2 * nc + tid; k < L; k += T) s_key[k] = 0xFFFFFFFFu;
        if (tid == 0) s_next = 0u;
        __syncthreads();
        rank_incidences<T>(s_key, s_rank, s_start, s_deg, L);

        // reverse sweep
        const uint32_t total = A.iters * nc;
        {
            Item it;
            bool has = false, fin = total == 0;
            uint32_t ra_pos = 0, rb_pos = 0;
            f3 t_la, t_wa, t_lb, t_wb;
            float t_old = 0.0f;
            uint64_t last = wall_clock64();
            for (;;) {
                uint64_t need = __ballot(!has && !fin);
                while (need) {
                    const uint32_t cntw = (uint32_t)__popcll(need);
                    const int leader = __builtin_ctzll(need);
                    uint32_t base = 0;
                    if (lane == leader)
                        base = __hip_atomic_fetch_add(&s_next, cntw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    base = __shfl(base, leader, 64);
                    if (!has && !fin) {
                        const uint32_t r = base + (uint32_t)__popcll(need & lt);
                        if (r >= total) {
                            fin = true;
                        } else {
                            const uint32_t q2 = total - 1 - r;
                            const uint32_t t = q2 / nc, c = q2 - t * nc;
                            const float *P = A.prep + (cseg + c) * CPREP_WORDS;
                            const uint32_t flags = __float_as_uint(P[P_FLAGS]);
                            if (!(flags & F_SKIP)) {
                                const uint32_t ga = __float_as_uint(P[P_A]), gb = __float_as_uint(P[P_B]);
                                it.c = c;
                                it.q = q2;
                                it.a = ga - bs;
                                it.b = gb - bs;
                                it.dyn_a = flags & F_DYN_A;
                                it.dyn_b = flags & F_DYN_B;
                                if (it.dyn_a)
                                    ra_pos = A.iters * s_deg[it.a] - 1u - (t * s_deg[it.a] + s_rank[2 * c]);
                                if (it.dyn_b)
                                    rb_pos = A.iters * s_deg[it.b] - 1u - (t * s_deg[it.b] + s_rank[2 * c + 1]);
                                it.n = mk(P[P_N], P[P_N + 1], P[P_N + 2]);
                                it.ra = mk(P[P_RA], P[P_RA + 1], P[P_RA + 2]);
                                it.rb = mk(P[P_RB], P[P_RB + 1], P[P_RB + 2]);
                                it.kn = P[P_KN];
                                it.bias = P[P_BIAS];
                                it.mu = P[P_MU];
                                it.ima = P[P_IMA];
                                it.imb = P[P_IMB];
                                it.Ra = quat_mat_at(quat0, ga);
                                it.Rb = quat_mat_at(quat0, gb);
                                it.iia = ld3(A.inv_i, ga);
                                it.iib = ld3(A.inv_i, gb);
                                const float *tp = A.items + ((size_t)sc * A.iters * A.cs + q2) * TAPE_ITEM;
                                t_la = ld3(tp, 0);
                                t_wa = ld3(tp, 1);
                                t_lb = ld3(tp, 2);
                                t_wb = ld3(tp, 3);
                                t_old = tp[12];
                                has = true;
                            }
                        }
                    }
                    need = __ballot(!has && !fin);
                }
                if (!__any(has)) break;
                bool progressed = false;
                if (has) {
                    const bool ready = (!it.dyn_a || __hip_atomic_load(&s_rt[it.a], __ATOMIC_RELAXED,
                                                                       __HIP_MEMORY_SCOPE_WORKGROUP) == ra_pos) &&
                                       (!it.dyn_b || __hip_atomic_load(&s_rt[it.b], __ATOMIC_RELAXED,
                                                                       __HIP_MEMORY_SCOPE_WORKGROUP) == rb_pos);
                    if (ready) {
                        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                        const f3 z = mk(0.0f, 0.0f, 0.0f);
                        ItemAdj o;
                        item_vjp(it, t_la, t_wa, t_lb, t_wb, t_old,
                                 it.dyn_a ? ld3(s_vb, it.a) : z, it.dyn_a ? ld3(s_wb, it.a) : z,
                                 it.dyn_b ? ld3(s_vb, it.b) : z, it.dyn_b ? ld3(s_wb, it.b) : z, s_lb[it.c], o);
                        if (it.dyn_a) {
                            st3(s_vb, it.a, o.la);
                            st3(s_wb, it.a, o.wa);
                            acc_rot(s_Rb + 9 * it.a, o.Ra);
                        }
                        if (it.dyn_b) {
                            st3(s_vb, it.b, o.lb);
                            st3(s_wb, it.b, o.wb);
                            acc_rot(s_Rb + 9 * it.b, o.Rb);
                        }
                        s_lb[it.c] = o.old;
                        float *acc = s_acc + it.c * ACC;
                        acc[0] += o.ra.x; acc[1] += o.ra.y; acc[2] += o.ra.z;
                        acc[3] += o.rb.x; acc[4] += o.rb.y; acc[5] += o.rb.z;
                        acc[6] += o.kn; acc[7] += o.bias; acc[8] += o.mu;
                        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                        if (it.dyn_a)
                            __hip_atomic_fetch_add(&s_rt[it.a], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                        if (it.dyn_b)
                            __hip_atomic_fetch_add(&s_rt[it.b], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                        has = false;
                        progressed = true;
                    }
                }
                if (__any(progressed)) {
                    last = wall_clock64();
                } else {
                    if (__hip_atomic_load(A.abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) break;
                    if (wall_clock64() - last > STALL_TICKS) {
                        __hip_atomic_store(A.abort_flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        if (lane == 0) atomicOr(A.sc_cnt + 3 * sc + 2, (uint32_t)SV_STATUS_SOLVER_TIMEOUT);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
        }
        __syncthreads();

        // contact constants -> per-incidence contributions and materials
        for (uint32_t c = tid; c < nc; c += T) {
            const float *P = A.prep + (cseg + c) * CPREP_WORDS;
            const uint32_t flags = __float_as_uint(P[P_FLAGS]);
            if (flags & F_SKIP) continue;
            const uint32_t ga_b = __float_as_uint(P[P_A]), gb_b = __float_as_uint(P[P_B]);
            const uint32_t gga = __float_as_uint(P[P_GEOM_A]), ggb = __float_as_uint(P[P_GEOM_B]);
            const uint32_t a = ga_b - bs, b = gb_b - bs;
            const float *acc = s_acc + c * ACC;
            f3 gra = mk(acc[0], acc[1], acc[2]), grb = mk(acc[3], acc[4], acc[5]);
            const float gkn = acc[6], gbias = acc[7], gmu = acc[8];
            const f3 n = mk(P[P_N], P[P_N + 1], P[P_N + 2]);
            const f3 ra = mk(P[P_RA], P[P_RA + 1], P[P_RA + 2]), rb = mk(P[P_RB], P[P_RB + 1], P[P_RB + 2]);
            const f3 w1a = ld3(s_w1, a), w1b = ld3(s_w1, b);
            const f3 va = ld3(s_v1, a) + cross3(w1a, ra), vb = ld3(s_v1, b) + cross3(w1b, rb);
            const float vn = dot3(vb - va, n);
            const float ea = A.rest[gga], eb = A.rest[ggb];
            const float e = gmax(ea, eb);
            const float term = -vn < 0.0f ? 0.0f : -vn;
            const float ge = gbias * term;
            const float gvn = -vn < 0.0f ? 0.0f : -gbias * e;
            atomicAdd(ea < eb ? A.arest + ggb : A.arest + gga, ge);
            const f3 gvb = gvn * n, gva = -gvn * n;
            const m3 Ra = quat_mat_at(quat0, ga_b), Rb = quat_mat_at(quat0, gb_b);
            const f3 iia = ld3(A.inv_i, ga_b), iib = ld3(A.inv_i, gb_b);
            const f3 ran = cross3(ra, n), rbn = cross3(rb, n);
            const f3 gran = (2.0f * gkn) * inv_I3(Ra, iia, ran), grbn = (2.0f * gkn) * inv_I3(Rb, iib, rbn);
            gra = gra + cross3(gva, w1a) + cross3(n, gran);
            grb = grb + cross3(gvb, w1b) + cross3(n, grbn);
            const float mu = P[P_MU];
            if (mu > 0.0f) {
                atomicAdd(A.afric + gga, gmu * A.fric[ggb] / (2.0f * mu));
                atomicAdd(A.afric + ggb, gmu * A.fric[gga] / (2.0f * mu));
            }
            float *ca = A.contrib + (2 * cseg + 2 * c) * CONTRIB;
            float *cb = ca + CONTRIB;
            m3 Rba = m3_zero(), Rbb = m3_zero();
            inv_I_rbar(Rba, Ra, iia, ran, ran, gkn);
            inv_I_rbar(Rbb, Rb, iib, rbn, rbn, gkn);
            st3(ca, 0, -gra);
            st3(ca, 1, gva);
            st3(ca, 2, cross3(ra, gva));
            st3(cb, 0, -grb);
            st3(cb, 1, gvb);
            st3(cb, 2, cross3(rb, gvb));
#pragma unroll
            for (int k = 0; k < 9; k++) {
                ca[9 + k] = 0.0f;
                cb[9 + k] = 0.0f;
            }
            acc_rot(ca + 9, Rba);
            acc_rot(cb + 9, Rbb);
        }
        __syncthreads();

        // per body: gather in incidence order, rotation and gravity adjoints
        for (uint32_t i = tid; i < ns; i += T) {
            const uint32_t b = bs + i;
            f3 gp = ld3(s_pb, i), gv = ld3(s_vb, i), gw = ld3(s_wb, i);
            float *R9 = s_Rb + 9 * i;
            const uint32_t deg = s_deg[i];
            if (deg) {
                const uint32_t k0 = s_start[i];
                for (uint32_t p = k0; p < k0 + deg; p++) {
                    const uint32_t slot = s_key[p] & 0xFFFu;
                    const float *cc = A.contrib + (2 * cseg + slot) * CONTRIB;
                    gp = gp + ld3(cc, 0);
                    gv = gv + ld3(cc, 1);
                    gw = gw + ld3(cc, 2);
#pragma unroll
                    for (int k = 0; k < 9; k++) R9[k] += cc[9 + k];
                }
            }
            float qb[4] = {s_qb[4 * i], s_qb[4 * i + 1], s_qb[4 * i + 2], s_qb[4 * i + 3]};
            quat_mat_vjp(quat0 + 4 * (size_t)b, ld_m3(R9), qb);
            if (A.inv_mass[b] > 0.0f) {
                gv = A.ld * gv;
                gw = A.ad * gw;
            }
            st3(ap, b, gp);
            st3(av, b, gv);
            st3(aw, b, gw);
#pragma unroll
            for (int k = 0; k < 4; k++) aq[4 * b + k] = qb[k];
        }
        __syncthreads();
    }
}

typedef void (*AdjKernel)(AdjArgs);

AdjKernel adj_kernel(uint32_t threads) {
    switch (threads) {
    case 64: return k_batch_adjoint<64>;
    case 128: return k_batch_adjoint<128>;
    case 256: return k_batch_adjoint<256>;
    case 512: return k_batch_adjoint<512>;
    default: return nullptr;
    }
}

}  // namespace
}  // namespace sv

namespace {

using sv::set_error;

template <typename T>
hipError_t dmalloc(T *&p, size_t count) {
    if (p) (void)hipFree(p);
    p = nullptr;
    return hipMalloc(reinterpret_cast<void **>(&p), (count ? count : 1) * sizeof(T));
}

// state arrays of the batch context, packed as pos | quat | vel | omega
hipError_t copy_state(sv_batch *b, float *packed, bool save) {
    sv_ctx *c = b->ctx;
    const size_t N = b->num_bodies;
    const int ids[4] = {SV_BODY_POS, SV_BODY_QUAT, SV_BODY_VEL, SV_BODY_OMEGA};
    const size_t w[4] = {3, 4, 3, 3};
    size_t off = 0;
    hipError_t e = hipSuccess;
    for (int k = 0; k < 4 && e == hipSuccess; k++) {
        float *dev = reinterpret_cast<float *>(c->buf[ids[k]]);
        e = save ? hipMemcpyAsync(packed + off, dev, w[k] * N * sizeof(float), hipMemcpyDeviceToDevice, c->stream)
                 : hipMemcpyAsync(dev, packed + off, w[k] * N * sizeof(float), hipMemcpyDeviceToDevice, c->stream);
        off += w[k] * N;
    }
    return e;
}

}  // namespace

extern "C" {

int sv_batch_diff_reserve(sv_batch *b, uint32_t max_frames, uint32_t substeps) {
    if (!b) return set_error(SV_E_INVALID, "sv_batch_diff_reserve: null batch");
    if (!max_frames || !substeps || substeps > 64)
        return set_error(SV_E_INVALID, "sv_batch_diff_reserve: frames > 0 and 1-64 substeps");
    if (b->caps.max_scene_contacts > sv::DC_MAX)
        return set_error(SV_E_CAPACITY, "sv_batch_diff_reserve: max_scene_contacts above 512 (the adjoint's LDS tile)");
    const size_t N = b->caps.max_bodies, S = b->caps.max_scenes, cs = b->caps.max_scene_contacts;
    const size_t iters = b->ctx->params.pgs_iterations;
    hipError_t e = dmalloc(b->d_ckpt, (max_frames + 1) * 13 * N);
    // two tape sets: one frame is recomputed while the previous one is reversed
    if (e == hipSuccess) e = dmalloc(b->d_state, 2 * substeps * 13 * N);
    if (e == hipSuccess) e = dmalloc(b->d_w2, 2 * substeps * 3 * N);
    if (e == hipSuccess) e = dmalloc(b->d_prep, 2 * substeps * S * cs * sv::CPREP_WORDS);
    if (e == hipSuccess) e = dmalloc(b->d_items, 2 * substeps * S * iters * cs * sv::TAPE_ITEM);
    if (e == hipSuccess) e = dmalloc(b->d_nc, 2 * substeps * S);
    if (e == hipSuccess && !b->d_stream) e = hipStreamCreateWithFlags(&b->d_stream, hipStreamNonBlocking);
    for (int k = 0; k < 5 && e == hipSuccess; k++)
        if (!b->d_ev[k]) e = hipEventCreateWithFlags(&b->d_ev[k], hipEventDisableTiming);
    if (e == hipSuccess) e = dmalloc(b->d_adj, 13 * N);
    if (e == hipSuccess) e = dmalloc(b->d_contrib, 2 * S * cs * sv::CONTRIB);
    if (e != hipSuccess) {
        b->d_frames = 0;
        return set_error(SV_E_NOMEM, "sv_batch_diff_reserve: device allocation failed");
    }
    b->d_frames = max_frames;
    b->d_substeps = substeps;
    b->d_iters = (uint32_t)iters;
    return SV_OK;
}

int sv_batch_diff_forward(sv_batch *b, uint32_t frames, uint32_t substeps, float dt) {
    if (!b) return set_error(SV_E_INVALID, "sv_batch_diff_forward: null batch");
    if (!b->d_frames || frames > b->d_frames || substeps != b->d_substeps ||
        b->d_iters != b->ctx->params.pgs_iterations)
        return set_error(SV_E_CAPACITY, "sv_batch_diff_forward: reserve frames/substeps/iterations first");
    const size_t N13 = 13 * (size_t)b->num_bodies;
    sv::fill_u32(b->ctx, b->sc_cnt, 0u, 3 * (size_t)b->num_scenes);
    for (uint32_t f = 0; f < frames; f++) {
        hipError_t e = copy_state(b, b->d_ckpt + f * N13, true);
        if (e != hipSuccess) return sv::batch_dev_fail(e, "sv_batch_diff_forward: checkpoint");
        sv::enqueue_substeps(b, substeps, dt / (float)substeps, b->threads, b->wgs_per_cu);
    }
    hipError_t e = copy_state(b, b->d_ckpt + frames * N13, true);
    if (e != hipSuccess) return sv::batch_dev_fail(e, "sv_batch_diff_forward: checkpoint");
    return sv::batch_finish(b);
}

int sv_batch_diff_backward(sv_batch *b, uint32_t frames, uint32_t substeps, float dt, const float *grad_state,
                           float *out_state, float *out_friction, float *out_restitution) {
    if (!b || !out_state || !out_friction || !out_restitution)
        return set_error(SV_E_INVALID, "sv_batch_diff_backward: null argument");
    if (!b->d_frames || frames > b->d_frames || substeps != b->d_substeps ||
        b->d_iters != b->ctx->params.pgs_iterations)
        return set_error(SV_E_CAPACITY, "sv_batch_diff_backward: reserve frames/substeps/iterations first");
    sv_ctx *c = b->ctx;
    hipStream_t st = c->stream;
    const size_t N = b->num_bodies, N13 = 13 * N, S = b->num_scenes, cs = b->caps.max_scene_contacts;
    const size_t iters = b->d_iters;
    const float sub_dt = dt / (float)substeps;
    hipError_t e = hipSuccess;
    if (grad_state) e = hipMemcpyAsync(b->d_adj, grad_state, N13 * sizeof(float), hipMemcpyDefault, st);
    else sv::fill_u32(c, reinterpret_cast<uint32_t *>(b->d_adj), 0u, N13);
    float *afric = nullptr, *arest = nullptr;
    if (e == hipSuccess) e = hipMalloc(&afric, (N ? N : 1) * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&arest, (N ? N : 1) * sizeof(float));
    if (e != hipSuccess) {
        if (afric) (void)hipFree(afric);
        if (arest) (void)hipFree(arest);
        return sv::batch_dev_fail(e, "sv_batch_diff_backward: setup");
    }
    sv::fill_u32(c, reinterpret_cast<uint32_t *>(afric), 0u, N);
    sv::fill_u32(c, reinterpret_cast<uint32_t *>(arest), 0u, N);
    sv::launch_hull_ranges(c, b->num_bodies);
    uint32_t grid = (uint32_t)b->cus * b->wgs_per_cu;
    if (grid > S) grid = (uint32_t)S;
    if (grid == 0) grid = 1;
    // Frame f is recomputed (tape set f & 1) on the context stream while the
    // adjoint of frame f + 1 runs on the second stream; a tape set is
    // rewritten only after the adjoint that read it has finished.
    hipStream_t ast = b->d_stream;
    hipEvent_t *ev_rec = b->d_ev, *ev_adj = b->d_ev + 2, ev_misc = b->d_ev[4];
    if (e == hipSuccess) e = hipEventRecord(ev_misc, st);
    if (e == hipSuccess) e = hipStreamWaitEvent(ast, ev_misc, 0);
    const size_t T_STATE = substeps * N13, T_W2 = substeps * 3 * N, T_PREP = substeps * S * cs * sv::CPREP_WORDS,
                 T_ITEMS = substeps * S * iters * cs * sv::TAPE_ITEM, T_NC = substeps * S;
    for (int64_t f = (int64_t)frames - 1; f >= 0 && e == hipSuccess; f--) {
        const int k = (int)(f & 1);
        float *t_state = b->d_state + k * T_STATE, *t_w2 = b->d_w2 + k * T_W2, *t_prep = b->d_prep + k * T_PREP;
        float *t_items = b->d_items + k * T_ITEMS;
        uint32_t *t_nc = b->d_nc + k * T_NC, *queue = b->queue + (size_t)k * 2 * substeps;
        if (f + 2 <= (int64_t)frames - 1) e = hipStreamWaitEvent(st, ev_adj[k], 0);
        if (e == hipSuccess) e = copy_state(b, b->d_ckpt + (size_t)f * N13, false);
        if (e != hipSuccess) break;
        sv::fill_u32(c, queue, 0u, 2 * (size_t)substeps);
        for (uint32_t s = 0; s < substeps; s++) {
            sv::BatchArgs A = sv::batch_args(b, sub_dt, queue + s);
            A.t_state = t_state + s * N13;
            A.cprep = t_prep + s * S * cs * sv::CPREP_WORDS;
            A.t_items = t_items + s * S * iters * cs * sv::TAPE_ITEM;
            A.t_w2 = t_w2 + s * 3 * N;
            A.t_nc = t_nc + s * S;
            sv::batch_launch_substep(A, b->threads, grid, true, st);
        }
        e = hipEventRecord(ev_rec[k], st);
        if (e == hipSuccess) e = hipStreamWaitEvent(ast, ev_rec[k], 0);
        for (int64_t s = (int64_t)substeps - 1; s >= 0 && e == hipSuccess; s--) {
            sv::AdjArgs J;
            J.sc_body = b->sc_body;
            J.sc_grav = b->sc_grav;
            J.num_scenes = (uint32_t)S;
            J.queue = queue + substeps + s;
            J.abort_flag = b->abort_flag;
            J.sc_cnt = b->sc_cnt;
            J.state = t_state + s * N13;
            J.w2 = t_w2 + s * 3 * N;
            J.prep = t_prep + s * S * cs * sv::CPREP_WORDS;
            J.items = t_items + s * S * iters * cs * sv::TAPE_ITEM;
            J.nc = t_nc + s * S;
            J.inv_mass = reinterpret_cast<const float *>(c->buf[SV_BODY_INV_MASS]);
            J.inv_i = reinterpret_cast<const float *>(c->buf[SV_BODY_INV_INERTIA]);
            J.fric = reinterpret_cast<const float *>(c->buf[SV_GEOM_FRICTION]);
            J.rest = reinterpret_cast<const float *>(c->buf[SV_GEOM_RESTITUTION]);
            J.adj = b->d_adj;
            J.afric = afric;
            J.arest = arest;
            J.contrib = b->d_contrib;
            J.n = (uint32_t)N;
            J.cs = (uint32_t)cs;
            J.iters = (uint32_t)iters;
            J.dt = sub_dt;
            J.ld = c->params.linear_damping;
            J.ad = c->params.angular_damping;
            hipLaunchKernelGGL(sv::adj_kernel(b->threads), dim3(grid), dim3(b->threads), 0, ast, J);
        }
        if (e == hipSuccess) e = hipEventRecord(ev_adj[k], ast);
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out_state, b->d_adj, N13 * sizeof(float), hipMemcpyDefault, ast);
    if (e == hipSuccess) e = hipMemcpyAsync(out_friction, afric, N * sizeof(float), hipMemcpyDefault, ast);
    if (e == hipSuccess) e = hipMemcpyAsync(out_restitution, arest, N * sizeof(float), hipMemcpyDefault, ast);
    if (e == hipSuccess) e = hipEventRecord(ev_misc, ast);
    if (e == hipSuccess) e = hipStreamWaitEvent(st, ev_misc, 0);
    if (e == hipSuccess) e = copy_state(b, b->d_ckpt + (size_t)frames * N13, false);
    int rc = e == hipSuccess ? sv::batch_finish(b) : sv::batch_dev_fail(e, "sv_batch_diff_backward");
    (void)hipStreamSynchronize(ast);
    (void)hipFree(afric);
    (void)hipFree(arest);
    return rc;
}

}  // extern "C"
