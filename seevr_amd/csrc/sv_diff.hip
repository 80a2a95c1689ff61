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
//               deterministic), and so are the friction and restitution
//               adjoints: each incidence carries its own geom's share and
//               the body's thread adds them, in that order, to an FP64 sum
//               kept over the whole backward pass (geoms are 1:1 with
//               bodies, src/data.zig:462-479)
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
constexpr uint32_t CONTRIB = 20;  // per incidence: p 3, v 3, w 3, R 9, own geom's friction and restitution

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
    double *adj;         // [13N] in: after the substep, out: before it
    double *afric, *arest;  // [N] per geom (= body), summed over every substep of the backward pass
    double *contrib;
    uint32_t n, cs, iters;
    float dt, ld, ad;
};

// FP32 R diag(d) R^T v, as the forward sweep evaluates it (apply_inv_I)
SVD f3 inv_I3(const m3 &R, f3 d, f3 v) { return mvec(R, hmul(d, mtvec(R, v))); }

// ---------------------------------------------------------------------------
// Adjoint values are FP64.  The forward quantities an adjoint needs are
// recomputed in FP32 from the recorded inputs, exactly as the forward sweep
// computed them, so every branch (impulse clamp, stick or slip, the tangent
// threshold) is the forward's own; only the reverse accumulation runs in
// double.  The friction and restitution gradients are sums of terms that
// cancel by orders of magnitude (a friction gradient of 1e-5 from terms of
// 1e-2), which FP32 adjoint values cannot resolve to 1e-3.
struct g3 {
    double x, y, z;
};
SVD g3 gk(double x, double y, double z) { g3 r; r.x = x; r.y = y; r.z = z; return r; }
SVD g3 G(f3 a) { return gk(a.x, a.y, a.z); }
SVD g3 gzero() { return gk(0.0, 0.0, 0.0); }
SVD g3 operator+(g3 a, g3 b) { return gk(a.x + b.x, a.y + b.y, a.z + b.z); }
SVD g3 operator-(g3 a, g3 b) { return gk(a.x - b.x, a.y - b.y, a.z - b.z); }
SVD g3 operator-(g3 a) { return gk(-a.x, -a.y, -a.z); }
SVD g3 operator*(double s, g3 a) { return gk(s * a.x, s * a.y, s * a.z); }
SVD double gdot(g3 a, g3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
SVD g3 gcross(g3 a, g3 b) { return gk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
SVD g3 ghmul(f3 a, g3 b) { return gk(a.x * b.x, a.y * b.y, a.z * b.z); }
SVD g3 ldg(const double *p, uint32_t i) { return gk(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
SVD void stg(double *p, uint32_t i, g3 v) { p[3 * i] = v.x; p[3 * i + 1] = v.y; p[3 * i + 2] = v.z; }
// R (an FP32 rotation) times v, and its transpose
SVD g3 gmvec(const m3 &m, g3 v) { return v.x * G(m.c0) + v.y * G(m.c1) + v.z * G(m.c2); }
SVD g3 gmtvec(const m3 &m, g3 v) { return gk(gdot(G(m.c0), v), gdot(G(m.c1), v), gdot(G(m.c2), v)); }
SVD g3 inv_I3g(const m3 &R, f3 d, g3 v) { return gmvec(R, ghmul(d, gmtvec(R, v))); }

struct gm3 {
    g3 c0, c1, c2;
};
SVD gm3 gm3_zero() {
    gm3 m;
    m.c0 = m.c1 = m.c2 = gzero();
    return m;
}
SVD void add_outer(gm3 &A, g3 u, g3 w, double s) {
    A.c0 = A.c0 + (s * w.x) * u;
    A.c1 = A.c1 + (s * w.y) * u;
    A.c2 = A.c2 + (s * w.z) * u;
}
// adjoint of R in y = R diag(d) R^T x for the output adjoint yb, times s
SVD void inv_I_rbar(gm3 &Rb, const m3 &R, f3 d, g3 x, g3 yb, double s) {
    add_outer(Rb, yb, ghmul(d, gmtvec(R, x)), s);
    add_outer(Rb, x, ghmul(d, gmtvec(R, yb)), s);
}

struct ItemAdj {
    g3 la, wa, lb, wb, ra, rb;
    double old, kn, bias, mu;
    gm3 Ra, Rb;
};

// Adjoint of apply_item (physics.comp:525-580) at recorded inputs: the
// forward part in FP32 (the sweep's values), the reverse part in FP64.
SVD void item_vjp(const Item &k, f3 la, f3 wa, f3 lb, f3 wb, float old, g3 gla2, g3 gwa2, g3 glb2, g3 gwb2,
                  double gnl, ItemAdj &o) {
    o.ra = o.rb = gzero();
    o.kn = o.bias = o.mu = 0.0;
    o.Ra = o.Rb = gm3_zero();
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
    const g3 n = G(k.n), ra = G(k.ra), rb = G(k.rb);
    g3 gla1 = gla2, gwa1 = gwa2, glb1 = glb2, gwb1 = gwb2, gdv = gzero();
    double gnl_acc = gnl;
    if (vtl > 1e-6f) {
        const f3 td = mk(vt.x / vtl, vt.y / vtl, vt.z / vtl);
        const f3 rat = cross3(k.ra, td), rbt = cross3(k.rb, td);
        const f3 Irat = inv_I3(k.Ra, k.iia, rat), Irbt = inv_I3(k.Rb, k.iib, rbt);
        const float kt = k.ima + k.imb + dot3(rat, Irat) + dot3(rbt, Irbt);
        const float s1 = vtl / kt, s2 = k.mu * nl;
        const float dlt = gmin(s1, s2);
        const f3 impt = dlt * td;
        const f3 ta = cross3(k.ra, impt), tb = cross3(k.rb, impt);
        const g3 tdg = G(td), impg = G(impt);
        const g3 gta = inv_I3g(k.Ra, k.iia, gwa2);
        inv_I_rbar(o.Ra, k.Ra, k.iia, G(ta), gwa2, 1.0);
        g3 gimpt = gcross(gta, ra);
        o.ra = o.ra + gcross(impg, gta);
        const g3 gtb = -inv_I3g(k.Rb, k.iib, gwb2);
        inv_I_rbar(o.Rb, k.Rb, k.iib, G(tb), gwb2, -1.0);
        gimpt = gimpt + gcross(gtb, rb);
        o.rb = o.rb + gcross(impg, gtb);
        gimpt = (gimpt + (double)k.ima * gla2) - (double)k.imb * glb2;
        const double gdlt = gdot(gimpt, tdg);
        g3 gtd = (double)dlt * gimpt;
        double gs1 = 0.0, gs2 = 0.0;
        if (s2 < s1) gs2 = gdlt; else gs1 = gdlt;
        o.mu += gs2 * (double)nl;
        gnl_acc += gs2 * (double)k.mu;
        double gvtl = gs1 / (double)kt;
        const double gkt = -gs1 * (double)s1 / (double)kt;
        const g3 grat = (2.0 * gkt) * G(Irat), grbt = (2.0 * gkt) * G(Irbt);
        inv_I_rbar(o.Ra, k.Ra, k.iia, G(rat), G(rat), gkt);
        inv_I_rbar(o.Rb, k.Rb, k.iib, G(rbt), G(rbt), gkt);
        o.ra = o.ra + gcross(tdg, grat);
        gtd = gtd + gcross(grat, ra);
        o.rb = o.rb + gcross(tdg, grbt);
        gtd = gtd + gcross(grbt, rb);
        const double il = 1.0 / (double)vtl;
        const g3 vtg = G(vt);
        g3 gvt = il * gtd;
        gvtl += -gdot(gtd, vtg) * il * il;
        gvt = gvt + (gvtl * il) * vtg;
        gdv = gvt - gdot(gvt, n) * n;
    }
    const g3 gvb2 = gdv, gva2 = -gdv;
    gla1 = gla1 + gva2;
    gwa1 = gwa1 + gcross(ra, gva2);
    o.ra = o.ra + gcross(gva2, G(wa1));
    glb1 = glb1 + gvb2;
    gwb1 = gwb1 + gcross(rb, gvb2);
    o.rb = o.rb + gcross(gvb2, G(wb1));
    g3 gwa = gwa1;
    const g3 gua = -inv_I3g(k.Ra, k.iia, gwa1);
    inv_I_rbar(o.Ra, k.Ra, k.iia, G(ua), gwa1, -1.0);
    const g3 impg = G(imp);
    g3 gimp = gcross(gua, ra);
    o.ra = o.ra + gcross(impg, gua);
    g3 gwb = gwb1;
    const g3 gub = inv_I3g(k.Rb, k.iib, gwb1);
    inv_I_rbar(o.Rb, k.Rb, k.iib, G(ub), gwb1, 1.0);
    gimp = gimp + gcross(gub, rb);
    o.rb = o.rb + gcross(impg, gub);
    g3 gla = gla1, glb = glb1;
    gimp = (gimp - (double)k.ima * gla1) + (double)k.imb * glb1;
    const double gd2 = gdot(gimp, n);
    gnl_acc += gd2;
    double gold = -gd2;
    const double gx = x < 0.0f ? 0.0 : gnl_acc;
    gold += gx;
    const double gdln = gx;
    o.bias += gdln / (double)k.kn;
    const double gvn = -gdln / (double)k.kn;
    o.kn += -gdln * (double)dln / (double)k.kn;
    const g3 gvb = gvn * n, gva = -gvn * n;
    gla = gla + gva;
    gwa = gwa + gcross(ra, gva);
    o.ra = o.ra + gcross(gva, G(wa));
    glb = glb + gvb;
    gwb = gwb + gcross(rb, gvb);
    o.rb = o.rb + gcross(gvb, G(wb));
    o.la = gla;
    o.wa = gwa;
    o.lb = glb;
    o.wb = gwb;
    o.old = gold;
}

// quaternion adjoint from the adjoint A of R = quat_mat(q) (physics.comp:74-81)
SVD void quat_mat_vjp(const float *q, const gm3 &A, double *qb) {
    const double x = q[0], y = q[1], z = q[2], w = q[3];
    qb[0] += A.c0.y * 2.0 * y + A.c0.z * 2.0 * z + A.c1.x * 2.0 * y - A.c1.y * 4.0 * x + A.c1.z * 2.0 * w +
             A.c2.x * 2.0 * z - A.c2.y * 2.0 * w - A.c2.z * 4.0 * x;
    qb[1] += -A.c0.x * 4.0 * y + A.c0.y * 2.0 * x - A.c0.z * 2.0 * w + A.c1.x * 2.0 * x + A.c1.z * 2.0 * z +
             A.c2.x * 2.0 * w + A.c2.y * 2.0 * z - A.c2.z * 4.0 * y;
    qb[2] += -A.c0.x * 4.0 * z + A.c0.y * 2.0 * w + A.c0.z * 2.0 * x - A.c1.x * 2.0 * w - A.c1.y * 4.0 * z +
             A.c1.z * 2.0 * y + A.c2.x * 2.0 * x + A.c2.y * 2.0 * y;
    qb[3] += A.c0.y * 2.0 * z - A.c0.z * 2.0 * y - A.c1.x * 2.0 * z + A.c1.z * 2.0 * x + A.c2.x * 2.0 * y -
             A.c2.y * 2.0 * x;
}

SVD void acc_rot(double *r9, const gm3 &A) {
    r9[0] += A.c0.x; r9[1] += A.c0.y; r9[2] += A.c0.z;
    r9[3] += A.c1.x; r9[4] += A.c1.y; r9[5] += A.c1.z;
    r9[6] += A.c2.x; r9[7] += A.c2.y; r9[8] += A.c2.z;
}
SVD gm3 ld_gm3(const double *r9) {
    gm3 m;
    m.c0 = gk(r9[0], r9[1], r9[2]);
    m.c1 = gk(r9[3], r9[4], r9[5]);
    m.c2 = gk(r9[6], r9[7], r9[8]);
    return m;
}

template <int T>
__global__ void __launch_bounds__(T) k_batch_adjoint(AdjArgs A) {
    __shared__ double s_pb[3 * NB_MAX], s_qb[4 * NB_MAX], s_vb[3 * NB_MAX], s_wb[3 * NB_MAX];
    __shared__ double s_Rb[9 * NB_MAX];
    __shared__ float s_v1[3 * NB_MAX], s_w1[3 * NB_MAX];
    __shared__ uint32_t s_key[DINC_MAX];
    __shared__ uint16_t s_rank[DINC_MAX];
    __shared__ uint16_t s_start[NB_MAX], s_deg[NB_MAX];
    __shared__ uint32_t s_rt[NB_MAX];
    __shared__ double s_lb[DC_MAX];
    __shared__ double s_acc[DC_MAX * ACC];
    __shared__ uint32_t s_q, s_next;
    const int tid = threadIdx.x, lane = tid & 63;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const size_t N = A.n;
    const float *pos0 = A.state, *quat0 = A.state + 3 * N, *vel0 = A.state + 7 * N, *omg0 = A.state + 10 * N;
    double *ap = A.adj, *aq = A.adj + 3 * N, *av = A.adj + 7 * N, *aw = A.adj + 10 * N;
    const float dt = A.dt;
    (void)pos0;

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

        // post-gravity velocities (FP32, as the forward), and the integrate adjoint
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
            g3 gp = ldg(ap, b), gv = ldg(av, b), gw = ldg(aw, b);
            double g[4] = {aq[4 * b], aq[4 * b + 1], aq[4 * b + 2], aq[4 * b + 3]};
            if (dyn) {
                const double hdt = dt;
                gv = gv + hdt * gp;
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
                    const double ux = px / l, uy = py / l, uz = pz / l, uw = pw / l, il = 1.0 / (double)l;
                    const double d = g[0] * ux + g[1] * uy + g[2] * uz + g[3] * uw;
                    g[0] = (g[0] - d * ux) * il;
                    g[1] = (g[1] - d * uy) * il;
                    g[2] = (g[2] - d * uz) * il;
                    g[3] = (g[3] - d * uw) * il;
                }
                const double hh = 0.5 * hdt;
                const double hx = hh * g[0], hy = hh * g[1], hz = hh * g[2], hw = hh * g[3];
                gw = gw + gk(hx * qw - hy * qz + hz * qy - hw * qx, hx * qz + hy * qw - hz * qx - hw * qy,
                             -hx * qy + hy * qx + hz * qw - hw * qz);
                const double n0 = g[0] + hy * oz - hz * oy - hw * ox;
                const double n1 = g[1] - hx * oz + hz * ox - hw * oy;
                const double n2 = g[2] + hx * oy - hy * ox - hw * oz;
                const double n3 = g[3] + hx * ox + hy * oy + hz * oz;
                g[0] = n0; g[1] = n1; g[2] = n2; g[3] = n3;
            }
            stg(s_pb, i, gp);
            stg(s_vb, i, gv);
            stg(s_wb, i, gw);
#pragma unroll
            for (int k = 0; k < 4; k++) s_qb[4 * i + k] = g[k];
#pragma unroll
            for (int k = 0; k < 9; k++) s_Rb[9 * i + k] = 0.0;
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
            s_lb[c] = 0.0;
#pragma unroll
            for (int k = 0; k < (int)ACC; k++) s_acc[c * ACC + k] = 0.0;
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
                        const g3 z = gzero();
                        ItemAdj o;
                        item_vjp(it, t_la, t_wa, t_lb, t_wb, t_old,
                                 it.dyn_a ? ldg(s_vb, it.a) : z, it.dyn_a ? ldg(s_wb, it.a) : z,
                                 it.dyn_b ? ldg(s_vb, it.b) : z, it.dyn_b ? ldg(s_wb, it.b) : z, s_lb[it.c], o);
                        if (it.dyn_a) {
                            stg(s_vb, it.a, o.la);
                            stg(s_wb, it.a, o.wa);
                            acc_rot(s_Rb + 9 * it.a, o.Ra);
                        }
                        if (it.dyn_b) {
                            stg(s_vb, it.b, o.lb);
                            stg(s_wb, it.b, o.wb);
                            acc_rot(s_Rb + 9 * it.b, o.Rb);
                        }
                        s_lb[it.c] = o.old;
                        double *acc = s_acc + it.c * ACC;
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
            const double *acc = s_acc + c * ACC;
            g3 gra = gk(acc[0], acc[1], acc[2]), grb = gk(acc[3], acc[4], acc[5]);
            const double gkn = acc[6], gbias = acc[7], gmu = acc[8];
            const f3 n = mk(P[P_N], P[P_N + 1], P[P_N + 2]);
            const f3 ra = mk(P[P_RA], P[P_RA + 1], P[P_RA + 2]), rb = mk(P[P_RB], P[P_RB + 1], P[P_RB + 2]);
            const f3 w1a = ld3(s_w1, a), w1b = ld3(s_w1, b);
            const f3 va = ld3(s_v1, a) + cross3(w1a, ra), vb = ld3(s_v1, b) + cross3(w1b, rb);
            const float vn = dot3(vb - va, n);
            const float ea = A.rest[gga], eb = A.rest[ggb];
            const float e = gmax(ea, eb);
            const float term = -vn < 0.0f ? 0.0f : -vn;
            const double ge = gbias * (double)term;
            const double gvn = -vn < 0.0f ? 0.0 : -gbias * (double)e;
            // e = max(e_a, e_b) (physics.comp:519): the larger one takes it
            // (a on ties, as GLSL max returns its first argument)
            const double ge_a = ea < eb ? 0.0 : ge, ge_b = ea < eb ? ge : 0.0;
            const g3 ng = G(n), rag = G(ra), rbg = G(rb);
            const g3 gvb = gvn * ng, gva = -gvn * ng;
            const m3 Ra = quat_mat_at(quat0, ga_b), Rb = quat_mat_at(quat0, gb_b);
            const f3 iia = ld3(A.inv_i, ga_b), iib = ld3(A.inv_i, gb_b);
            const f3 ran = cross3(ra, n), rbn = cross3(rb, n);
            const g3 gran = (2.0 * gkn) * G(inv_I3(Ra, iia, ran)), grbn = (2.0 * gkn) * G(inv_I3(Rb, iib, rbn));
            gra = gra + gcross(gva, G(w1a)) + gcross(ng, gran);
            grb = grb + gcross(gvb, G(w1b)) + gcross(ng, grbn);
            // mu = sqrt(f_a f_b) (physics.comp:557): dmu/df_a = f_b / (2 mu)
            const float mu = P[P_MU];
            const double gf_a = mu > 0.0f ? gmu * (double)A.fric[ggb] / (2.0 * (double)mu) : 0.0;
            const double gf_b = mu > 0.0f ? gmu * (double)A.fric[gga] / (2.0 * (double)mu) : 0.0;
            double *ca = A.contrib + (2 * cseg + 2 * c) * CONTRIB;
            double *cb = ca + CONTRIB;
            gm3 Rba = gm3_zero(), Rbb = gm3_zero();
            inv_I_rbar(Rba, Ra, iia, G(ran), G(ran), gkn);
            inv_I_rbar(Rbb, Rb, iib, G(rbn), G(rbn), gkn);
            stg(ca, 0, -gra);
            stg(ca, 1, gva);
            stg(ca, 2, gcross(rag, gva));
            stg(cb, 0, -grb);
            stg(cb, 1, gvb);
            stg(cb, 2, gcross(rbg, gvb));
#pragma unroll
            for (int k = 0; k < 9; k++) {
                ca[9 + k] = 0.0;
                cb[9 + k] = 0.0;
            }
            acc_rot(ca + 9, Rba);
            acc_rot(cb + 9, Rbb);
            ca[18] = gf_a;
            ca[19] = ge_a;
            cb[18] = gf_b;
            cb[19] = ge_b;
        }
        __syncthreads();

        // per body: gather in incidence order, rotation and gravity adjoints
        for (uint32_t i = tid; i < ns; i += T) {
            const uint32_t b = bs + i;
            g3 gp = ldg(s_pb, i), gv = ldg(s_vb, i), gw = ldg(s_wb, i);
            double *R9 = s_Rb + 9 * i;
            const uint32_t deg = s_deg[i];
            if (deg) {
                const uint32_t k0 = s_start[i];
                double gfr = A.afric[b], gre = A.arest[b];
                for (uint32_t p = k0; p < k0 + deg; p++) {
                    const uint32_t slot = s_key[p] & 0xFFFu;
                    const double *cc = A.contrib + (2 * cseg + slot) * CONTRIB;
                    gp = gp + ldg(cc, 0);
                    gv = gv + ldg(cc, 1);
                    gw = gw + ldg(cc, 2);
#pragma unroll
                    for (int k = 0; k < 9; k++) R9[k] += cc[9 + k];
                    gfr += cc[18];
                    gre += cc[19];
                }
                A.afric[b] = gfr;
                A.arest[b] = gre;
            } else if (A.inv_mass[b] == 0.0f && A.inv_i[3 * b] == 0.0f && A.inv_i[3 * b + 1] == 0.0f &&
                       A.inv_i[3 * b + 2] == 0.0f) {
                // a static body has no incidences (it orders nothing): its
                // material shares are gathered from the contacts, in order
                double gfr = A.afric[b], gre = A.arest[b];
                for (uint32_t c = 0; c < nc; c++) {
                    const float *P = A.prep + (cseg + c) * CPREP_WORDS;
                    if (__float_as_uint(P[P_FLAGS]) & F_SKIP) continue;
                    const double *ca = A.contrib + (2 * cseg + 2 * c) * CONTRIB;
                    if (__float_as_uint(P[P_A]) == b) {
                        gfr += ca[18];
                        gre += ca[19];
                    }
                    if (__float_as_uint(P[P_B]) == b) {
                        gfr += ca[CONTRIB + 18];
                        gre += ca[CONTRIB + 19];
                    }
                }
                A.afric[b] = gfr;
                A.arest[b] = gre;
            }
            double qb[4] = {s_qb[4 * i], s_qb[4 * i + 1], s_qb[4 * i + 2], s_qb[4 * i + 3]};
            quat_mat_vjp(quat0 + 4 * (size_t)b, ld_gm3(R9), qb);
            if (A.inv_mass[b] > 0.0f) {
                gv = (double)A.ld * gv;
                gw = (double)A.ad * gw;
            }
            stg(ap, b, gp);
            stg(av, b, gv);
            stg(aw, b, gw);
#pragma unroll
            for (int k = 0; k < 4; k++) aq[4 * b + k] = qb[k];
        }
        __syncthreads();
    }
}

__global__ void k_to_float(const double *in, float *out, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (float)in[i];
}

__global__ void k_to_double(const float *in, double *out, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (double)in[i];
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
    if (e == hipSuccess) e = dmalloc(b->d_adj32, 13 * N);
    if (e == hipSuccess) e = dmalloc(b->d_contrib, 2 * S * cs * sv::CONTRIB);
    if (e == hipSuccess) e = dmalloc(b->d_amat, 2 * N);
    if (e == hipSuccess) e = dmalloc(b->d_amat32, 2 * N);
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
    if (grad_state) {
        e = hipMemcpyAsync(b->d_adj32, grad_state, N13 * sizeof(float), hipMemcpyDefault, st);
        if (e == hipSuccess)
            hipLaunchKernelGGL(sv::k_to_double, dim3(sv::div_up(N13, 256)), dim3(256), 0, st, b->d_adj32, b->d_adj,
                               (uint32_t)N13);
    } else {
        sv::fill_u32(c, reinterpret_cast<uint32_t *>(b->d_adj), 0u, 2 * N13);  // N13 doubles
    }
    if (e != hipSuccess) return sv::batch_dev_fail(e, "sv_batch_diff_backward: setup");
    double *afric = b->d_amat, *arest = b->d_amat + N;
    sv::fill_u32(c, reinterpret_cast<uint32_t *>(b->d_amat), 0u, 4 * N);  // 2N doubles
    sv::launch_hull_ranges(c, b->num_bodies, true);  // exact offsets, as the forward (enqueue_substeps)
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
    if (e == hipSuccess) {
        hipLaunchKernelGGL(sv::k_to_float, dim3(sv::div_up(N13, 256)), dim3(256), 0, ast, b->d_adj, b->d_adj32,
                           (uint32_t)N13);
        e = hipMemcpyAsync(out_state, b->d_adj32, N13 * sizeof(float), hipMemcpyDefault, ast);
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(sv::k_to_float, dim3(sv::div_up(2 * N, 256)), dim3(256), 0, ast, b->d_amat, b->d_amat32,
                           (uint32_t)(2 * N));
        e = hipMemcpyAsync(out_friction, b->d_amat32, N * sizeof(float), hipMemcpyDefault, ast);
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out_restitution, b->d_amat32 + N, N * sizeof(float), hipMemcpyDefault, ast);
    if (e == hipSuccess) e = hipEventRecord(ev_misc, ast);
    if (e == hipSuccess) e = hipStreamWaitEvent(st, ev_misc, 0);
    if (e == hipSuccess) e = copy_state(b, b->d_ckpt + (size_t)frames * N13, false);
    int rc = e == hipSuccess ? sv::batch_finish(b) : sv::batch_dev_fail(e, "sv_batch_diff_backward");
    (void)hipStreamSynchronize(ast);
    return rc;
}

}  // extern "C"
