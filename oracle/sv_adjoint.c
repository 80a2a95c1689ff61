/*
 * sv_adjoint.c - FP64 CPU adjoint of one physics substep with the contact
 * set held fixed.  TEST INFRASTRUCTURE ONLY: the checker of the library's
 * differentiable step (its gradients are compared with this adjoint and with
 * finite differences of adj_forward).  The product never links or calls it.
 *
 * The substep map differentiated is the part of physics.comp that follows
 * collision detection, with the contacts (point, normal, depth, bodies) of
 * the substep treated as constants:
 *   gravity   physics.comp:418-428   v1 = v0*ld + g*dt, w1 = w0*ad (inv_mass > 0)
 *   sweep     physics.comp:500-583   bias/k_n per contact from the state after
 *                                    gravity, then iters Gauss-Seidel passes
 *                                    over the contacts (normal impulse with
 *                                    projection, then non-accumulated Coulomb
 *                                    friction with mu = sqrt(f_a f_b))
 *   integrate physics.comp:586-604   p += v dt, q = normalize(q + dt/2 (w,0) q)
 * The reverse pass runs the adjoint of every sweep item in reverse order
 * (the exact derivative of the iters passes actually run; at a converged
 * sweep this is the implicit-function adjoint of the Gauss-Seidel fixed
 * point truncated after iters transposed passes), then the adjoints of the
 * per-contact constants, of the rotation matrices (through the quaternion),
 * of gravity and of integration.  Material adjoints are per geom (geom i is
 * body i).
 *
 * adj_set_implicit(1) linearizes every reversed pass at the last pass's
 * inputs instead (the converged sweep, as implicit differentiation at the
 * Gauss-Seidel fixed point would): the reverse sweep is then the fixed-point
 * iteration y <- J^T y + g of (I - J)^T y = g.  For this sweep that is not
 * the derivative of the computation: at the fixed point every item's normal
 * update is zero, so the linearization there loses the derivative of the
 * impulse already applied through the lever arms, rotations and k_n, and the
 * friction update, which is recomputed each pass rather than accumulated,
 * is switched off there (|v_t| <= 1e-6) although it acted in earlier
 * passes.  It exists so that tests can show this against finite
 * differences; the library's adjoint is the unrolled one.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef double R;
typedef struct { R x, y, z; } v3;
typedef struct { v3 c0, c1, c2; } m3;

static inline v3 V(R x, R y, R z) { v3 r = {x, y, z}; return r; }
static inline v3 add(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 scl(R s, v3 a) { return V(s * a.x, s * a.y, s * a.z); }
static inline v3 hmul(v3 a, v3 b) { return V(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline R dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 cross(v3 a, v3 b) { return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static inline R gmin(R x, R y) { return y < x ? y : x; }
static inline R gmax(R x, R y) { return x < y ? y : x; }
static inline v3 ld3(const R *p, uint32_t i) { return V(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
static inline void st3(R *p, uint32_t i, v3 v) { p[3 * i] = v.x; p[3 * i + 1] = v.y; p[3 * i + 2] = v.z; }
static inline void acc3(R *p, uint32_t i, v3 v) { p[3 * i] += v.x; p[3 * i + 1] += v.y; p[3 * i + 2] += v.z; }

static v3 mvec(m3 m, v3 v) {
    return V(m.c0.x * v.x + m.c1.x * v.y + m.c2.x * v.z, m.c0.y * v.x + m.c1.y * v.y + m.c2.y * v.z,
             m.c0.z * v.x + m.c1.z * v.y + m.c2.z * v.z);
}
static v3 mtvec(m3 m, v3 v) { return V(dot(m.c0, v), dot(m.c1, v), dot(m.c2, v)); }
/* physics.comp:88-94: R diag(d) R^T v */
static v3 inv_I(m3 R_, v3 d, v3 v) { return mvec(R_, hmul(d, mtvec(R_, v))); }
/* A += s u w^T (A stored by columns) */
static void add_outer(m3 *A, v3 u, v3 w, R s) {
    A->c0 = add(A->c0, scl(s * w.x, u));
    A->c1 = add(A->c1, scl(s * w.y, u));
    A->c2 = add(A->c2, scl(s * w.z, u));
}
/* adjoint of R in y = inv_I(R, d, x) for output adjoint yb, scaled by s */
static void inv_I_rbar(m3 *Rb, m3 R_, v3 d, v3 x, v3 yb, R s) {
    add_outer(Rb, yb, hmul(d, mtvec(R_, x)), s);
    add_outer(Rb, x, hmul(d, mtvec(R_, yb)), s);
}

/* Precision of the forward recomputation.  By default every forward value
 * is FP64.  adj_set_forward32(1) rounds the result of every forward
 * operation to float.  For float operands that is the float operation
 * itself (a double result of +, -, *, / or sqrt rounded once to float is the
 * correctly rounded float result, since 53 >= 2 * 24 + 2), so the forward
 * values and every branch they decide (impulse clamp, stick or slip, the
 * tangent threshold) are those of the FP32 sweep of physics.comp; the
 * reverse stays FP64.  That is the exact derivative of the FP32 computation,
 * evaluated in double. */
static int g_f32 = 0;
void adj_set_forward32(int on) { g_f32 = on != 0; }
static int g_implicit = 0;
void adj_set_implicit(int on) { g_implicit = on != 0; }
static inline R rf(R x) { return g_f32 ? (R)(float)x : x; }
static inline v3 fadd(v3 a, v3 b) { return V(rf(a.x + b.x), rf(a.y + b.y), rf(a.z + b.z)); }
static inline v3 fsub(v3 a, v3 b) { return V(rf(a.x - b.x), rf(a.y - b.y), rf(a.z - b.z)); }
static inline v3 fscl(R s, v3 a) { return V(rf(s * a.x), rf(s * a.y), rf(s * a.z)); }
static inline v3 fhmul(v3 a, v3 b) { return V(rf(a.x * b.x), rf(a.y * b.y), rf(a.z * b.z)); }
static inline R fdot(v3 a, v3 b) { return rf(rf(rf(a.x * b.x) + rf(a.y * b.y)) + rf(a.z * b.z)); }
static inline v3 fcross(v3 a, v3 b) {
    return V(rf(rf(a.y * b.z) - rf(a.z * b.y)), rf(rf(a.z * b.x) - rf(a.x * b.z)), rf(rf(a.x * b.y) - rf(a.y * b.x)));
}
static inline R fsqrt(R x) { return rf(sqrt(x)); }
/* physics.comp:74-81, each operation rounded like the forward */
static m3 fquat_mat(const R *q) {
    R x = q[0], y = q[1], z = q[2], w = q[3];
    m3 m;
    m.c0 = V(rf(1 - rf(2 * rf(rf(y * y) + rf(z * z)))), rf(2 * rf(rf(x * y) + rf(w * z))),
             rf(2 * rf(rf(x * z) - rf(w * y))));
    m.c1 = V(rf(2 * rf(rf(x * y) - rf(w * z))), rf(1 - rf(2 * rf(rf(x * x) + rf(z * z)))),
             rf(2 * rf(rf(y * z) + rf(w * x))));
    m.c2 = V(rf(2 * rf(rf(x * z) + rf(w * y))), rf(2 * rf(rf(y * z) - rf(w * x))),
             rf(1 - rf(2 * rf(rf(x * x) + rf(y * y)))));
    return m;
}
static v3 fmvec(m3 m, v3 v) {
    return V(rf(rf(rf(m.c0.x * v.x) + rf(m.c1.x * v.y)) + rf(m.c2.x * v.z)),
             rf(rf(rf(m.c0.y * v.x) + rf(m.c1.y * v.y)) + rf(m.c2.y * v.z)),
             rf(rf(rf(m.c0.z * v.x) + rf(m.c1.z * v.y)) + rf(m.c2.z * v.z)));
}
static v3 fmtvec(m3 m, v3 v) { return V(fdot(m.c0, v), fdot(m.c1, v), fdot(m.c2, v)); }
static v3 finv_I(m3 R_, v3 d, v3 v) { return fmvec(R_, fhmul(d, fmtvec(R_, v))); }

typedef struct adj_scene {
    uint32_t n;
    R *pos, *quat, *vel, *omega;
    const R *inv_mass, *inv_i, *fric, *rest;
    uint32_t nc;
    const uint32_t *ca, *cb;
    const R *cp, *cn, *pen;
    uint32_t iters;
    R dt, g[3], baumgarte, slop, ld, ad;
} adj_scene;

/* per-contact constants of the sweep (physics.comp:504-521,544-546) */
typedef struct {
    v3 n, ra, rb;
    R kn, bias, mu, ima, imb;
    int skip, dyn_a, dyn_b;
} cst;

static int body_dyn(const adj_scene *s, uint32_t b) {
    v3 d = ld3(s->inv_i, b);
    return !(s->inv_mass[b] == 0 && d.x == 0 && d.y == 0 && d.z == 0);
}

static cst contact_cst(const adj_scene *s, uint32_t c, const R *vel1, const R *w1) {
    cst k;
    uint32_t a = s->ca[c], b = s->cb[c];
    k.n = ld3(s->cn, c);
    v3 cp = ld3(s->cp, c);
    k.ra = fsub(cp, ld3(s->pos, a));
    k.rb = fsub(cp, ld3(s->pos, b));
    v3 va = fadd(ld3(vel1, a), fcross(ld3(w1, a), k.ra));
    v3 vb = fadd(ld3(vel1, b), fcross(ld3(w1, b), k.rb));
    R vn = fdot(fsub(vb, va), k.n);
    R e = gmax(s->rest[a], s->rest[b]);
    k.bias = rf(rf(rf(rf(s->baumgarte) * gmax(rf(s->pen[c] - rf(s->slop)), 0)) / s->dt) + rf(e * gmax(-vn, 0)));
    k.ima = s->inv_mass[a];
    k.imb = s->inv_mass[b];
    m3 Ra = fquat_mat(s->quat + 4 * a), Rb = fquat_mat(s->quat + 4 * b);
    v3 ran = fcross(k.ra, k.n), rbn = fcross(k.rb, k.n);
    k.kn = rf(rf(rf(k.ima + k.imb) + fdot(ran, finv_I(Ra, ld3(s->inv_i, a), ran))) +
              fdot(rbn, finv_I(Rb, ld3(s->inv_i, b), rbn)));
    k.mu = fsqrt(rf(s->fric[a] * s->fric[b]));
    k.skip = (k.ima == 0 && k.imb == 0) || a == b;
    k.dyn_a = body_dyn(s, a);
    k.dyn_b = body_dyn(s, b);
    return k;
}

size_t adj_tape_doubles(const adj_scene *s) { return (size_t)s->iters * s->nc * 13 + 9 * (size_t)s->n + 1; }

/* One substep with fixed contacts.  tape (adj_tape_doubles) receives the
 * inputs of every sweep item and the velocities after gravity and after the
 * sweep; NULL records nothing. */
void adj_forward(adj_scene *s, R *tape) {
    const uint32_t n = s->n, nc = s->nc;
    /* gravity */
    for (uint32_t i = 0; i < n; i++) {
        if (s->inv_mass[i] <= 0) continue;
        for (int k = 0; k < 3; k++) {
            s->vel[3 * i + k] = rf(rf(s->vel[3 * i + k] * rf(s->ld)) + rf(s->g[k] * s->dt));
            s->omega[3 * i + k] = rf(s->omega[3 * i + k] * rf(s->ad));
        }
    }
    if (tape) {
        R *t1 = tape + (size_t)s->iters * nc * 13;
        memcpy(t1, s->vel, 3 * n * sizeof(R));
        memcpy(t1 + 3 * n, s->omega, 3 * n * sizeof(R));
    }
    cst *K = (cst *)malloc(sizeof(cst) * (nc ? nc : 1));
    R *lam = (R *)calloc(nc ? nc : 1, sizeof(R));
    for (uint32_t c = 0; c < nc; c++) K[c] = contact_cst(s, c, s->vel, s->omega);
    for (uint32_t t = 0; t < s->iters; t++) {
        for (uint32_t c = 0; c < nc; c++) {
            const cst k = K[c];
            if (k.skip) continue;
            uint32_t a = s->ca[c], b = s->cb[c];
            v3 la = ld3(s->vel, a), wa = ld3(s->omega, a), lb = ld3(s->vel, b), wb = ld3(s->omega, b);
            if (tape) {
                R *tp = tape + ((size_t)t * nc + c) * 13;
                st3(tp, 0, la); st3(tp, 1, wa); st3(tp, 2, lb); st3(tp, 3, wb);
                tp[12] = lam[c];
            }
            m3 Ra = fquat_mat(s->quat + 4 * a), Rb = fquat_mat(s->quat + 4 * b);
            v3 iia = ld3(s->inv_i, a), iib = ld3(s->inv_i, b);
            v3 va = fadd(la, fcross(wa, k.ra)), vb = fadd(lb, fcross(wb, k.rb));
            R vn = fdot(fsub(vb, va), k.n);
            R dln = rf(rf(-vn + k.bias) / k.kn);
            R old = lam[c];
            R nl = gmax(rf(old + dln), 0);
            dln = rf(nl - old);
            lam[c] = nl;
            v3 imp = fscl(dln, k.n);
            la = fsub(la, fscl(k.ima, imp));
            lb = fadd(lb, fscl(k.imb, imp));
            wa = fsub(wa, finv_I(Ra, iia, fcross(k.ra, imp)));
            wb = fadd(wb, finv_I(Rb, iib, fcross(k.rb, imp)));
            va = fadd(la, fcross(wa, k.ra));
            vb = fadd(lb, fcross(wb, k.rb));
            v3 dv = fsub(vb, va);
            v3 vt = fsub(dv, fscl(fdot(dv, k.n), k.n));
            R vtl = fsqrt(fdot(vt, vt));
            if (vtl > rf(1e-6)) {
                v3 td = V(rf(vt.x / vtl), rf(vt.y / vtl), rf(vt.z / vtl));
                v3 rat = fcross(k.ra, td), rbt = fcross(k.rb, td);
                R kt = rf(rf(rf(k.ima + k.imb) + fdot(rat, finv_I(Ra, iia, rat))) + fdot(rbt, finv_I(Rb, iib, rbt)));
                R dlt = gmin(rf(vtl / kt), rf(k.mu * nl));
                v3 impt = fscl(dlt, td);
                la = fadd(la, fscl(k.ima, impt));
                lb = fsub(lb, fscl(k.imb, impt));
                wa = fadd(wa, finv_I(Ra, iia, fcross(k.ra, impt)));
                wb = fsub(wb, finv_I(Rb, iib, fcross(k.rb, impt)));
            }
            if (k.dyn_a) { st3(s->vel, a, la); st3(s->omega, a, wa); }
            if (k.dyn_b) { st3(s->vel, b, lb); st3(s->omega, b, wb); }
        }
    }
    if (tape) memcpy(tape + (size_t)s->iters * nc * 13 + 6 * n, s->omega, 3 * n * sizeof(R));
    free(K);
    free(lam);
    /* integrate */
    for (uint32_t i = 0; i < n; i++) {
        if (s->inv_mass[i] <= 0) continue;
        for (int k = 0; k < 3; k++) s->pos[3 * i + k] = rf(s->pos[3 * i + k] + rf(s->vel[3 * i + k] * s->dt));
        R ox = s->omega[3 * i], oy = s->omega[3 * i + 1], oz = s->omega[3 * i + 2];
        R *q = s->quat + 4 * i;
        R qx = q[0], qy = q[1], qz = q[2], qw = q[3];
        R dx = rf(rf(rf(ox * qw) + rf(oy * qz)) - rf(oz * qy));
        R dy = rf(rf(rf(-ox * qz) + rf(oy * qw)) + rf(oz * qx));
        R dz = rf(rf(rf(ox * qy) - rf(oy * qx)) + rf(oz * qw));
        R dw = rf(rf(rf(-ox * qx) - rf(oy * qy)) - rf(oz * qz));
        qx = rf(qx + rf(rf(0.5 * dx) * s->dt)); qy = rf(qy + rf(rf(0.5 * dy) * s->dt));
        qz = rf(qz + rf(rf(0.5 * dz) * s->dt)); qw = rf(qw + rf(rf(0.5 * dw) * s->dt));
        R l = fsqrt(rf(rf(rf(rf(qx * qx) + rf(qy * qy)) + rf(qz * qz)) + rf(qw * qw)));
        if (l > rf(1e-8)) { qx = rf(qx / l); qy = rf(qy / l); qz = rf(qz / l); qw = rf(qw / l); }
        q[0] = qx; q[1] = qy; q[2] = qz; q[3] = qw;
    }
}

/* adjoints of one sweep item's inputs and constants */
typedef struct {
    v3 la, wa, lb, wb;
    R old;
    v3 ra, rb;
    R kn, bias, mu;
    m3 Ra, Rb;
} item_adj;

static void item_vjp(const cst *k, m3 Ra, m3 Rb, v3 iia, v3 iib, v3 la, v3 wa, v3 lb, v3 wb, R old, v3 gla2,
                     v3 gwa2, v3 glb2, v3 gwb2, R gnl, item_adj *o) {
    memset(o, 0, sizeof(*o));
    /* forward recomputation */
    v3 va = fadd(la, fcross(wa, k->ra)), vb = fadd(lb, fcross(wb, k->rb));
    R vn = fdot(fsub(vb, va), k->n);
    R dln = rf(rf(-vn + k->bias) / k->kn);
    R x = rf(old + dln);
    R nl = gmax(x, 0);
    R d2 = rf(nl - old);
    v3 imp = fscl(d2, k->n);
    v3 la1 = fsub(la, fscl(k->ima, imp)), lb1 = fadd(lb, fscl(k->imb, imp));
    v3 ua = fcross(k->ra, imp), ub = fcross(k->rb, imp);
    v3 wa1 = fsub(wa, finv_I(Ra, iia, ua)), wb1 = fadd(wb, finv_I(Rb, iib, ub));
    v3 va2 = fadd(la1, fcross(wa1, k->ra)), vb2 = fadd(lb1, fcross(wb1, k->rb));
    v3 dv = fsub(vb2, va2);
    v3 vt = fsub(dv, fscl(fdot(dv, k->n), k->n));
    R vtl = fsqrt(fdot(vt, vt));
    /* reverse */
    v3 gla1 = gla2, gwa1 = gwa2, glb1 = glb2, gwb1 = gwb2, gdv = V(0, 0, 0);
    R gnl_acc = gnl;
    if (vtl > rf(1e-6)) {
        v3 td = V(rf(vt.x / vtl), rf(vt.y / vtl), rf(vt.z / vtl));
        v3 rat = fcross(k->ra, td), rbt = fcross(k->rb, td);
        v3 Irat = finv_I(Ra, iia, rat), Irbt = finv_I(Rb, iib, rbt);
        R kt = rf(rf(rf(k->ima + k->imb) + fdot(rat, Irat)) + fdot(rbt, Irbt));
        R s1 = rf(vtl / kt), s2 = rf(k->mu * nl);
        R dlt = gmin(s1, s2);
        v3 impt = fscl(dlt, td);
        v3 ta = fcross(k->ra, impt), tb = fcross(k->rb, impt);
        v3 gta = inv_I(Ra, iia, gwa2);
        inv_I_rbar(&o->Ra, Ra, iia, ta, gwa2, 1);
        v3 gimpt = cross(gta, k->ra);
        o->ra = add(o->ra, cross(impt, gta));
        v3 gtb = scl(-1, inv_I(Rb, iib, gwb2));
        inv_I_rbar(&o->Rb, Rb, iib, tb, gwb2, -1);
        gimpt = add(gimpt, cross(gtb, k->rb));
        o->rb = add(o->rb, cross(impt, gtb));
        gimpt = sub(add(gimpt, scl(k->ima, gla2)), scl(k->imb, glb2));
        R gdlt = dot(gimpt, td);
        v3 gtd = scl(dlt, gimpt);
        R gs1 = 0, gs2 = 0;
        if (s2 < s1) gs2 = gdlt; else gs1 = gdlt;
        o->mu += gs2 * nl;
        gnl_acc += gs2 * k->mu;
        R gvtl = gs1 / kt;
        R gkt = -gs1 * s1 / kt;
        v3 grat = scl(2 * gkt, Irat), grbt = scl(2 * gkt, Irbt);
        inv_I_rbar(&o->Ra, Ra, iia, rat, rat, gkt);
        inv_I_rbar(&o->Rb, Rb, iib, rbt, rbt, gkt);
        o->ra = add(o->ra, cross(td, grat));
        gtd = add(gtd, cross(grat, k->ra));
        o->rb = add(o->rb, cross(td, grbt));
        gtd = add(gtd, cross(grbt, k->rb));
        v3 gvt = scl(1 / vtl, gtd);
        gvtl += -dot(gtd, vt) / (vtl * vtl);
        gvt = add(gvt, scl(gvtl / vtl, vt));
        gdv = sub(gvt, scl(dot(gvt, k->n), k->n));
    }
    v3 gvb2 = gdv, gva2 = scl(-1, gdv);
    gla1 = add(gla1, gva2);
    gwa1 = add(gwa1, cross(k->ra, gva2));
    o->ra = add(o->ra, cross(gva2, wa1));
    glb1 = add(glb1, gvb2);
    gwb1 = add(gwb1, cross(k->rb, gvb2));
    o->rb = add(o->rb, cross(gvb2, wb1));
    v3 gwa = gwa1;
    v3 gua = scl(-1, inv_I(Ra, iia, gwa1));
    inv_I_rbar(&o->Ra, Ra, iia, ua, gwa1, -1);
    v3 gimp = cross(gua, k->ra);
    o->ra = add(o->ra, cross(imp, gua));
    v3 gwb = gwb1;
    v3 gub = inv_I(Rb, iib, gwb1);
    inv_I_rbar(&o->Rb, Rb, iib, ub, gwb1, 1);
    gimp = add(gimp, cross(gub, k->rb));
    o->rb = add(o->rb, cross(imp, gub));
    v3 gla = gla1, glb = glb1;
    gimp = add(sub(gimp, scl(k->ima, gla1)), scl(k->imb, glb1));
    R gd2 = dot(gimp, k->n);
    gnl_acc += gd2;
    R gold = -gd2;
    R gx = x < 0 ? 0 : gnl_acc;
    gold += gx;
    R gdln = gx;
    o->bias += gdln / k->kn;
    R gvn = -gdln / k->kn;
    o->kn += -gdln * dln / k->kn;
    v3 gvb = scl(gvn, k->n), gva = scl(-gvn, k->n);
    gla = add(gla, gva);
    gwa = add(gwa, cross(k->ra, gva));
    o->ra = add(o->ra, cross(gva, wa));
    glb = add(glb, gvb);
    gwb = add(gwb, cross(k->rb, gvb));
    o->rb = add(o->rb, cross(gvb, wb));
    o->la = gla; o->wa = gwa; o->lb = glb; o->wb = gwb; o->old = gold;
}

/* q-bar from R-bar for R = quat_mat(q) */
static void quat_mat_vjp(const R *q, m3 A, R *qb) {
    R x = q[0], y = q[1], z = q[2], w = q[3];
    qb[0] += A.c0.y * 2 * y + A.c0.z * 2 * z + A.c1.x * 2 * y - A.c1.y * 4 * x + A.c1.z * 2 * w + A.c2.x * 2 * z -
             A.c2.y * 2 * w - A.c2.z * 4 * x;
    qb[1] += -A.c0.x * 4 * y + A.c0.y * 2 * x - A.c0.z * 2 * w + A.c1.x * 2 * x + A.c1.z * 2 * z + A.c2.x * 2 * w +
             A.c2.y * 2 * z - A.c2.z * 4 * y;
    qb[2] += -A.c0.x * 4 * z + A.c0.y * 2 * w + A.c0.z * 2 * x - A.c1.x * 2 * w - A.c1.y * 4 * z + A.c1.z * 2 * y +
             A.c2.x * 2 * x + A.c2.y * 2 * y;
    qb[3] += A.c0.y * 2 * z - A.c0.z * 2 * y - A.c1.x * 2 * z + A.c1.z * 2 * x + A.c2.x * 2 * y - A.c2.y * 2 * x;
}

/* Reverse of adj_forward.  s holds the state at the substep start; tape is
 * the forward's record.  ap/aq/av/aw (3n, 4n, 3n, 3n) hold the adjoint of the
 * state after the substep on entry and of the state before it on return;
 * af/ae (n) accumulate the friction and restitution adjoints per geom. */
void adj_backward(const adj_scene *s, const R *tape, R *ap, R *aq, R *av, R *aw, R *af, R *ae) {
    const uint32_t n = s->n, nc = s->nc;
    const R *v1 = tape + (size_t)s->iters * nc * 13, *w1 = v1 + 3 * n, *w2 = v1 + 6 * n;
    const R dt = s->dt;
    m3 *Rbar = (m3 *)calloc(n ? n : 1, sizeof(m3));
    /* integrate */
    for (uint32_t i = 0; i < n; i++) {
        if (s->inv_mass[i] <= 0) continue;
        v3 gp = ld3(ap, i);
        acc3(av, i, scl(dt, gp)); /* p3 = p0 + v2 dt */
        const R *q = s->quat + 4 * i;
        R ox = w2[3 * i], oy = w2[3 * i + 1], oz = w2[3 * i + 2];
        R qx = q[0], qy = q[1], qz = q[2], qw = q[3];
        R dx = rf(rf(rf(ox * qw) + rf(oy * qz)) - rf(oz * qy));
        R dy = rf(rf(rf(-ox * qz) + rf(oy * qw)) + rf(oz * qx));
        R dz = rf(rf(rf(ox * qy) - rf(oy * qx)) + rf(oz * qw));
        R dw = rf(rf(rf(-ox * qx) - rf(oy * qy)) - rf(oz * qz));
        R px = rf(qx + rf(rf(0.5 * dx) * dt)), py = rf(qy + rf(rf(0.5 * dy) * dt)), pz = rf(qz + rf(rf(0.5 * dz) * dt)),
          pw = rf(qw + rf(rf(0.5 * dw) * dt));
        R l = fsqrt(rf(rf(rf(rf(px * px) + rf(py * py)) + rf(pz * pz)) + rf(pw * pw)));
        R g[4] = {aq[4 * i], aq[4 * i + 1], aq[4 * i + 2], aq[4 * i + 3]};
        if (l > rf(1e-8)) {
            R ux = rf(px / l), uy = rf(py / l), uz = rf(pz / l), uw = rf(pw / l);
            R d = g[0] * ux + g[1] * uy + g[2] * uz + g[3] * uw;
            g[0] = (g[0] - d * ux) / l; g[1] = (g[1] - d * uy) / l;
            g[2] = (g[2] - d * uz) / l; g[3] = (g[3] - d * uw) / l;
        }
        R gdx = 0.5 * dt * g[0], gdy = 0.5 * dt * g[1], gdz = 0.5 * dt * g[2], gdw = 0.5 * dt * g[3];
        aw[3 * i] += gdx * qw - gdy * qz + gdz * qy - gdw * qx;
        aw[3 * i + 1] += gdx * qz + gdy * qw - gdz * qx - gdw * qy;
        aw[3 * i + 2] += -gdx * qy + gdy * qx + gdz * qw - gdw * qz;
        aq[4 * i] = g[0] + gdy * oz - gdz * oy - gdw * ox;
        aq[4 * i + 1] = g[1] - gdx * oz + gdz * ox - gdw * oy;
        aq[4 * i + 2] = g[2] + gdx * oy - gdy * ox - gdw * oz;
        aq[4 * i + 3] = g[3] + gdx * ox + gdy * oy + gdz * oz;
    }
    /* sweep, reversed; av/aw now hold the adjoint of the post-sweep velocities */
    adj_scene s1 = *s;
    cst *K = (cst *)malloc(sizeof(cst) * (nc ? nc : 1));
    R *glam = (R *)calloc(nc ? nc : 1, sizeof(R));
    R *acc = (R *)calloc(9 * (size_t)(nc ? nc : 1), sizeof(R)); /* ra 3, rb 3, kn, bias, mu */
    for (uint32_t c = 0; c < nc; c++) K[c] = contact_cst(&s1, c, v1, w1);
    for (int64_t t = (int64_t)s->iters - 1; t >= 0; t--) {
        for (int64_t cc = (int64_t)nc - 1; cc >= 0; cc--) {
            uint32_t c = (uint32_t)cc;
            const cst *k = &K[c];
            if (k->skip) continue;
            uint32_t a = s->ca[c], b = s->cb[c];
            const R *tp = tape + ((size_t)(g_implicit ? (int64_t)s->iters - 1 : t) * nc + c) * 13;
            m3 Ra = fquat_mat(s->quat + 4 * a), Rb = fquat_mat(s->quat + 4 * b);
            v3 zero = V(0, 0, 0);
            item_adj o;
            item_vjp(k, Ra, Rb, ld3(s->inv_i, a), ld3(s->inv_i, b), ld3(tp, 0), ld3(tp, 1), ld3(tp, 2), ld3(tp, 3),
                     tp[12], k->dyn_a ? ld3(av, a) : zero, k->dyn_a ? ld3(aw, a) : zero,
                     k->dyn_b ? ld3(av, b) : zero, k->dyn_b ? ld3(aw, b) : zero, glam[c], &o);
            if (k->dyn_a) {
                st3(av, a, o.la); st3(aw, a, o.wa);
                Rbar[a].c0 = add(Rbar[a].c0, o.Ra.c0); Rbar[a].c1 = add(Rbar[a].c1, o.Ra.c1);
                Rbar[a].c2 = add(Rbar[a].c2, o.Ra.c2);
            }
            if (k->dyn_b) {
                st3(av, b, o.lb); st3(aw, b, o.wb);
                Rbar[b].c0 = add(Rbar[b].c0, o.Rb.c0); Rbar[b].c1 = add(Rbar[b].c1, o.Rb.c1);
                Rbar[b].c2 = add(Rbar[b].c2, o.Rb.c2);
            }
            glam[c] = o.old;
            R *A = acc + 9 * c;
            A[0] += o.ra.x; A[1] += o.ra.y; A[2] += o.ra.z;
            A[3] += o.rb.x; A[4] += o.rb.y; A[5] += o.rb.z;
            A[6] += o.kn; A[7] += o.bias; A[8] += o.mu;
        }
    }
    /* per-contact constants: bias, k_n, mu, ra, rb */
    for (uint32_t c = 0; c < nc; c++) {
        const cst *k = &K[c];
        if (k->skip) continue;
        uint32_t a = s->ca[c], b = s->cb[c];
        const R *A = acc + 9 * c;
        v3 gra = V(A[0], A[1], A[2]), grb = V(A[3], A[4], A[5]);
        R gkn = A[6], gbias = A[7], gmu = A[8];
        v3 va = fadd(ld3(v1, a), fcross(ld3(w1, a), k->ra));
        v3 vb = fadd(ld3(v1, b), fcross(ld3(w1, b), k->rb));
        R vn = fdot(fsub(vb, va), k->n);
        R e = gmax(s->rest[a], s->rest[b]);
        R term = -vn < 0 ? 0 : -vn;
        R ge = gbias * term;
        R gvn = -vn < 0 ? 0 : -gbias * e;
        if (s->rest[a] < s->rest[b]) ae[b] += ge; else ae[a] += ge;
        v3 gvb = scl(gvn, k->n), gva = scl(-gvn, k->n);
        m3 Ra = fquat_mat(s->quat + 4 * a), Rb = fquat_mat(s->quat + 4 * b);
        v3 ran = fcross(k->ra, k->n), rbn = fcross(k->rb, k->n);
        v3 gran = scl(2 * gkn, finv_I(Ra, ld3(s->inv_i, a), ran));
        v3 grbn = scl(2 * gkn, finv_I(Rb, ld3(s->inv_i, b), rbn));
        gra = add(gra, cross(gva, ld3(w1, a)));
        grb = add(grb, cross(gvb, ld3(w1, b)));
        gra = add(gra, cross(k->n, gran));
        grb = add(grb, cross(k->n, grbn));
        if (k->mu > 0) {
            af[a] += gmu * s->fric[b] / (2 * k->mu);
            af[b] += gmu * s->fric[a] / (2 * k->mu);
        }
        if (k->dyn_a) {
            acc3(av, a, gva);
            acc3(aw, a, cross(k->ra, gva));
            acc3(ap, a, scl(-1, gra));
            m3 t = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
            inv_I_rbar(&t, Ra, ld3(s->inv_i, a), ran, ran, gkn);
            Rbar[a].c0 = add(Rbar[a].c0, t.c0); Rbar[a].c1 = add(Rbar[a].c1, t.c1); Rbar[a].c2 = add(Rbar[a].c2, t.c2);
        }
        if (k->dyn_b) {
            acc3(av, b, gvb);
            acc3(aw, b, cross(k->rb, gvb));
            acc3(ap, b, scl(-1, grb));
            m3 t = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
            inv_I_rbar(&t, Rb, ld3(s->inv_i, b), rbn, rbn, gkn);
            Rbar[b].c0 = add(Rbar[b].c0, t.c0); Rbar[b].c1 = add(Rbar[b].c1, t.c1); Rbar[b].c2 = add(Rbar[b].c2, t.c2);
        }
    }
    for (uint32_t i = 0; i < n; i++) quat_mat_vjp(s->quat + 4 * i, Rbar[i], aq + 4 * i);
    /* gravity */
    for (uint32_t i = 0; i < n; i++) {
        if (s->inv_mass[i] <= 0) continue;
        for (int k = 0; k < 3; k++) {
            av[3 * i + k] *= s->ld;
            aw[3 * i + k] *= s->ad;
        }
    }
    free(K);
    free(glam);
    free(acc);
    free(Rbar);
}
