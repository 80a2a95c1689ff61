// sv_dev.h - device functions shared by the per-stage kernels of the plain
// path (sv_physics.hip, sv_pgs.hip) and the batched engine (sv_batch.hip).
// Both paths evaluate every float expression through these functions, so a
// scene stepped in a batch matches the same scene stepped alone bit for bit.
#pragma once
#include "sv_internal.h"
#include "sv_math.h"

namespace sv {

// Step 0 (physics.comp:418-428) for body i: damped velocities plus gravity,
// dynamic bodies only.
SVD void gravity_body(uint32_t i, float dt, float gx, float gy, float gz, float ld, float ad,
                      const float *__restrict__ inv_mass, float *__restrict__ vel, float *__restrict__ omega) {
    if (inv_mass[i] <= 0.0f) return;
    vel[i * 3 + 0] = vel[i * 3 + 0] * ld + gx * dt;
    vel[i * 3 + 1] = vel[i * 3 + 1] * ld + gy * dt;
    vel[i * 3 + 2] = vel[i * 3 + 2] * ld + gz * dt;
    omega[i * 3 + 0] *= ad;
    omega[i * 3 + 1] *= ad;
    omega[i * 3 + 2] *= ad;
}

// Step 6 (physics.comp:586-604) for body i: x += v dt; q += dt/2 (w,0) q,
// renormalised when |q| > 1e-8.
SVD void integrate_body(uint32_t i, float dt, const float *__restrict__ inv_mass, float *__restrict__ pos,
                        float *__restrict__ quat, const float *__restrict__ vel, const float *__restrict__ omega) {
    if (inv_mass[i] <= 0.0f) return;
    pos[i * 3 + 0] += vel[i * 3 + 0] * dt;
    pos[i * 3 + 1] += vel[i * 3 + 1] * dt;
    pos[i * 3 + 2] += vel[i * 3 + 2] * dt;
    float ox = omega[i * 3], oy = omega[i * 3 + 1], oz = omega[i * 3 + 2];
    float qx = quat[i * 4], qy = quat[i * 4 + 1], qz = quat[i * 4 + 2], qw = quat[i * 4 + 3];
    float dx = ox * qw + oy * qz - oz * qy;
    float dy = -ox * qz + oy * qw + oz * qx;
    float dz = ox * qy - oy * qx + oz * qw;
    float dw = -ox * qx - oy * qy - oz * qz;
    qx += 0.5f * dx * dt;
    qy += 0.5f * dy * dt;
    qz += 0.5f * dz * dt;
    qw += 0.5f * dw * dt;
    float l = __builtin_sqrtf(qx * qx + qy * qy + qz * qz + qw * qw);
    if (l > 1e-8f) {
        qx /= l;
        qy /= l;
        qz /= l;
        qw /= l;
    }
    quat[i * 4] = qx;
    quat[i * 4 + 1] = qy;
    quat[i * 4 + 2] = qz;
    quat[i * 4 + 3] = qw;
}

// Per-body AABB by one wave (the semantics of physics.comp:431-448): min/max
// over p + R*v of the hull verts, seeded with +-1e10.  Lane L folds verts
// L, L+64, L+128, ... in that order (loads issued four at a time), then an
// xor-shuffle tree combines the lanes; every lane ends with the result.
SVD void body_aabb_wave(uint32_t b, int lane, const float *__restrict__ pos, const float *__restrict__ quat,
                        const uint32_t *__restrict__ hrange, const float *__restrict__ hull, m3 &R, f3 &mn, f3 &mx) {
    R = quat_mat_at(quat, b);
    const f3 p = ld3(pos, b);
    const uint32_t vs = hrange[2 * b], vc = hrange[2 * b + 1];
    mn = mk(1e10f, 1e10f, 1e10f);
    mx = mk(-1e10f, -1e10f, -1e10f);
    for (uint32_t v0 = (uint32_t)lane; v0 < vc; v0 += 256) {
        f3 hv[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint32_t v = v0 + 64u * k;
            hv[k] = v < vc ? ld3(hull, vs + v) : mk(0.0f, 0.0f, 0.0f);
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (v0 + 64u * k < vc) {
                const f3 wv = p + mvec(R, hv[k]);
                mn = vmin3(mn, wv);
                mx = vmax3(mx, wv);
            }
        }
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        f3 omn = mk(__shfl_xor(mn.x, off, 64), __shfl_xor(mn.y, off, 64), __shfl_xor(mn.z, off, 64));
        f3 omx = mk(__shfl_xor(mx.x, off, 64), __shfl_xor(mx.y, off, 64), __shfl_xor(mx.z, off, 64));
        mn = vmin3(mn, omn);
        mx = vmax3(mx, omx);
    }
}

// The same box with 16 lanes (one DPP row) per body, four bodies per wave.
// Lane g of the row takes vertices g, g + 16, ...; each round loads 6 per
// lane before using any (indices past the last vertex load the last one
// again, which leaves min and max as they are), and the row reduces by
// right shifts of 1, 2, 4, 8 lanes, so lane 15 of the row ends with the
// box.  min and max select one of their operands, so any order of
// reduction gives the serial loop's box.
constexpr int AABB_ROW = 16, AABB_UNROLL = 10;

template <int CTRL>
SVD float dpp_keep(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), CTRL, 0xF, 0xF, false));
}

SVD void body_aabb_row(uint32_t b, int g, const float *__restrict__ pos, const float *__restrict__ quat,
                       const uint32_t *__restrict__ hrange, const float *__restrict__ hull, m3 &R, f3 &mn,
                       f3 &mx) {
    R = quat_mat_at(quat, b);
    const f3 p = ld3(pos, b);
    const uint32_t vs = hrange[2 * b], vc = hrange[2 * b + 1];
    mn = mk(1e10f, 1e10f, 1e10f);
    mx = mk(-1e10f, -1e10f, -1e10f);
    // loads past the hull's end are not issued (a small hull needs few per
    // lane); the lane's vertices are still taken in index order
    for (uint32_t v0 = (uint32_t)g; v0 < vc; v0 += AABB_ROW * AABB_UNROLL) {
        f3 hv[AABB_UNROLL];
#pragma unroll
        for (int k = 0; k < AABB_UNROLL; k++) {
            const uint32_t v = v0 + (uint32_t)(AABB_ROW * k);
            hv[k] = v < vc ? ld3(hull, vs + v) : mk(0.0f, 0.0f, 0.0f);
        }
#pragma unroll
        for (int k = 0; k < AABB_UNROLL; k++) {
            if (v0 + (uint32_t)(AABB_ROW * k) >= vc) break;
            const f3 wv = p + mvec(R, hv[k]);
            mn = vmin3(mn, wv);
            mx = vmax3(mx, wv);
        }
    }
#define SV_ROW_STEP(CTRL)                                                                          \
    mn = vmin3(mn, mk(dpp_keep<CTRL>(mn.x), dpp_keep<CTRL>(mn.y), dpp_keep<CTRL>(mn.z)));        \
    mx = vmax3(mx, mk(dpp_keep<CTRL>(mx.x), dpp_keep<CTRL>(mx.y), dpp_keep<CTRL>(mx.z)));
    SV_ROW_STEP(0x111)  // row_shr:1
    SV_ROW_STEP(0x112)  // row_shr:2
    SV_ROW_STEP(0x114)  // row_shr:4
    SV_ROW_STEP(0x118)  // row_shr:8
#undef SV_ROW_STEP
}

// ---------------------------------------------------------------------------
// MPR (physics.comp:107-399) for one pair.
struct Body {
    f3 p;
    m3 R;
    uint32_t vs, vc, lk;
};

struct Geo {
    const float *hull;
    const uint32_t *lookup;
    // optional: the lookup's candidate vertices themselves, 3 floats per
    // lookup entry (hull[start + lookup[e]]), so a lookup support needs one
    // memory round trip instead of two; null = read through the lookup
    const float *lxyz = nullptr;
};

SVD Body load_body(uint32_t b, const float *pos, const float *rot, const uint32_t *hrange,
                   const uint32_t *sl_off) {
    Body B;
    B.p = ld3(pos, b);
    B.R = ld_rot(rot, b);
    B.vs = hrange[2 * b];
    B.vc = hrange[2 * b + 1];
    B.lk = sl_off[b];
    return B;
}

// physics.comp:107-162: farthest hull vertex along world direction d.
SVD f3 support(const Body &B, f3 d, const Geo &g) {
    f3 ld = mtvec(B.R, d);
    float best = -1e10f;
    f3 bv = mk(0.0f, 0.0f, 0.0f);
    if (B.lk == 0) {
        // chunks of 8 vertices: all 8 loads issued before the first compare
        // (a box's whole hull in one round trip), then compared in index
        // order with the same strict '>', so the same vertex wins
        constexpr uint32_t CH = 8;
        for (uint32_t i0 = 0; i0 < B.vc; i0 += CH) {
            f3 v[CH];
#pragma unroll
            for (uint32_t k = 0; k < CH; k++) v[k] = ld3(g.hull, B.vs + (i0 + k < B.vc ? i0 + k : i0));
#pragma unroll
            for (uint32_t k = 0; k < CH; k++) {
                if (i0 + k < B.vc) {
                    float dd = dot3(v[k], ld);
                    if (dd > best) {
                        best = dd;
                        bv = v[k];
                    }
                }
            }
        }
        return B.p + mvec(B.R, bv);
    }
    float ang[3];
    ang[0] = sv_atan2f(ld.z, ld.y);
    ang[1] = sv_atan2f(ld.z, ld.x);
    ang[2] = sv_atan2f(ld.y, ld.x);
#pragma unroll
    for (uint32_t ax = 0; ax < 3; ax++) {
        float a = ang[ax];
        if (a < 0.0f) a += 6.2831853f;
        float bin_f = a / 6.2831853f * 64.0f;
        uint32_t b0 = ((uint32_t)bin_f) % 64u;
        uint32_t b1 = (b0 + 1u) % 64u;
        f3 v0, v1;
        if (g.lxyz) {
            v0 = ld3(g.lxyz, B.lk + ax * 64u + b0);
            v1 = ld3(g.lxyz, B.lk + ax * 64u + b1);
        } else {
            v0 = ld3(g.hull, B.vs + g.lookup[B.lk + ax * 64u + b0]);
            v1 = ld3(g.hull, B.vs + g.lookup[B.lk + ax * 64u + b1]);
        }
        float d0 = dot3(v0, ld);
        if (d0 > best) {
            best = d0;
            bv = v0;
        }
        float d1 = dot3(v1, ld);
        if (d1 > best) {
            best = d1;
            bv = v1;
        }
    }
    return B.p + mvec(B.R, bv);
}

constexpr float CCD_EPS = 1e-6f;
constexpr float CCD_TOL = 1e-6f;
constexpr int CCD_ITERS = 50;

// Minkowski portal: Minkowski points and the support points on A and B.
struct Portal {
    f3 v[4], sa[4], sb[4];
};

struct Pair {
    Body A, Bb;
};

SVD void mink(const Pair &P, const Geo &g, f3 d, f3 &v, f3 &sa, f3 &sb) {
    sa = support(P.A, d, g);
    sb = support(P.Bb, -d, g);
    v = sa - sb;
}

SVD f3 portal_dir(f3 v1, f3 v2, f3 v3) { return norm3(cross3(v2 - v1, v3 - v1)); }

SVD bool reach_tol(f3 v1, f3 v2, f3 v3, f3 v4, f3 dir) {
    float dv1 = dot3(v1, dir), dv2 = dot3(v2, dir), dv3 = dot3(v3, dir), dv4 = dot3(v4, dir);
    float d1 = gmin(gmin(dv4 - dv1, dv4 - dv2), dv4 - dv3);
    return d1 < CCD_TOL + CCD_EPS * gmax(1.0f, d1);
}

// physics.comp:200-221
SVD void expand_portal(Portal &s, f3 v4, f3 s4a, f3 s4b) {
    f3 v4v0 = cross3(v4, s.v[0]);
    int slot;
    if (dot3(s.v[1], v4v0) > 0.0f) {
        slot = dot3(s.v[2], v4v0) > 0.0f ? 1 : 3;
    } else {
        slot = dot3(s.v[3], v4v0) > 0.0f ? 2 : 1;
    }
    // one select per element: stores chosen by slot would become a store at
    // a per-lane index, which moves the portal out of registers
    const bool k1 = slot == 1, k2 = slot == 2, k3 = slot == 3;
    s.v[1] = sel3(k1, v4, s.v[1]); s.sa[1] = sel3(k1, s4a, s.sa[1]); s.sb[1] = sel3(k1, s4b, s.sb[1]);
    s.v[2] = sel3(k2, v4, s.v[2]); s.sa[2] = sel3(k2, s4a, s.sa[2]); s.sb[2] = sel3(k2, s4b, s.sb[2]);
    s.v[3] = sel3(k3, v4, s.v[3]); s.sa[3] = sel3(k3, s4a, s.sa[3]); s.sb[3] = sel3(k3, s4b, s.sb[3]);
}

// physics.comp:224-247: barycentric contact point, averaged over A and B.
SVD f3 find_pos(const Portal &s) {
    f3 v0 = s.v[0], v1 = s.v[1], v2 = s.v[2], v3 = s.v[3];
    float b0 = dot3(cross3(v1, v2), v3);
    float b1 = -dot3(cross3(v0, v2), v3);
    float b2 = dot3(cross3(v0, v1), v3);
    float b3 = -dot3(cross3(v0, v1), v2);
    float sum = b0 + b1 + b2 + b3;
    if (__builtin_fabsf(sum) < CCD_EPS) {
        f3 dir = portal_dir(v1, v2, v3);
        b0 = 0.0f;
        b1 = dot3(cross3(v2, v3), dir);
        b2 = dot3(cross3(v3, v1), dir);
        b3 = dot3(cross3(v1, v2), dir);
        sum = b0 + b1 + b2 + b3;
    }
    f3 p1 = ((b0 * s.sa[0] + b1 * s.sa[1]) + b2 * s.sa[2]) + b3 * s.sa[3];
    f3 p2 = ((b0 * s.sb[0] + b1 * s.sb[1]) + b2 * s.sb[2]) + b3 * s.sb[3];
    return (0.5f / sum) * (p1 + p2);
}

// physics.comp:250-320: -1 separated, 0 portal found, 1 touching, 2 collinear
SVD int discover_portal(const Pair &P, const Geo &g, Portal &s) {
    s.sa[0] = P.A.p;
    s.sb[0] = P.Bb.p;
    s.v[0] = s.sa[0] - s.sb[0];
    if (__builtin_fabsf(s.v[0].x) < CCD_EPS && __builtin_fabsf(s.v[0].y) < CCD_EPS &&
        __builtin_fabsf(s.v[0].z) < CCD_EPS)
        s.v[0].x += 10.0f * CCD_EPS;
    f3 dir = norm3(-s.v[0]);
    mink(P, g, dir, s.v[1], s.sa[1], s.sb[1]);
    if (dot3(s.v[1], dir) < CCD_EPS) return -1;
    dir = cross3(s.v[0], s.v[1]);
    if (dot3(dir, dir) < CCD_EPS) {
        if (__builtin_fabsf(s.v[1].x) < CCD_EPS && __builtin_fabsf(s.v[1].y) < CCD_EPS &&
            __builtin_fabsf(s.v[1].z) < CCD_EPS)
            return 1;
        return 2;
    }
    dir = norm3(dir);
    mink(P, g, dir, s.v[2], s.sa[2], s.sb[2]);
    if (dot3(s.v[2], dir) < CCD_EPS) return -1;
    f3 va = s.v[1] - s.v[0];
    f3 vb = s.v[2] - s.v[0];
    dir = norm3(cross3(va, vb));
    if (dot3(dir, s.v[0]) > 0.0f) {
        f3 t = s.v[1]; s.v[1] = s.v[2]; s.v[2] = t;
        t = s.sa[1]; s.sa[1] = s.sa[2]; s.sa[2] = t;
        t = s.sb[1]; s.sb[1] = s.sb[2]; s.sb[2] = t;
        dir = -dir;
    }
    for (int trials = 0; trials < 15;) {
        f3 v3c, s3a, s3b;
        mink(P, g, dir, v3c, s3a, s3b);
        if (dot3(v3c, dir) < CCD_EPS) return -1;
        // replace v2 if the origin is outside (v0, v1, v3c), else v1 if it is
        // outside (v0, v3c, v2); selects per element (see expand_portal)
        const bool r2 = dot3(cross3(s.v[1], v3c), s.v[0]) < -CCD_EPS;
        const bool r1 = !r2 && dot3(cross3(v3c, s.v[2]), s.v[0]) < -CCD_EPS;
        if (r2 || r1) {
            s.v[2] = sel3(r2, v3c, s.v[2]); s.sa[2] = sel3(r2, s3a, s.sa[2]); s.sb[2] = sel3(r2, s3b, s.sb[2]);
            s.v[1] = sel3(r1, v3c, s.v[1]); s.sa[1] = sel3(r1, s3a, s.sa[1]); s.sb[1] = sel3(r1, s3b, s.sb[1]);
            dir = norm3(cross3(s.v[1] - s.v[0], s.v[2] - s.v[0]));
            trials++;
            continue;
        }
        s.v[3] = v3c; s.sa[3] = s3a; s.sb[3] = s3b;
        return 0;
    }
    return -1;
}

// physics.comp:323-343
SVD int refine_portal(const Pair &P, const Geo &g, Portal &s) {
    for (int it = 0; it < CCD_ITERS; it++) {
        f3 dir = portal_dir(s.v[1], s.v[2], s.v[3]);
        if (dot3(s.v[1], dir) > -CCD_EPS) return 0;
        f3 v4, s4a, s4b;
        mink(P, g, dir, v4, s4a, s4b);
        if (!(dot3(v4, dir) > -CCD_EPS) || reach_tol(s.v[1], s.v[2], s.v[3], v4, dir)) return -1;
        expand_portal(s, v4, s4a, s4b);
    }
    return 1;
}

// physics.comp:346-365 (always terminates with a result by the last iteration)
SVD void find_penetration(const Pair &P, const Geo &g, Portal &s, f3 &normal, float &depth) {
    for (int it = 0; it < CCD_ITERS; it++) {
        f3 dir = portal_dir(s.v[1], s.v[2], s.v[3]);
        f3 v4, s4a, s4b;
        mink(P, g, dir, v4, s4a, s4b);
        if (reach_tol(s.v[1], s.v[2], s.v[3], v4, dir) || it >= CCD_ITERS - 1) {
            depth = dot3(dir, s.v[1]);
            normal = dir;
            return;
        }
        expand_portal(s, v4, s4a, s4b);
    }
}

// physics.comp:368-399
SVD bool mpr(const Pair &P, const Geo &g, f3 &normal, float &depth, f3 &cp) {
    Portal s;
    int res = discover_portal(P, g, s);
    if (res == 1) {
        normal = norm3(s.v[0]);
        depth = 0.0f;
        cp = 0.5f * (s.sa[1] + s.sb[1]);
        return true;
    }
    if (res == 2) {
        depth = len3(s.v[1]);
        normal = norm3(s.v[1]);
        cp = 0.5f * (s.sa[1] + s.sb[1]);
        return true;
    }
    if (res != 0) return false;
    if (refine_portal(P, g, s) < 0) return false;
    find_penetration(P, g, s, normal, depth);
    cp = find_pos(s);
    return true;
}


// ---------------------------------------------------------------------------
// Contact sweep pieces (physics.comp:500-583).

// cprep word offsets
enum : int {
    P_N = 0, P_RA = 3, P_RB = 6, P_KN = 9, P_BIAS = 10, P_MU = 11, P_IMA = 12, P_IMB = 13,
    P_A = 14, P_B = 15, P_FLAGS = 16, P_SUCC_A = 17, P_SUCC_B = 18,
    // the batched path keeps the contact's geoms in the words the flow
    // schedule uses for successors
    P_GEOM_A = 17, P_GEOM_B = 18
};
// Successor words (flow form): the contact that next updates the same body,
// in the same iteration or (WRAP) in the next one; HASPRED marks a contact
// with an earlier update of that body in iteration 0.
constexpr uint32_t SUCC_NONE = 0xFFFFFFFFu;
constexpr uint32_t SUCC_WRAP = 0x80000000u;
constexpr uint32_t SUCC_HASPRED = 0x40000000u;
constexpr uint32_t SUCC_ID = 0x3FFFFFFFu;
// F_WARM: the contact starts from a carried-over impulse (warm start),
// applied to both bodies before the sweep.
enum : uint32_t { F_SKIP = 1u, F_DYN_A = 2u, F_DYN_B = 4u, F_WARM = 8u };

SVD f3 apply_inv_I(const m3 &R, f3 inv_i, f3 v) { return mvec(R, hmul(inv_i, mtvec(R, v))); }

// Per-contact constants of the sweep into P[0..CPREP_WORDS).  The bias uses
// the velocities before the sweep (physics.comp:504-521); k_n depends only on
// positions and orientations, which the sweep does not change, so it equals
// the value the serial loop recomputes every iteration (physics.comp:544-546).
// a, b are body indices into pos/vel/omega/inv_mass/inv_i/rot; ga, gb geom
// indices into fric/rest.  Returns the flags word.
SVD uint32_t contact_prep(uint32_t a, uint32_t b, uint32_t ga, uint32_t gb, f3 n, f3 cp, float pen, float dt,
                          float baumgarte, float slop, const float *__restrict__ pos, const float *vel,
                          const float *omega, const float *__restrict__ inv_mass, const float *__restrict__ inv_i,
                          const float *__restrict__ rot, const float *__restrict__ fric,
                          const float *__restrict__ rest, float *__restrict__ P) {
    f3 ra = cp - ld3(pos, a), rb = cp - ld3(pos, b);
    f3 va = ld3(vel, a) + cross3(ld3(omega, a), ra);
    f3 vb = ld3(vel, b) + cross3(ld3(omega, b), rb);
    float vn = dot3(vb - va, n);
    float e = gmax(rest[ga], rest[gb]);
    float bias = baumgarte * gmax(pen - slop, 0.0f) / dt + e * gmax(-vn, 0.0f);
    float ima = inv_mass[a], imb = inv_mass[b];
    f3 iia = ld3(inv_i, a), iib = ld3(inv_i, b);
    m3 Ra = ld_rot(rot, a), Rb = ld_rot(rot, b);
    f3 ran = cross3(ra, n), rbn = cross3(rb, n);
    float kn = ima + imb + dot3(ran, apply_inv_I(Ra, iia, ran)) + dot3(rbn, apply_inv_I(Rb, iib, rbn));
    float mu = __builtin_sqrtf(fric[ga] * fric[gb]);
    bool skip = (ima == 0.0f && imb == 0.0f) || a == b;
    bool dyn_a = !(ima == 0.0f && iia.x == 0.0f && iia.y == 0.0f && iia.z == 0.0f);
    bool dyn_b = !(imb == 0.0f && iib.x == 0.0f && iib.y == 0.0f && iib.z == 0.0f);
    uint32_t flags = (skip ? F_SKIP : 0u) | (dyn_a ? F_DYN_A : 0u) | (dyn_b ? F_DYN_B : 0u);
    P[P_N + 0] = n.x; P[P_N + 1] = n.y; P[P_N + 2] = n.z;
    P[P_RA + 0] = ra.x; P[P_RA + 1] = ra.y; P[P_RA + 2] = ra.z;
    P[P_RB + 0] = rb.x; P[P_RB + 1] = rb.y; P[P_RB + 2] = rb.z;
    P[P_KN] = kn;
    P[P_BIAS] = bias;
    P[P_MU] = mu;
    P[P_IMA] = ima;
    P[P_IMB] = imb;
    P[P_A] = __uint_as_float(a);
    P[P_B] = __uint_as_float(b);
    P[P_FLAGS] = __uint_as_float(flags);
    P[P_SUCC_A] = __uint_as_float(SUCC_NONE);
    P[P_SUCC_B] = __uint_as_float(SUCC_NONE);
    return flags;
}

// One item of the sweep with its constants in registers.
struct Item {
    f3 n, ra, rb;
    float kn, bias, mu, ima, imb;
    m3 Ra, Rb;
    f3 iia, iib;
    uint32_t c, a, b, ta, tb, q;
    bool dyn_a, dyn_b;
};

// Access to the words a sweep item reads and writes (velocities, angular
// velocities, impulses).  PlainMem: ordinary loads and stores (LDS or one
// CU, or after cache maintenance).  CohMem: device-scope relaxed atomics,
// which read and write at the device's point of coherence, so items on
// different XCDs hand these words over without L2 write-back or invalidation.
struct PlainMem {
    static SVD float ld(const float *p) { return *p; }
    static SVD void st(float *p, float v) { *p = v; }
};
struct CohMem {
    static SVD float ld(const float *p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
    static SVD void st(float *p, float v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
};
template <class M>
SVD f3 mld3(const float *p, uint32_t i) { return mk(M::ld(p + i * 3), M::ld(p + i * 3 + 1), M::ld(p + i * 3 + 2)); }
template <class M>
SVD void mst3(float *p, uint32_t i, f3 v) { M::st(p + i * 3, v.x); M::st(p + i * 3 + 1, v.y); M::st(p + i * 3 + 2, v.z); }

// physics.comp:525-580 for one contact on the values of its two bodies and
// its impulse: normal impulse with projection, then Coulomb friction with
// the geometric-mean coefficient.  Returns the new impulse.
SVD float item_update(const Item &it, f3 &la, f3 &wa, f3 &lb, f3 &wb, float old) {
    f3 va = la + cross3(wa, it.ra);
    f3 vb = lb + cross3(wb, it.rb);
    float vn = dot3(vb - va, it.n);
    float dln = (-vn + it.bias) / it.kn;
    float nl = gmax(old + dln, 0.0f);
    dln = nl - old;
    f3 imp = dln * it.n;
    la = la - it.ima * imp;
    lb = lb + it.imb * imp;
    wa = wa - apply_inv_I(it.Ra, it.iia, cross3(it.ra, imp));
    wb = wb + apply_inv_I(it.Rb, it.iib, cross3(it.rb, imp));

    va = la + cross3(wa, it.ra);
    vb = lb + cross3(wb, it.rb);
    f3 dv = vb - va;
    f3 vt = dv - dot3(dv, it.n) * it.n;
    float vtl = len3(vt);
    if (vtl > 1e-6f) {
        f3 td = mk(vt.x / vtl, vt.y / vtl, vt.z / vtl);
        f3 rat = cross3(it.ra, td), rbt = cross3(it.rb, td);
        float kt = it.ima + it.imb + dot3(rat, apply_inv_I(it.Ra, it.iia, rat)) +
                   dot3(rbt, apply_inv_I(it.Rb, it.iib, rbt));
        float dlt = gmin(vtl / kt, it.mu * nl);
        f3 impt = dlt * td;
        la = la + it.ima * impt;
        lb = lb - it.imb * impt;
        wa = wa + apply_inv_I(it.Ra, it.iia, cross3(it.ra, impt));
        wb = wb - apply_inv_I(it.Rb, it.iib, cross3(it.rb, impt));
    }
    return nl;
}

// One sweep item through memory: load, update, store.
template <class M = PlainMem>
SVD void apply_item(const Item &it, float *__restrict__ vel, float *__restrict__ omega,
                    float *__restrict__ lambda) {
    f3 la = mld3<M>(vel, it.a), wa = mld3<M>(omega, it.a);
    f3 lb = mld3<M>(vel, it.b), wb = mld3<M>(omega, it.b);
    float nl = item_update(it, la, wa, lb, wb, M::ld(lambda + it.c));
    M::st(lambda + it.c, nl);
    if (it.dyn_a) { mst3<M>(vel, it.a, la); mst3<M>(omega, it.a, wa); }
    if (it.dyn_b) { mst3<M>(vel, it.b, lb); mst3<M>(omega, it.b, wb); }
}

}  // namespace sv
