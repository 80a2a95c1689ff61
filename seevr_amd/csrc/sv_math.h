// sv_math.h - float3 helpers with GLSL 4.50 semantics, shared by the
// gfx950 kernels.  Every expression keeps the association order of the
// reference shader and is compiled without floating-point contraction
// (-ffp-contract=off plus the pragma below), so results are reproducible
// bit for bit on any IEEE-754 binary32 implementation with correctly
// rounded division and square root.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

#define SVD __device__ __forceinline__

// XCC (XCD) the wave runs on: hardware register HW_REG_XCC_ID, bits 0-15.
#define xcc_id() ((uint32_t)__builtin_amdgcn_s_getreg((15 << 11) | 20))

struct f3 {
    float x, y, z;
};

SVD f3 mk(float x, float y, float z) { f3 r; r.x = x; r.y = y; r.z = z; return r; }
// GLSL min/max: min(x,y) = y < x ? y : x ; max(x,y) = x < y ? y : x
SVD float gmin(float x, float y) { return y < x ? y : x; }
SVD float gmax(float x, float y) { return x < y ? y : x; }
SVD f3 operator+(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
SVD f3 operator-(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
SVD f3 operator-(f3 a) { return mk(-a.x, -a.y, -a.z); }
SVD f3 operator*(float s, f3 a) { return mk(s * a.x, s * a.y, s * a.z); }
SVD f3 hmul(f3 a, f3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
SVD float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
SVD f3 cross3(f3 a, f3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
SVD float len3(f3 a) { return __builtin_sqrtf(dot3(a, a)); }
SVD f3 norm3(f3 a) { float inv = 1.0f / __builtin_sqrtf(dot3(a, a)); return inv * a; }
// c ? a : b per component (a select of whole structs can become a select of
// their addresses, which puts them in memory)
SVD f3 sel3(bool c, f3 a, f3 b) { return mk(c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z); }
SVD f3 vmin3(f3 a, f3 b) { return mk(gmin(a.x, b.x), gmin(a.y, b.y), gmin(a.z, b.z)); }
SVD f3 vmax3(f3 a, f3 b) { return mk(gmax(a.x, b.x), gmax(a.y, b.y), gmax(a.z, b.z)); }

SVD f3 ld3(const float *p, uint32_t i) { return mk(p[i * 3], p[i * 3 + 1], p[i * 3 + 2]); }
SVD void st3(float *p, uint32_t i, f3 v) { p[i * 3] = v.x; p[i * 3 + 1] = v.y; p[i * 3 + 2] = v.z; }

// Rotation matrix as three columns (GLSL mat3 is column-major).
struct m3 {
    f3 c0, c1, c2;
};

// Columns of the rotation of quaternion (x, y, z, w) (physics.comp:74-81).
SVD m3 quat_mat(float x, float y, float z, float w) {
    m3 m;
    m.c0 = mk(1.0f - 2.0f * (y * y + z * z), 2.0f * (x * y + w * z), 2.0f * (x * z - w * y));
    m.c1 = mk(2.0f * (x * y - w * z), 1.0f - 2.0f * (x * x + z * z), 2.0f * (y * z + w * x));
    m.c2 = mk(2.0f * (x * z + w * y), 2.0f * (y * z - w * x), 1.0f - 2.0f * (x * x + y * y));
    return m;
}
SVD m3 quat_mat_at(const float *q, uint32_t i) {
    return quat_mat(q[i * 4], q[i * 4 + 1], q[i * 4 + 2], q[i * 4 + 3]);
}
// R * v
SVD f3 mvec(const m3 &m, f3 v) {
    return mk(m.c0.x * v.x + m.c1.x * v.y + m.c2.x * v.z,
              m.c0.y * v.x + m.c1.y * v.y + m.c2.y * v.z,
              m.c0.z * v.x + m.c1.z * v.y + m.c2.z * v.z);
}
// transpose(R) * v
SVD f3 mtvec(const m3 &m, f3 v) { return mk(dot3(m.c0, v), dot3(m.c1, v), dot3(m.c2, v)); }

// Cached rotation: 9 floats per body, columns c0, c1, c2.
SVD m3 ld_rot(const float *rot, uint32_t i) {
    const float *r = rot + (size_t)i * 9;
    m3 m;
    m.c0 = mk(r[0], r[1], r[2]);
    m.c1 = mk(r[3], r[4], r[5]);
    m.c2 = mk(r[6], r[7], r[8]);
    return m;
}
SVD void st_rot(float *rot, uint32_t i, const m3 &m) {
    float *r = rot + (size_t)i * 9;
    r[0] = m.c0.x; r[1] = m.c0.y; r[2] = m.c0.z;
    r[3] = m.c1.x; r[4] = m.c1.y; r[5] = m.c1.z;
    r[6] = m.c2.x; r[7] = m.c2.y; r[8] = m.c2.z;
}

// atan2f of this library.  Reduction to |z| <= tan(pi/8):
//   t = min/max of |x|,|y|;  atan t = pi/4 + atan((t-1)/(t+1)) for t > tan(pi/8);
//   atan z ~= z + z^3 P(z^2), P of degree 4 fitted by gen_atan_poly.py
// (relative error <= 6.3e-8 on the reduced range); then the octant is
// restored with pi/2 - r, pi - r and the sign of y (C99 atan2 zero rules).
#define SV_ATAN_C0 -0x1.55554ap-2f
#define SV_ATAN_C1 0x1.999198p-3f
#define SV_ATAN_C2 -0x1.23b532p-3f
#define SV_ATAN_C3 0x1.b1ecd2p-4f
#define SV_ATAN_C4 -0x1.f1efd6p-5f
#define SV_PI 0x1.921fb6p+1f
#define SV_PIO2 0x1.921fb6p+0f
#define SV_PIO4 0x1.921fb6p-1f
#define SV_TAN_PI8 0x1.a827dap-2f

SVD float sv_atan_reduced(float z) {
    float u = z * z;
    float p = SV_ATAN_C4;
    p = p * u + SV_ATAN_C3;
    p = p * u + SV_ATAN_C2;
    p = p * u + SV_ATAN_C1;
    p = p * u + SV_ATAN_C0;
    return z + (z * u) * p;
}

SVD float sv_atan2f(float y, float x) {
    float ax = __builtin_fabsf(x), ay = __builtin_fabsf(y);
    if (__builtin_isnan(x) || __builtin_isnan(y)) return x + y;
    bool swap = ay > ax;
    float mx = swap ? ay : ax, mn = swap ? ax : ay;
    float r;
    if (mx == 0.0f) {
        r = 0.0f;
    } else if (__builtin_isinf(mx)) {
        bool both = __builtin_isinf(mn);
        r = both ? SV_PIO4 : 0.0f;
        if (both) swap = false;
    } else {
        float t = mn / mx;
        if (t > SV_TAN_PI8) {
            r = SV_PIO4 + sv_atan_reduced((t - 1.0f) / (t + 1.0f));
        } else {
            r = sv_atan_reduced(t);
        }
    }
    if (swap) r = SV_PIO2 - r;
    if (__builtin_signbit(x)) r = SV_PI - r;
    return __builtin_copysignf(r, y);
}

// Broadphase separation (physics.comp:461-466): the boxes are disjoint iff on
// some axis max_a < min_b or max_b < min_a (touching boxes overlap; a NaN
// bound separates nothing, as in the reference).  Six compares whose lane
// masks combine in scalar registers.  Only x, y, z are read.
SVD bool boxes_apart(float4 alo, float4 ahi, float4 blo, float4 bhi) {
    return ahi.x < blo.x || bhi.x < alo.x || ahi.y < blo.y || bhi.y < alo.y || ahi.z < blo.z || bhi.z < alo.z;
}

// The same test as one vector value for the all-pairs loop: the largest of
// the six separations min_b - max_a, min_a - max_b is positive iff the boxes
// are apart, for finite bounds (a difference of two floats is positive iff
// the first is larger, with gradual underflow) and for the empty box
// (+inf lower, -inf upper bounds), whose separations are +inf.  Six
// subtractions and three maxima, so the lane mask comes from one compare.
SVD float boxes_gap(float4 alo, float4 ahi, float4 blo, float4 bhi) {
    // left-nested so the compiler forms two three-input maxima
    const float g = fmaxf(fmaxf(blo.x - ahi.x, alo.x - bhi.x), blo.y - ahi.y);
    return fmaxf(fmaxf(fmaxf(g, alo.y - bhi.y), blo.z - ahi.z), alo.z - bhi.z);
}

// Device-scope release that is complete on return: the release fence writes
// back this XCD's L2, and the explicit wait holds the wave until that
// write-back and every earlier load and store have completed.  hipcc (ROCm
// 7.2) was seen to emit the fence's buffer_wbl2 with no following
// s_waitcnt vmcnt(0) when the stores came from an earlier loop iteration,
// which lets the flag update that follows overtake the data it publishes.
// s_waitcnt 0 = vmcnt(0) expcnt(0) lgkmcnt(0).
SVD void release_agent() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __builtin_amdgcn_s_waitcnt(0);
}
