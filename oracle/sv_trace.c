/*
 * sv_trace.c - CPU restatement of the reference path tracer's closest-hit
 * query (raytrace.comp:56-133).  TEST INFRASTRUCTURE ONLY: it is the checker
 * for the BVH layout contract the tracer relies on (node_count > 0 marks a
 * leaf, node_left of a leaf is its first sorted primitive, prim_idx maps a
 * sorted position to a triangle, node_left/node_right of an internal node
 * are its children, node 0 is the root).  The product never links it.
 *
 * Float arithmetic, no contraction, GLSL min/max (y < x ? y : x) and the
 * tracer's fixed 32-entry stack with its "push only while sp < 31" guard
 * (raytrace.comp:99,127-128), which silently drops subtrees of deep trees;
 * each ray reports whether that happened.
 */
#include <math.h>
#include <stdint.h>

typedef struct { float x, y, z; } v3;

static inline v3 V(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 sub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 mul(v3 a, v3 b) { return V(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 cross(v3 a, v3 b) {
    return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline float gmin(float x, float y) { return y < x ? y : x; }
static inline float gmax(float x, float y) { return x < y ? y : x; }
static inline v3 vmin(v3 a, v3 b) { return V(gmin(a.x, b.x), gmin(a.y, b.y), gmin(a.z, b.z)); }
static inline v3 vmax(v3 a, v3 b) { return V(gmax(a.x, b.x), gmax(a.y, b.y), gmax(a.z, b.z)); }
static inline v3 ld3(const float *p, uint32_t i) { return V(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }

/* raytrace.comp:56-70 (Moller-Trumbore; hits closer than 0.001 rejected) */
float ort_ray_tri(const float ro[3], const float rd[3], const float *v0p, const float *v1p,
                  const float *v2p) {
    v3 o = V(ro[0], ro[1], ro[2]), d = V(rd[0], rd[1], rd[2]);
    v3 v0 = V(v0p[0], v0p[1], v0p[2]), v1 = V(v1p[0], v1p[1], v1p[2]), v2 = V(v2p[0], v2p[1], v2p[2]);
    v3 e1 = sub(v1, v0), e2 = sub(v2, v0);
    v3 h = cross(d, e2);
    float a = dot(e1, h);
    if (fabsf(a) < 1e-8f) return -1.0f;
    float f = 1.0f / a;
    v3 s = sub(o, v0);
    float u = f * dot(s, h);
    if (u < 0.0f || u > 1.0f) return -1.0f;
    v3 q = cross(s, e1);
    float v = f * dot(d, q);
    if (v < 0.0f || u + v > 1.0f) return -1.0f;
    float t = f * dot(e2, q);
    return t > 0.001f ? t : -1.0f;
}

/* raytrace.comp:74-82 */
static int ray_aabb(v3 ro, v3 inv_d, v3 bmin, v3 bmax, float closest) {
    v3 t1 = mul(sub(bmin, ro), inv_d);
    v3 t2 = mul(sub(bmax, ro), inv_d);
    v3 tmin = vmin(t1, t2), tmax = vmax(t1, t2);
    float enter = gmax(gmax(tmin.x, tmin.y), tmin.z);
    float exit_ = gmin(gmin(tmax.x, tmax.y), tmax.z);
    return exit_ >= enter && enter < closest && exit_ > 0.0f;
}

static float tri_t(const float *verts, const uint32_t *idx, uint32_t ti, const float ro[3], const float rd[3]) {
    uint32_t i0 = idx[3 * ti], i1 = idx[3 * ti + 1], i2 = idx[3 * ti + 2];
    return ort_ray_tri(ro, rd, verts + 3 * i0, verts + 3 * i1, verts + 3 * i2);
}

/* raytrace.comp:86-133: closest hit through the BVH.  out_t = 1e10 and
 * out_tri = 0xFFFFFFFF for a miss; out_dropped = 1 when the stack guard
 * dropped a push.  visited counts node tests (a cost figure). */
void ort_trace(const float *verts, const uint32_t *idx, const float *node_min, const float *node_max,
               const uint32_t *node_left, const uint32_t *node_right, const uint32_t *node_count,
               const uint32_t *prim_idx, uint32_t n_rays, const float *ro, const float *rd, float *out_t,
               uint32_t *out_tri, uint32_t *out_dropped, uint64_t *visited) {
    uint64_t vis = 0;
    for (uint32_t r = 0; r < n_rays; r++) {
        const float *o = ro + 3 * r, *d = rd + 3 * r;
        v3 ov = V(o[0], o[1], o[2]);
        v3 inv_d = V(1.0f / d[0], 1.0f / d[1], 1.0f / d[2]);
        float closest = 1e10f;
        uint32_t hit = 0xFFFFFFFFu, dropped = 0;
        int stack[32];
        stack[0] = 0;
        int sp = 1;
        while (sp > 0) {
            sp--;
            int node = stack[sp];
            vis++;
            if (!ray_aabb(ov, inv_d, ld3(node_min, (uint32_t)node), ld3(node_max, (uint32_t)node), closest)) continue;
            uint32_t cnt = node_count[node], first = node_left[node];
            if (cnt > 0) {
                for (uint32_t i = 0; i < cnt; i++) {
                    uint32_t ti = prim_idx[first + i];
                    float t = tri_t(verts, idx, ti, o, d);
                    if (t > 0.0f && t < closest) {
                        closest = t;
                        hit = ti;
                    }
                }
            } else {
                uint32_t right = node_right[node];
                if (sp < 31) stack[sp++] = (int)first; else dropped = 1;
                if (sp < 31) stack[sp++] = (int)right; else dropped = 1;
            }
        }
        out_t[r] = closest;
        out_tri[r] = hit;
        out_dropped[r] = dropped;
    }
    if (visited) *visited = vis;
}

/* Closest hit over every triangle in index order (first index wins ties):
 * the answer the traversal must reproduce when it drops nothing. */
void ort_brute(const float *verts, const uint32_t *idx, uint32_t n_tris, uint32_t n_rays, const float *ro,
               const float *rd, float *out_t, uint32_t *out_tri) {
    for (uint32_t r = 0; r < n_rays; r++) {
        float closest = 1e10f;
        uint32_t hit = 0xFFFFFFFFu;
        for (uint32_t ti = 0; ti < n_tris; ti++) {
            float t = tri_t(verts, idx, ti, ro + 3 * r, rd + 3 * r);
            if (t > 0.0f && t < closest) {
                closest = t;
                hit = ti;
            }
        }
        out_t[r] = closest;
        out_tri[r] = hit;
    }
}
