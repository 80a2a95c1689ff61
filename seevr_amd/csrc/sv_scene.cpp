// sv_scene.cpp - host-side scene construction with the layouts and float
// arithmetic of seeVR's data.zig (addMaterial / addBox / addSphere /
// buildSupportLookup / upload), plus the seeded benchmark scenes.
// Compiled without floating-point contraction: Zig's default strict float
// mode evaluates each expression left to right with one rounding per
// operation, and so does this file.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/seevr.h"

#pragma STDC FP_CONTRACT OFF

namespace {

constexpr uint32_t kLookupThreshold = 32;  // data.zig:22
constexpr uint32_t kBins = 64;             // data.zig:21
constexpr uint32_t kEntry = 3 * kBins;     // data.zig:23

}  // namespace

struct sv_scene {
    sv_caps caps{};
    uint32_t num_vertices = 0, num_triangles = 0, num_bodies = 0, num_geoms = 0;
    uint32_t num_hull_verts = 0, num_materials = 0, num_support_entries = 0;
    std::vector<float> vertices, orig_vertices;
    std::vector<uint32_t> indices;
    std::vector<float> body_pos, body_quat, body_vel, body_omega, body_inv_mass, body_inertia, body_inv_inertia;
    std::vector<uint32_t> body_vert_start, body_vert_count;
    std::vector<uint32_t> geom_body_idx;
    std::vector<float> geom_data, geom_friction, geom_restitution;
    std::vector<uint32_t> geom_vert_start, geom_vert_count, geom_hull_range;
    std::vector<float> hull_verts;
    std::vector<uint32_t> support_offset, support_lookup;
    std::vector<float> material_density;
};

namespace {

int bad(const char *) { return SV_E_INVALID; }

// data.zig:299-353
void build_support_lookup(sv_scene *s, uint32_t gi, uint32_t hs, uint32_t hc) {
    if (hc <= kLookupThreshold) {
        s->support_offset[gi] = 0;
        return;
    }
    const uint32_t start = s->num_support_entries == 0 ? 1u : s->num_support_entries;
    s->support_offset[gi] = start;
    if (s->support_lookup.size() < (size_t)start + kEntry) s->support_lookup.resize((size_t)start + kEntry, 0);
    static const float perps[3][2][3] = {{{0, 1, 0}, {0, 0, 1}}, {{1, 0, 0}, {0, 0, 1}}, {{1, 0, 0}, {0, 1, 0}}};
    const float pi = 3.14159265358979323846f;  // std.math.pi as f32
    const float *hv = s->hull_verts.data() + (size_t)hs * 3;
    for (uint32_t ax = 0; ax < 3; ax++) {
        const float *p1 = perps[ax][0], *p2 = perps[ax][1];
        for (uint32_t bin = 0; bin < kBins; bin++) {
            float angle = (float)bin / (float)kBins * 2.0f * pi;
            float ca = cosf(angle), sa = sinf(angle);
            float dx = ca * p1[0] + sa * p2[0];
            float dy = ca * p1[1] + sa * p2[1];
            float dz = ca * p1[2] + sa * p2[2];
            float best = -1e10f;
            uint32_t best_i = 0;
            for (uint32_t v = 0; v < hc; v++) {
                const float *h = hv + (size_t)v * 3;
                float d = dx * h[0] + dy * h[1] + dz * h[2];
                if (d > best) {
                    best = d;
                    best_i = v;
                }
            }
            s->support_lookup[start + ax * kBins + bin] = best_i;
        }
    }
    s->num_support_entries = start + kEntry;
}

bool room(const sv_scene *s, uint32_t verts, uint32_t tris, uint32_t hull) {
    const sv_caps &k = s->caps;
    return s->num_bodies < k.max_bodies && (uint64_t)s->num_vertices + verts <= k.max_vertices &&
           (uint64_t)s->num_triangles + tris <= k.max_triangles &&
           (uint64_t)s->num_hull_verts + hull <= k.max_hull_verts &&
           (hull <= kLookupThreshold ||
            (uint64_t)(s->num_support_entries ? s->num_support_entries : 1u) + kEntry <= k.max_support_entries);
}

void push_body(sv_scene *s, const float c[3], float inv_mass, const float inertia[3], const float inv_i[3]) {
    for (int k = 0; k < 3; k++) {
        s->body_pos.push_back(c[k]);
        s->body_vel.push_back(0.0f);
        s->body_omega.push_back(0.0f);
        s->body_inertia.push_back(inertia[k]);
        s->body_inv_inertia.push_back(inv_i[k]);
    }
    s->body_quat.insert(s->body_quat.end(), {0.0f, 0.0f, 0.0f, 1.0f});
    s->body_inv_mass.push_back(inv_mass);
    s->body_vert_start.push_back(0);
    s->body_vert_count.push_back(0);
    s->num_bodies++;
}

void push_geom(sv_scene *s, uint32_t bi, uint32_t hv, uint32_t hc, float fr, float re, uint32_t vs, uint32_t vc) {
    s->geom_body_idx.push_back(bi);
    s->geom_data.insert(s->geom_data.end(), {(float)hv, (float)hc, 0, 0, 0, 0, 0, 0});
    s->geom_friction.push_back(fr);
    s->geom_restitution.push_back(re);
    s->geom_vert_start.push_back(vs);
    s->geom_vert_count.push_back(vc);
    s->geom_hull_range.push_back(hv);
    s->geom_hull_range.push_back(hc);
    s->support_offset.push_back(0);
}

// ---- seeded workloads -----------------------------------------------------
struct Rng {
    uint64_t x;
    uint64_t next() {  // splitmix64
        uint64_t z = (x += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    float uni() { return (float)(next() >> 40) * (1.0f / 16777216.0f); }  // [0, 1)
    float uni(float lo, float hi) { return lo + (hi - lo) * uni(); }
};

// Uniform random rotation (Shoemake, Graphics Gems III), as (x, y, z, w).
void random_quat(Rng &r, float q[4]) {
    double u1 = r.uni(), u2 = r.uni(), u3 = r.uni();
    double a = std::sqrt(1.0 - u1), b = std::sqrt(u1), t1 = 2.0 * M_PI * u2, t2 = 2.0 * M_PI * u3;
    q[0] = (float)(a * std::sin(t1));
    q[1] = (float)(a * std::cos(t1));
    q[2] = (float)(b * std::sin(t2));
    q[3] = (float)(b * std::cos(t2));
}

int add_random_body(sv_scene *s, Rng &r, float p_sphere, float hlo, float hhi, float rlo, float rhi,
                    uint32_t segs, const float pos[3], uint32_t mat) {
    static const float col[3] = {0.8f, 0.8f, 0.8f};
    uint32_t bi;
    int rc;
    if (r.uni() < p_sphere) {
        float rad = r.uni(rlo, rhi);
        rc = sv_scene_add_sphere(s, pos, rad, col, segs, 0.5f, 0.1f, mat, &bi);
    } else {
        float h[3] = {r.uni(hlo, hhi), r.uni(hlo, hhi), r.uni(hlo, hhi)};
        rc = sv_scene_add_box(s, pos, h, col, 0.5f, 0.1f, mat, &bi);
    }
    if (rc) return rc;
    float q[4];
    random_quat(r, q);
    static const float zero[3] = {0, 0, 0};
    return sv_scene_set_body_state(s, bi, q, zero, zero);
}

int static_box(sv_scene *s, float cx, float cy, float cz, float hx, float hy, float hz, uint32_t mat) {
    static const float col[3] = {0.35f, 0.35f, 0.30f};
    float c[3] = {cx, cy, cz}, h[3] = {hx, hy, hz};
    uint32_t bi;
    return sv_scene_add_box(s, c, h, col, 0.5f, 0.1f, mat, &bi);
}

}  // namespace

extern "C" {

int sv_scene_create(const sv_caps *caps, sv_scene **out) {
    if (!caps || !out) return bad("null");
    sv_scene *s = new (std::nothrow) sv_scene();
    if (!s) return SV_E_NOMEM;
    s->caps = *caps;
    *out = s;
    return SV_OK;
}

int sv_scene_destroy(sv_scene *s) {
    delete s;
    return SV_OK;
}

// data.zig:356-370 (albedo, roughness, metallic, emission and ior feed only
// the path tracer; density is the one physics input)
int sv_scene_add_material(sv_scene *s, const float *, float, float, const float *, float, float density,
                          uint32_t *out) {
    if (!s) return bad("null");
    if (s->num_materials >= 256) return SV_E_CAPACITY;  // data.zig:18
    s->material_density.push_back(density);
    if (out) *out = s->num_materials;
    s->num_materials++;
    return SV_OK;
}

// data.zig:392-485
int sv_scene_add_box(sv_scene *s, const float c[3], const float h[3], const float *, float friction,
                     float restitution, uint32_t material, uint32_t *out) {
    if (!s || !c || !h || material >= s->num_materials) return bad("arg");
    if (!room(s, 8, 12, 8)) return SV_E_CAPACITY;
    const uint32_t vs = s->num_vertices, ts = s->num_triangles, bi = s->num_bodies;
    const float bv[8][3] = {{-h[0], -h[1], -h[2]}, {h[0], -h[1], -h[2]}, {h[0], h[1], -h[2]}, {-h[0], h[1], -h[2]},
                            {-h[0], -h[1], h[2]},  {h[0], -h[1], h[2]},  {h[0], h[1], h[2]},  {-h[0], h[1], h[2]}};
    static const uint32_t bf[12][3] = {{0, 2, 1}, {0, 3, 2}, {5, 7, 4}, {5, 6, 7}, {4, 3, 0}, {4, 7, 3},
                                       {1, 6, 5}, {1, 2, 6}, {3, 6, 2}, {3, 7, 6}, {4, 1, 5}, {4, 0, 1}};
    for (int i = 0; i < 8; i++)
        for (int j = 0; j < 3; j++) {
            s->orig_vertices.push_back(bv[i][j]);
            s->vertices.push_back(c[j] + bv[i][j]);
        }
    for (int i = 0; i < 12; i++)
        for (int j = 0; j < 3; j++) s->indices.push_back(vs + bf[i][j]);
    s->num_vertices += 8;
    s->num_triangles += 12;
    const float density = s->material_density[material];
    const float volume = h[0] * 2 * h[1] * 2 * h[2] * 2;
    const float mass = density * volume;
    float inertia[3] = {0, 0, 0}, inv_i[3] = {0, 0, 0};
    if (mass > 0) {
        const float wx = h[0] * 2, wy = h[1] * 2, wz = h[2] * 2;
        inertia[0] = mass / 12.0f * (wy * wy + wz * wz);
        inertia[1] = mass / 12.0f * (wx * wx + wz * wz);
        inertia[2] = mass / 12.0f * (wx * wx + wy * wy);
        for (int k = 0; k < 3; k++) inv_i[k] = 1.0f / inertia[k];
    }
    push_body(s, c, mass > 0 ? 1.0f / mass : 0.0f, inertia, inv_i);
    const uint32_t hv = s->num_hull_verts;
    for (int i = 0; i < 8; i++)
        for (int j = 0; j < 3; j++) s->hull_verts.push_back(bv[i][j]);
    s->num_hull_verts += 8;
    const uint32_t gi = s->num_geoms;
    push_geom(s, bi, hv, 8, friction, restitution, vs, 8);
    build_support_lookup(s, gi, hv, 8);
    s->num_geoms++;
    if (out) *out = bi;
    return SV_OK;
}

// data.zig:487-597
int sv_scene_add_sphere(sv_scene *s, const float c[3], float radius, const float *, uint32_t segments,
                        float friction, float restitution, uint32_t material, uint32_t *out) {
    if (!s || !c || segments == 0 || material >= s->num_materials) return bad("arg");
    const uint32_t nv = (segments + 1) * (segments + 1), nt = segments * segments * 2;
    if (!room(s, nv, nt, nv)) return SV_E_CAPACITY;
    const uint32_t vs = s->num_vertices, ts = s->num_triangles, bi = s->num_bodies;
    const float pi = 3.14159265f;
    const float segs = (float)segments;
    for (uint32_t i = 0; i <= segments; i++) {
        const float fi = (float)i;
        const float lat = pi * fi / segs;
        const float sl = sinf(lat), cl = cosf(lat);
        for (uint32_t j = 0; j <= segments; j++) {
            const float fj = (float)j;
            const float lon = 2.0f * pi * fj / segs;
            const float lx = radius * cosf(lon) * sl;
            const float ly = radius * cl;
            const float lz = radius * sinf(lon) * sl;
            s->orig_vertices.insert(s->orig_vertices.end(), {lx, ly, lz});
            s->vertices.insert(s->vertices.end(), {c[0] + lx, c[1] + ly, c[2] + lz});
        }
    }
    s->num_vertices += nv;
    for (uint32_t i = 0; i < segments; i++)
        for (uint32_t j = 0; j < segments; j++) {
            const uint32_t cur = vs + i * (segments + 1) + j, nxt = cur + segments + 1;
            s->indices.insert(s->indices.end(), {cur, nxt, cur + 1, cur + 1, nxt, nxt + 1});
        }
    (void)ts;
    s->num_triangles += nt;
    const float density = s->material_density[material];
    const float volume = (4.0f / 3.0f) * pi * radius * radius * radius;
    const float mass = density * volume;
    float inertia[3] = {0, 0, 0}, inv_i[3] = {0, 0, 0};
    if (mass > 0) {
        const float I = 0.4f * mass * radius * radius;
        for (int k = 0; k < 3; k++) {
            inertia[k] = I;
            inv_i[k] = 1.0f / I;
        }
    }
    push_body(s, c, mass > 0 ? 1.0f / mass : 0.0f, inertia, inv_i);
    const uint32_t hv = s->num_hull_verts;
    s->hull_verts.insert(s->hull_verts.end(), s->orig_vertices.begin() + (size_t)vs * 3,
                         s->orig_vertices.begin() + (size_t)(vs + nv) * 3);
    s->num_hull_verts += nv;
    const uint32_t gi = s->num_geoms;
    push_geom(s, bi, hv, nv, friction, restitution, vs, nv);
    build_support_lookup(s, gi, hv, nv);
    s->num_geoms++;
    if (out) *out = bi;
    return SV_OK;
}

int sv_scene_set_body_state(sv_scene *s, uint32_t b, const float q[4], const float v[3], const float w[3]) {
    if (!s || b >= s->num_bodies) return bad("arg");
    if (q) std::memcpy(&s->body_quat[(size_t)b * 4], q, 4 * sizeof(float));
    if (v) std::memcpy(&s->body_vel[(size_t)b * 3], v, 3 * sizeof(float));
    if (w) std::memcpy(&s->body_omega[(size_t)b * 3], w, 3 * sizeof(float));
    return SV_OK;
}

int sv_scene_get_counts(const sv_scene *s, sv_scene_counts *o) {
    if (!s || !o) return bad("null");
    o->num_vertices = s->num_vertices;
    o->num_triangles = s->num_triangles;
    o->num_bodies = s->num_bodies;
    o->num_geoms = s->num_geoms;
    o->num_hull_verts = s->num_hull_verts;
    o->num_materials = s->num_materials;
    o->num_support_entries = s->num_support_entries;
    return SV_OK;
}

static void derive(sv_scene *s) {
    // data.zig:633-640: body vert range from the body's geom
    for (uint32_t gi = 0; gi < s->num_geoms; gi++) {
        uint32_t bi = s->geom_body_idx[gi];
        s->body_vert_start[bi] = s->geom_vert_start[gi];
        s->body_vert_count[bi] = s->geom_vert_count[gi];
    }
}

int sv_scene_array(const sv_scene *cs, sv_buf id, const void **host, size_t *bytes) {
    if (!cs || !host || !bytes) return bad("null");
    sv_scene *s = const_cast<sv_scene *>(cs);
    derive(s);
    auto F = [&](const std::vector<float> &v) { *host = v.data(); *bytes = v.size() * sizeof(float); return SV_OK; };
    auto U = [&](const std::vector<uint32_t> &v) { *host = v.data(); *bytes = v.size() * sizeof(uint32_t); return SV_OK; };
    switch (id) {
    case SV_BODY_POS: return F(s->body_pos);
    case SV_BODY_QUAT: return F(s->body_quat);
    case SV_BODY_VEL: return F(s->body_vel);
    case SV_BODY_OMEGA: return F(s->body_omega);
    case SV_BODY_INV_MASS: return F(s->body_inv_mass);
    case SV_BODY_INV_INERTIA: return F(s->body_inv_inertia);
    case SV_BODY_VERT_START: return U(s->body_vert_start);
    case SV_BODY_VERT_COUNT: return U(s->body_vert_count);
    case SV_GEOM_BODY_IDX: return U(s->geom_body_idx);
    case SV_GEOM_DATA: return F(s->geom_data);
    case SV_GEOM_FRICTION: return F(s->geom_friction);
    case SV_GEOM_RESTITUTION: return F(s->geom_restitution);
    case SV_HULL_VERTS: return F(s->hull_verts);
    case SV_VERTICES: return F(s->vertices);
    case SV_ORIGINAL_VERTICES: return F(s->orig_vertices);
    case SV_SUPPORT_LOOKUP_OFFSET: return U(s->support_offset);
    case SV_SUPPORT_LOOKUP: return U(s->support_lookup);
    case SV_INDICES: return U(s->indices);
    case SV_GEOM_HULL_RANGE: return U(s->geom_hull_range);
    default: return bad("buffer id has no host staging array");
    }
}

// data.zig:599-655 (the physics and BVH inputs)
int sv_scene_upload(sv_scene *s, sv_ctx *ctx) {
    if (!s || !ctx) return bad("null");
    derive(s);
    static const sv_buf ids[] = {SV_VERTICES, SV_INDICES, SV_ORIGINAL_VERTICES, SV_BODY_POS, SV_BODY_QUAT,
                                 SV_BODY_VEL, SV_BODY_OMEGA, SV_BODY_INV_MASS, SV_BODY_INV_INERTIA,
                                 SV_GEOM_BODY_IDX, SV_GEOM_DATA, SV_GEOM_FRICTION, SV_GEOM_RESTITUTION,
                                 SV_HULL_VERTS, SV_SUPPORT_LOOKUP_OFFSET, SV_SUPPORT_LOOKUP, SV_BODY_VERT_START,
                                 SV_BODY_VERT_COUNT, SV_GEOM_HULL_RANGE};
    for (sv_buf id : ids) {
        const void *p;
        size_t n;
        int rc = sv_scene_array(s, id, &p, &n);
        if (rc) return rc;
        rc = sv_upload(ctx, id, p, 0, n);
        if (rc) return rc;
    }
    return SV_OK;
}

int sv_scene_build_workload(sv_scene *s, int workload, uint32_t count, uint32_t segs, uint64_t seed) {
    if (!s || s->num_bodies) return bad("scene must be empty");
    Rng r{seed * 0x9E3779B97F4A7C15ull + (uint64_t)workload};
    const float zero3[3] = {0, 0, 0};
    uint32_t m_static, m_dyn;
    int rc = sv_scene_add_material(s, zero3, 0.95f, 0.0f, zero3, 0.0f, 0.0f, &m_static);
    if (!rc) rc = sv_scene_add_material(s, zero3, 0.5f, 0.0f, zero3, 0.0f, 1000.0f, &m_dyn);
    if (rc) return rc;
    switch (workload) {
    case 1: {  // 1,000 bodies in a 10x10 m column over a 50x1x50 m ground
        uint32_t n = count ? count : 1000, sg = segs ? segs : 16;
        if ((rc = static_box(s, 0, -0.5f, 0, 25, 0.5f, 25, m_static))) return rc;
        for (uint32_t i = 0; i < n && !rc; i++) {
            float p[3] = {r.uni(-5, 5), r.uni(1, 20), r.uni(-5, 5)};
            rc = add_random_body(s, r, 0.5f, 0.2f, 0.6f, 0.2f, 0.5f, sg, p, m_dyn);
        }
        return rc;
    }
    case 2: {  // 32,768 bodies in a 30x30 m container, ground + 4 walls
        uint32_t n = count ? count : 32768, sg = segs ? segs : 16;
        if ((rc = static_box(s, 0, -0.5f, 0, 16, 0.5f, 16, m_static))) return rc;
        if ((rc = static_box(s, 15.5f, 30, 0, 0.5f, 30, 16, m_static))) return rc;
        if ((rc = static_box(s, -15.5f, 30, 0, 0.5f, 30, 16, m_static))) return rc;
        if ((rc = static_box(s, 0, 30, 15.5f, 16, 30, 0.5f, m_static))) return rc;
        if ((rc = static_box(s, 0, 30, -15.5f, 16, 30, 0.5f, m_static))) return rc;
        for (uint32_t i = 0; i < n && !rc; i++) {
            float p[3] = {r.uni(-14.4f, 14.4f), r.uni(1, 60), r.uni(-14.4f, 14.4f)};
            rc = add_random_body(s, r, 0.5f, 0.2f, 0.6f, 0.2f, 0.5f, sg, p, m_dyn);
        }
        return rc;
    }
    case 4: {  // small scene over its own ground (differentiable workload)
        uint32_t n = count ? count : 64, sg = segs ? segs : 16;
        if ((rc = static_box(s, 0, -0.5f, 0, 10, 0.5f, 10, m_static))) return rc;
        for (uint32_t i = 0; i < n && !rc; i++) {
            float p[3] = {r.uni(-3, 3), r.uni(0.7f, 6), r.uni(-3, 3)};
            rc = add_random_body(s, r, 0.5f, 0.2f, 0.6f, 0.2f, 0.5f, sg, p, m_dyn);
        }
        return rc;
    }
    case 5: {  // 1,048,576 bodies, 70% spheres, 200x200x50 m over 220x1x220 m
        uint32_t n = count ? count : 1048576, sg = segs ? segs : 8;
        if ((rc = static_box(s, 0, -0.5f, 0, 110, 0.5f, 110, m_static))) return rc;
        for (uint32_t i = 0; i < n && !rc; i++) {
            float p[3] = {r.uni(-100, 100), r.uni(1, 51), r.uni(-100, 100)};
            rc = add_random_body(s, r, 0.7f, 0.1f, 0.4f, 0.1f, 0.3f, sg, p, m_dyn);
        }
        return rc;
    }
    default:
        return bad("unknown workload");
    }
}

// Independent scenes for the batched engine, each with its own ground at the
// origin: scene k has n_k = floor(exp(U[ln min_n, ln max_n])) bodies (the
// box/sphere mix of the 1,000-body column) in a 10x10 m column at heights
// 1 m to `top` m (20 m when 0) over a 50x1x50 m ground, and a gravity scale U[0.8, 1.2].  Scene k
// draws from its own stream (seed, k), so building scenes [first,
// first+count) alone gives the same bodies as building them in a larger
// batch.  Writes body_start[0..count] (first body of each scene, then the
// end) and gravity_scale[0..count).
int sv_scene_build_batch(sv_scene *s, uint32_t first, uint32_t count, uint32_t min_n, uint32_t max_n, uint32_t segs,
                         float top, uint64_t seed, uint32_t *body_start, float *gravity_scale) {
    const float hi = top > 1.0f ? top : 20.0f;
    if (!s || !body_start || !gravity_scale) return bad("null");
    if (s->num_bodies) return bad("scene must be empty");
    if (min_n < 1 || max_n < min_n) return bad("body range");
    const float zero3[3] = {0, 0, 0};
    uint32_t m_static, m_dyn;
    int rc = sv_scene_add_material(s, zero3, 0.95f, 0.0f, zero3, 0.0f, 0.0f, &m_static);
    if (!rc) rc = sv_scene_add_material(s, zero3, 0.5f, 0.0f, zero3, 0.0f, 1000.0f, &m_dyn);
    const uint32_t sg = segs ? segs : 16;
    for (uint32_t k = first; k < first + count && !rc; k++) {
        Rng r{(seed + 1) * 0xD1B54A32D192ED03ull ^ ((uint64_t)k * 0x9E3779B97F4A7C15ull)};
        float ln = r.uni(std::log((float)min_n), std::log((float)max_n));
        uint32_t n = (uint32_t)std::floor(std::exp(ln));
        if (n < min_n) n = min_n;
        if (n > max_n) n = max_n;
        gravity_scale[k - first] = r.uni(0.8f, 1.2f);
        body_start[k - first] = s->num_bodies;
        rc = static_box(s, 0, -0.5f, 0, 25, 0.5f, 25, m_static);
        for (uint32_t i = 0; i < n && !rc; i++) {
            float p[3] = {r.uni(-5, 5), r.uni(1, hi), r.uni(-5, 5)};
            rc = add_random_body(s, r, 0.5f, 0.2f, 0.6f, 0.2f, 0.5f, sg, p, m_dyn);
        }
    }
    body_start[count] = s->num_bodies;
    return rc;
}

}  // extern "C"

// ---- scene files (library extension; README.md:99 lists save/load) -------
//
// Little-endian, one file per scene:
//   0  char[8]  magic "SVSCENE\0"
//   8  u32      version (1)
//  12  u32      section count S
//  16  u32[7]   counts: vertices, triangles, bodies, geoms, hull verts,
//               materials, support entries (sv_scene_counts order)
//  44  u32      0
//  48  S x 32-byte entries {char tag[4]; u32 element bytes (4);
//                            u64 payload offset; u64 payload bytes;
//                            u64 FNV-1a 64 of the payload}
//  payloads at 64-byte aligned offsets, in table order.
// Sections are the scene's source arrays; body_vert_start/count and the
// hull ranges are derived again on load.  Every count, index and range is
// checked before anything is replaced.
namespace sv {
int set_error(int code, const char *msg);
}

namespace {

constexpr char kMagic[8] = {'S', 'V', 'S', 'C', 'E', 'N', 'E', '\0'};
constexpr uint32_t kVersion = 1;
constexpr size_t kHeader = 48, kEntryBytes = 32, kAlign = 64;

struct Section {
    char tag[5];
    bool is_float;
    std::vector<float> *f;
    std::vector<uint32_t> *u;
    uint64_t elems;  // expected element count from the counts
};

std::vector<Section> sections(sv_scene *s, const uint32_t n[7]) {
    const uint64_t nv = n[0], nt = n[1], nb = n[2], ng = n[3], nh = n[4], nm = n[5], ns = n[6];
    return {
        {"VERT", true, &s->vertices, nullptr, 3 * nv},         {"OVRT", true, &s->orig_vertices, nullptr, 3 * nv},
        {"INDX", false, nullptr, &s->indices, 3 * nt},          {"BPOS", true, &s->body_pos, nullptr, 3 * nb},
        {"BQUT", true, &s->body_quat, nullptr, 4 * nb},         {"BVEL", true, &s->body_vel, nullptr, 3 * nb},
        {"BOMG", true, &s->body_omega, nullptr, 3 * nb},        {"BIMS", true, &s->body_inv_mass, nullptr, nb},
        {"BINE", true, &s->body_inertia, nullptr, 3 * nb},      {"BINI", true, &s->body_inv_inertia, nullptr, 3 * nb},
        {"GBDY", false, nullptr, &s->geom_body_idx, ng},        {"GDAT", true, &s->geom_data, nullptr, 8 * ng},
        {"GFRC", true, &s->geom_friction, nullptr, ng},         {"GRST", true, &s->geom_restitution, nullptr, ng},
        {"GVST", false, nullptr, &s->geom_vert_start, ng},      {"GVCT", false, nullptr, &s->geom_vert_count, ng},
        {"HULL", true, &s->hull_verts, nullptr, 3 * nh},        {"SLOF", false, nullptr, &s->support_offset, ng},
        {"SLUT", false, nullptr, &s->support_lookup, ns},       {"MDEN", true, &s->material_density, nullptr, nm},
    };
}

uint64_t fnv1a(const uint8_t *p, size_t n) {
    uint64_t h = 0xCBF29CE484222325ull;
    for (size_t i = 0; i < n; i++) {
        h ^= p[i];
        h *= 0x100000001B3ull;
    }
    return h;
}

void put32(uint8_t *p, uint32_t v) { std::memcpy(p, &v, 4); }
void put64(uint8_t *p, uint64_t v) { std::memcpy(p, &v, 8); }
uint32_t get32(const uint8_t *p) { uint32_t v; std::memcpy(&v, p, 4); return v; }
uint64_t get64(const uint8_t *p) { uint64_t v; std::memcpy(&v, p, 8); return v; }
size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

int file_error(const char *msg) { return sv::set_error(SV_E_INVALID, msg); }

// Structural checks of a loaded scene: every index the physics and BVH
// steps follow stays inside its array.
const char *check_scene(const sv_scene &s) {
    const uint32_t nb = s.num_bodies, nv = s.num_vertices, nh = s.num_hull_verts;
    if (s.num_geoms != nb) return "scene file: geoms and bodies must be 1:1";
    for (uint32_t i : s.indices)
        if (i >= nv) return "scene file: triangle index out of range";
    for (uint32_t g = 0; g < s.num_geoms; g++) {
        if (s.geom_body_idx[g] >= nb) return "scene file: geom body index out of range";
        const float hs = s.geom_data[8 * g], hc = s.geom_data[8 * g + 1];
        if (!(hs >= 0 && hc >= 1 && hs == (float)(uint32_t)hs && hc == (float)(uint32_t)hc &&
              (uint64_t)hs + (uint64_t)hc <= nh))
            return "scene file: hull range out of range";
        if ((uint64_t)s.geom_vert_start[g] + s.geom_vert_count[g] > nv) return "scene file: vertex range out of range";
        const uint32_t so = s.support_offset[g];
        if (so != 0 && (uint64_t)so + kEntry > s.num_support_entries) return "scene file: support lookup out of range";
        if (so != 0)
            for (uint32_t k = 0; k < kEntry; k++)
                if (s.support_lookup[so + k] >= (uint32_t)hc) return "scene file: support lookup entry out of range";
    }
    if (s.num_materials > 256) return "scene file: more than 256 materials";
    return nullptr;
}

}  // namespace

extern "C" {

int sv_scene_serialize(const sv_scene *cs, void *dst, size_t cap, size_t *bytes_out) {
    if (!cs || !bytes_out) return bad("null");
    sv_scene *s = const_cast<sv_scene *>(cs);
    const uint32_t n[7] = {s->num_vertices, s->num_triangles, s->num_bodies, s->num_geoms,
                           s->num_hull_verts, s->num_materials, s->num_support_entries};
    auto secs = sections(s, n);
    size_t off = align_up(kHeader + kEntryBytes * secs.size());
    std::vector<size_t> offs;
    for (auto &x : secs) {
        offs.push_back(off);
        off = align_up(off + x.elems * 4);
    }
    *bytes_out = off;
    if (!dst) return SV_OK;
    if (cap < off) return sv::set_error(SV_E_RANGE, "scene serialize: destination smaller than the scene file");
    uint8_t *p = static_cast<uint8_t *>(dst);
    std::memset(p, 0, off);
    std::memcpy(p, kMagic, 8);
    put32(p + 8, kVersion);
    put32(p + 12, (uint32_t)secs.size());
    for (int i = 0; i < 7; i++) put32(p + 16 + 4 * i, n[i]);
    for (size_t i = 0; i < secs.size(); i++) {
        const Section &x = secs[i];
        const void *src = x.is_float ? (const void *)x.f->data() : (const void *)x.u->data();
        const size_t sz = (x.is_float ? x.f->size() : x.u->size());
        if (sz != x.elems) return sv::set_error(SV_E_INVALID, "scene serialize: array size disagrees with counts");
        uint8_t *e = p + kHeader + kEntryBytes * i;
        std::memcpy(e, x.tag, 4);
        put32(e + 4, 4);
        put64(e + 8, offs[i]);
        put64(e + 16, x.elems * 4);
        if (sz) std::memcpy(p + offs[i], src, x.elems * 4);
        put64(e + 24, fnv1a(p + offs[i], x.elems * 4));
    }
    return SV_OK;
}

int sv_scene_deserialize(sv_scene *s, const void *src, size_t bytes) {
    if (!s || (!src && bytes)) return bad("null");
    const uint8_t *p = static_cast<const uint8_t *>(src);
    if (bytes < kHeader || std::memcmp(p, kMagic, 8) != 0) return file_error("scene file: bad magic");
    if (get32(p + 8) != kVersion) return file_error("scene file: unsupported version");
    const uint32_t ns = get32(p + 12);
    uint32_t n[7];
    for (int i = 0; i < 7; i++) n[i] = get32(p + 16 + 4 * i);
    if (kHeader + (uint64_t)kEntryBytes * ns > bytes) return file_error("scene file: truncated section table");
    const sv_caps &k = s->caps;
    if (n[2] > k.max_bodies || n[0] > k.max_vertices || n[1] > k.max_triangles || n[4] > k.max_hull_verts ||
        n[6] > k.max_support_entries)
        return sv::set_error(SV_E_CAPACITY, "scene file: counts exceed the scene's capacities");
    sv_scene t;
    t.caps = s->caps;
    auto secs = sections(&t, n);
    std::vector<bool> seen(secs.size(), false);
    for (uint32_t i = 0; i < ns; i++) {
        const uint8_t *e = p + kHeader + kEntryBytes * i;
        size_t j = 0;
        while (j < secs.size() && std::memcmp(e, secs[j].tag, 4) != 0) j++;
        if (j == secs.size()) continue;  // unknown sections are skipped (forward compatible)
        if (seen[j]) return file_error("scene file: duplicate section");
        const uint64_t off = get64(e + 8), len = get64(e + 16);
        if (get32(e + 4) != 4) return file_error("scene file: element size must be 4");
        if (len != secs[j].elems * 4) return file_error("scene file: section length disagrees with counts");
        if (off > bytes || len > bytes - off) return file_error("scene file: section outside the file");
        if (fnv1a(p + off, len) != get64(e + 24)) return file_error("scene file: section checksum mismatch");
        if (secs[j].is_float) {
            secs[j].f->resize(secs[j].elems);
            if (len) std::memcpy(secs[j].f->data(), p + off, len);
        } else {
            secs[j].u->resize(secs[j].elems);
            if (len) std::memcpy(secs[j].u->data(), p + off, len);
        }
        seen[j] = true;
    }
    for (size_t j = 0; j < secs.size(); j++)
        if (!seen[j]) return file_error("scene file: missing section");
    t.num_vertices = n[0];
    t.num_triangles = n[1];
    t.num_bodies = n[2];
    t.num_geoms = n[3];
    t.num_hull_verts = n[4];
    t.num_materials = n[5];
    t.num_support_entries = n[6];
    if (const char *why = check_scene(t)) return file_error(why);
    t.body_vert_start.assign(t.num_bodies, 0);
    t.body_vert_count.assign(t.num_bodies, 0);
    t.geom_hull_range.resize(2 * (size_t)t.num_geoms);
    for (uint32_t g = 0; g < t.num_geoms; g++) {
        t.geom_hull_range[2 * g] = (uint32_t)t.geom_data[8 * g];
        t.geom_hull_range[2 * g + 1] = (uint32_t)t.geom_data[8 * g + 1];
    }
    derive(&t);
    *s = std::move(t);
    return SV_OK;
}

int sv_scene_save(const sv_scene *s, const char *path) {
    if (!s || !path) return bad("null");
    size_t n = 0;
    int rc = sv_scene_serialize(s, nullptr, 0, &n);
    if (rc) return rc;
    std::vector<uint8_t> buf(n);
    rc = sv_scene_serialize(s, buf.data(), n, &n);
    if (rc) return rc;
    FILE *f = std::fopen(path, "wb");
    if (!f) return sv::set_error(SV_E_INVALID, "scene save: cannot open the file for writing");
    const bool ok = std::fwrite(buf.data(), 1, n, f) == n;
    const bool closed = std::fclose(f) == 0;
    return ok && closed ? SV_OK : sv::set_error(SV_E_INVALID, "scene save: write failed");
}

int sv_scene_load(sv_scene *s, const char *path) {
    if (!s || !path) return bad("null");
    FILE *f = std::fopen(path, "rb");
    if (!f) return sv::set_error(SV_E_INVALID, "scene load: cannot open the file");
    std::vector<uint8_t> buf;
    uint8_t chunk[1 << 16];
    size_t got;
    while ((got = std::fread(chunk, 1, sizeof chunk, f)) > 0) buf.insert(buf.end(), chunk, chunk + got);
    const bool err = std::ferror(f) != 0;
    std::fclose(f);
    if (err) return sv::set_error(SV_E_INVALID, "scene load: read failed");
    return sv_scene_deserialize(s, buf.data(), buf.size());
}

// Copies the body state and render vertices of a context back into the
// scene (positions, orientations, velocities after simulated frames), so a
// saved scene resumes the simulation where the context stands.
int sv_scene_read_state(sv_scene *s, sv_ctx *ctx) {
    if (!s || !ctx) return bad("null");
    struct { sv_buf id; std::vector<float> *v; } io[] = {
        {SV_BODY_POS, &s->body_pos}, {SV_BODY_QUAT, &s->body_quat}, {SV_BODY_VEL, &s->body_vel},
        {SV_BODY_OMEGA, &s->body_omega}, {SV_VERTICES, &s->vertices}};
    for (auto &x : io) {
        if (x.v->empty()) continue;
        int rc = sv_download(ctx, x.id, x.v->data(), 0, x.v->size() * sizeof(float));
        if (rc) return rc;
    }
    return SV_OK;
}

}  // extern "C"
