/*
 * sv_oracle.h - CPU oracle for the seeVR physics (physics.comp) and LBVH
 * (bvh.comp) path.  TEST INFRASTRUCTURE ONLY: the oracle is the checker the
 * parity tests, the smoke check and the benchmark's CPU-baseline leg compare
 * against.  The product (libseevr_hip.so) never links or calls it.
 *
 * It restates, in serial C with no floating-point contraction, the
 * reference's algorithm in the canonical schedule (invocations in index
 * order, so pairs come out in lexicographic order and contacts in pair
 * order).  Every function cites the reference file:line it follows.
 *
 * OR_REAL selects float (default, the reference's precision) or double
 * (-DOR_REAL=double, symbols prefixed or64_).
 */
#ifndef SV_ORACLE_H
#define SV_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifndef OR_REAL
#define OR_REAL float
#endif

#ifdef OR_DOUBLE
#define OR(name) or64_##name
#else
#define OR(name) or_##name
#endif

typedef OR_REAL or_real;

/* Binding table of physics.comp:16-57 plus the run-time constants the
 * shader hard-codes (physics.comp:66-70). */
typedef struct or_bufs {
    or_real *body_pos, *body_quat, *body_vel, *body_omega;
    or_real *body_inv_mass, *body_inv_inertia;
    uint32_t *body_vert_start, *body_vert_count;
    uint32_t *geom_body_idx;
    or_real *geom_data, *geom_friction, *geom_restitution;
    or_real *hull_verts, *vertices, *orig_vertices;
    or_real *contact_pos, *contact_normal, *contact_pen;
    uint32_t *contact_geom_a, *contact_geom_b;
    or_real *lambda_n;
    uint32_t *counters; /* [4] */
    or_real *aabb_min, *aabb_max;
    uint32_t *pairs;
    uint32_t *support_offset, *support_lookup;
    uint32_t *hull_range; /* optional uint (start, count) per geom, else NULL */
    uint32_t max_pairs, max_contacts, pgs_iterations;
    or_real baumgarte, slop, lin_damping, ang_damping;
} or_bufs;

/* LBVH bindings of bvh.comp:11-28 (temp buffers are internal here). */
typedef struct or_bvh {
    const or_real *vertices; const uint32_t *indices;
    or_real *centroids; uint32_t *morton, *sort_idx; or_real *bounds;
    or_real *node_min, *node_max;
    uint32_t *node_left, *node_right, *node_count, *node_parent;
    uint32_t *prim_idx, *flags;
} or_bvh;

/* ---- math shared with the product by definition (not by code) -------- */
float OR(atan2f)(float y, float x);

/* ---- physics stages (physics.comp steps) ------------------------------ */
void OR(step_gravity)(or_bufs *b, uint32_t count, or_real dt, const or_real g[3]);
void OR(step_compute_aabb)(or_bufs *b, uint32_t count);
void OR(step_broad_phase)(or_bufs *b, uint32_t count);
void OR(step_narrow_phase)(or_bufs *b);
void OR(step_clear)(or_bufs *b);
void OR(step_solve)(or_bufs *b, or_real dt);
void OR(step_integrate)(or_bufs *b, uint32_t count, or_real dt);
void OR(step_update_verts)(or_bufs *b, uint32_t count, uint32_t num_bodies);
void OR(step_clear_counters)(or_bufs *b);
/* Physics.step sequencing (physics.zig:80-117). */
/* threads of the parallel stages: omp_get_max_threads() in the OpenMP build, else 1 */
int OR(threads)(void);
void OR(physics_step)(or_bufs *b, uint32_t num_bodies, uint32_t num_vertices,
                      uint32_t num_hull_verts, uint32_t substeps, float dt,
                      const float g[3]);
/* Warm start (library extension beside the reference path). */
void OR(step_solve_warm)(or_bufs *b, or_real dt);
void OR(step_solve_in_order)(or_bufs *b, or_real dt, int warm, const uint32_t *order);
void OR(warm_lookup)(or_bufs *b, const uint64_t *keys, const or_real *lam, uint32_t n_prev, or_real factor);
uint32_t OR(warm_save)(const or_bufs *b, uint64_t *keys, or_real *lam);
void OR(physics_step_warm)(or_bufs *b, uint32_t num_bodies, uint32_t num_vertices, uint32_t substeps, float dt,
                           const float g[3], or_real factor, uint64_t *keys, or_real *lam, uint32_t *n_prev);
/* Per-pair pieces, exposed for known-answer tests. */
void OR(support)(const or_bufs *b, uint32_t bi, const or_real d[3], or_real out[3]);
int OR(mpr)(const or_bufs *b, uint32_t a, uint32_t bb, or_real normal[3],
            or_real *depth, or_real contact[3]);

/* ---- LBVH (bvh.comp steps, bvh.zig:63-172) ---------------------------- */
uint32_t OR(expand_bits)(uint32_t v);
uint32_t OR(morton3d)(or_real x, or_real y, or_real z);
int OR(clz32)(uint32_t x);
void OR(bvh_centroids)(or_bvh *t, uint32_t n);
void OR(bvh_bounds)(or_bvh *t, uint32_t n);
void OR(bvh_morton)(or_bvh *t, uint32_t n);
void OR(bvh_radix_sort)(or_bvh *t, uint32_t n);
void OR(bvh_hierarchy)(or_bvh *t, uint32_t n);
void OR(bvh_leaf_aabb)(or_bvh *t, uint32_t n);
void OR(bvh_propagate)(or_bvh *t, uint32_t n);
void OR(bvh_build)(or_bvh *t, uint32_t n);
/* Body-AABB LBVH (extension): centroid = (min + max) * 0.5, leaf = AABB. */
void OR(bvh_build_bodies)(or_bvh *t, const or_real *aabb_min, const or_real *aabb_max,
                          uint32_t n);

/* ---- scene construction (data.zig) ------------------------------------ */
typedef struct or_scene {
    uint32_t cap_bodies, cap_verts, cap_tris, cap_hull, cap_lookup, cap_mat;
    uint32_t num_vertices, num_triangles, num_bodies, num_geoms, num_hull_verts;
    uint32_t num_materials, num_support_entries;
    float *vertices, *orig_vertices; uint32_t *indices;
    float *body_pos, *body_quat, *body_vel, *body_omega, *body_inv_mass;
    float *body_inertia, *body_inv_inertia;
    uint32_t *body_vert_start, *body_vert_count;
    uint32_t *geom_body_idx; float *geom_data, *geom_friction, *geom_restitution;
    uint32_t *geom_vert_start, *geom_vert_count;
    float *hull_verts; uint32_t *support_offset, *support_lookup;
    float *material_density;
} or_scene;

int OR(scene_init)(or_scene *s, uint32_t cap_bodies, uint32_t cap_verts,
                   uint32_t cap_tris, uint32_t cap_hull);
void OR(scene_free)(or_scene *s);
uint32_t OR(scene_add_material)(or_scene *s, float density);
uint32_t OR(scene_add_box)(or_scene *s, const float center[3], const float half[3],
                           float friction, float restitution, uint32_t material);
uint32_t OR(scene_add_sphere)(or_scene *s, const float center[3], float radius,
                              uint32_t segments, float friction, float restitution,
                              uint32_t material);
void OR(scene_derive)(or_scene *s);

#endif
