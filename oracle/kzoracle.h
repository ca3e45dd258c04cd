/*
 * kzoracle.h - C-ABI of the CPU oracle (oracle/kzoracle.cpp).
 *
 * TEST INFRASTRUCTURE ONLY: a clean-room CPU restatement of the reference's
 * contact path, used as the checker by tests/, by __graft_entry__.smoke() and
 * by bench.py's cpu_baseline leg.  The product never links or loads it.
 *
 * Inputs are the scene of include/kz_scene.h plus the derived per-shape and
 * per-body data that the reference computes on body creation (built hull
 * vertices/triangles, inverse mass, inverse local inertia), so the oracle
 * covers exactly the per-step path: narrowphase, constraint build with the
 * warm-start cache, velocity integration, the island solve and integration of
 * positions.  Pairs are supplied per step in broadphase order.
 *
 * Every function exists for Real=float ("f32") and Real=double ("f64").
 */
#ifndef KZORACLE_H
#define KZORACLE_H

#include <stdint.h>
#include "kz_scene.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Built convex hull data per shape: vertices (float3 pool) and triangles. */
typedef struct kzo_hull_ref {
    uint32_t vert_offset, vert_count; /* into the built-vertex pool */
    uint32_t tri_offset, tri_count;   /* into the triangle index pool (3 u32 each) */
} kzo_hull_ref;

/* precision: 0 = float, 1 = double.  mass: 10 floats per body
 * (inverse mass, inverse local inertia column-major). */
void* kzo_world_create(int precision, const kz_world_desc* world,
                       const kz_shape_desc* shapes, uint32_t num_shapes,
                       const kzo_hull_ref* hulls, const float* hull_verts, const uint32_t* hull_tris,
                       const kz_body_desc* bodies, uint32_t num_bodies, const float* mass);
void kzo_world_destroy(void* w);

/* warm_start: 0 cache (reference default), 1 cold (the reference's KZ_NOWS). */
void kzo_set_warm_start(void* w, int mode);

/* One step over `pairs` (packed id A, packed id B per pair).  stop_after:
 * 1 after the narrowphase, 2 after the constraint build, 3 after velocity
 * integration, 4 after the solve, 5 whole step. */
void kzo_step(void* w, float dt, const uint32_t* pairs, uint32_t num_pairs, int stop_after);

uint32_t kzo_num_constraints(void* w);
void kzo_get_states(void* w, kz_body_state* out);          /* rounded to float */
void kzo_get_states_f64(void* w, double* out);             /* 13 doubles per body */
void kzo_set_states(void* w, const kz_body_state* in);
void kzo_get_manifolds(void* w, uint32_t* counts, kz_contact_point* points, uint64_t* keys);
void kzo_get_constraints(void* w, uint32_t* hdr, float* pts);  /* same layout as the reference dump */
/* PGS residual after each velocity iteration of the last solve (see kzoracle.cpp). */
uint32_t kzo_get_residuals(void* w, double* out, uint32_t cap);

/* Standalone narrowphase over pairs of (shape, transform): xf = quat xyzw +
 * translation, 7 floats.  Output: counts[n], points[n*8]. */
void kzo_collide_batch(int precision, const kz_shape_desc* shapes, uint32_t num_shapes,
                       const kzo_hull_ref* hulls, const float* hull_verts, const uint32_t* hull_tris,
                       const uint32_t* shape_a, const uint32_t* shape_b,
                       const float* xf_a, const float* xf_b, uint32_t n, float max_separation,
                       uint32_t* counts, kz_contact_point* points, int threads);

#ifdef __cplusplus
}
#endif

#endif
