/*
 * kzw.h - C-ABI of the host world around the device contact path
 * (libkizuri_amd.so): the flat counterpart of the reference's
 * PhysicsWorld / BodySettings / WorldSettings API
 * (include/Kizuri/World/PhysicsWorld.h:82-239, include/Kizuri/Body/Body.h:62-90).
 *
 * A kzw world owns the shapes (convex hulls built as ConvexHullShape::Create
 * builds them), the bodies with their mass properties
 * (BodyManager::RecomputeMassProperties), the dynamic AABB tree broadphase,
 * and a kzg context that runs the narrowphase, constraint build, solve and
 * integration on the GPU.  kzw_step is PhysicsWorld::StepFixed.
 *
 * device = -1 creates a host-only world (no GPU): shapes, bodies, mass data
 * and the broadphase work, kzw_step returns KZG_ERR_DEVICE.
 * Status codes are those of kzg.h.
 */
#ifndef KZW_H
#define KZW_H

#include <stdint.h>
#include "kz_scene.h"
#include "kzg.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct kzw_world kzw_world;

typedef struct kzw_stats {
    uint32_t num_bodies;
    uint32_t num_broadphase_pairs;
    uint32_t num_contact_constraints;
    uint32_t num_contact_points;     /* WorldStats::numContactPoints */
    double broadphase_ms;            /* host: broadphase update + pair list */
    double contact_path_ms;          /* host clock around narrowphase + build + solve (device synchronised) */
    double step_ms;                  /* whole StepFixed, host clock */
    /* Inside broadphase_ms: the proxy moves (tree upkeep), and the pair
     * query (host walk, or tree image + device query with its transfers). */
    double broadphase_move_ms, broadphase_query_ms;
    uint32_t num_active_bodies;      /* WorldStats::numActiveBodies */
    uint32_t broadphase_reinserted;  /* proxies that left their fat boxes this step (MoveProxy's re-insertions) */
} kzw_stats;

int kzw_create(kzw_world** out, const kz_world_desc* world, int device);
void kzw_destroy(kzw_world* w);
const char* kzw_last_error(const kzw_world* w);

/* Append shapes; hull shapes read hull_count points at hull_offset of
 * hull_points (float3).  KZG_ERR_ARGUMENT when ConvexHullShape::Create would
 * reject a point set. */
int kzw_add_shapes(kzw_world* w, const kz_shape_desc* shapes, uint32_t n, const float* hull_points,
                   uint32_t num_hull_points);
/* Append bodies (PhysicsWorld::CreateBody); body i gets packed id = slot. */
int kzw_add_bodies(kzw_world* w, const kz_body_desc* bodies, uint32_t n);

/* PhysicsWorld::StepFixed(dt). */
int kzw_step(kzw_world* w, float dt);
/* The broadphase part of a step alone (UpdateBroadPhase + UpdatePairs). */
int kzw_update_broadphase(kzw_world* w);

int kzw_get_states(kzw_world* w, kz_body_state* out, uint32_t n);
int kzw_set_states(kzw_world* w, const kz_body_state* in, uint32_t n);
/* Pairs of the last broadphase update as (slot A, slot B); returns the count. */
int kzw_get_pairs(kzw_world* w, uint32_t* out, uint32_t cap);
/* kzw_step's broadphase: 1 (default with a device) the pair query on the
 * device over an image of the tree kept here, with the proxies' bounds and
 * fat-box tests on the device too (kzg_broadphase_move; the re-insertions
 * stay here, in order); 2 the same query with the proxy moves on the host;
 * 0 everything on the host (the walk on host threads).  The pairs and their
 * order are the same. */
int kzw_set_device_broadphase(kzw_world* w, int on);
/* Host threads for the broadphase update's per-body loops (the proxies'
 * tight boxes and fat-box tests; the re-insertions stay in list order) and
 * the tree image, the caller's included; default 1.  Like the reference's
 * JobSystem worker count, it changes no result.  More threads pay where
 * cores are free (32k bodies: 1.5 -> 0.46 ms on 4-8 threads of an idle
 * host); under a tight CPU quota they cost more than they save. */
int kzw_set_host_threads(kzw_world* w, int threads);
int kzw_get_stats(kzw_world* w, kzw_stats* out);
/* Per body: inverse mass + inverse local inertia (column-major), 10 floats. */
int kzw_get_mass(kzw_world* w, float* out, uint32_t n);
/* Built hull of shape `shape`: vertex count / triangle count returned via
 * the pointers; arrays may be null to query sizes. */
int kzw_get_hull(kzw_world* w, uint32_t shape, float* verts, uint32_t* tris, uint32_t* num_verts,
                 uint32_t* num_tris);
/* The device context of the world (null for a host-only world). */
kzg_ctx* kzw_device(kzw_world* w);

/* Snapshots (SerializeWorld / DeserializeWorld, src/Serialization/Serializer.cpp:182-234).
 * kzw_save writes the reference's KZP1 world record (bodies with their
 * shapes, settings and current state; readable by DeserializeWorld) followed
 * by an extension that makes the snapshot resumable bit for bit: world
 * settings, the shape table with its built hulls, body descriptions and
 * states, the broadphase tree, and the warm-start cache.  With out null it
 * only reports the size.
 * kzw_load creates a world from such a snapshot, or from a plain KZP1 record
 * written by the reference; a plain record carries no world settings, so
 * `settings` is required then (and ignored otherwise).  A plain record
 * rebuilds hulls, broadphase and an empty cache exactly as DeserializeWorld
 * does. */
int kzw_save(kzw_world* w, uint8_t* out, uint64_t cap, uint64_t* size);
int kzw_load(kzw_world** out, const uint8_t* data, uint64_t size, const kz_world_desc* settings, int device);

#ifdef __cplusplus
}
#endif

#endif /* KZW_H */
