/*
 * kz_scene.h - plain-data scene description shared by every C-ABI entry point.
 *
 * A scene is a table of shapes, a flat float3 pool of convex-hull input
 * points, and a table of bodies.  The fields mirror the reference engine's
 * creation structs field for field:
 *   kz_body_desc  <-> kizuri::BodySettings  (include/Kizuri/Body/Body.h:62-90)
 *   kz_world_desc <-> kizuri::WorldSettings + SolverSettings
 *                     (include/Kizuri/World/PhysicsWorld.h:20-43,
 *                      include/Kizuri/Solver/ContactSolver.h:67-82)
 * Shape type numbers are the reference's ShapeType enumerators
 * (include/Kizuri/Geometry/Shape.h:13-25).
 *
 * Bodies are created in table order with no destruction, so a body's slot
 * index equals its table index and its packed BodyID equals that index
 * (sequence 0; include/Kizuri/Body/Body.h:14-30).
 */
#ifndef KZ_SCENE_H
#define KZ_SCENE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum kz_shape_type {
    KZ_SHAPE_SPHERE = 0,      /* params: radius */
    KZ_SHAPE_BOX = 1,         /* params: half extent x, y, z */
    KZ_SHAPE_CAPSULE = 2,     /* params: half height, radius */
    KZ_SHAPE_CONVEX_HULL = 4, /* hull_offset/hull_count: input points (ConvexHullShape::Create) */
    KZ_SHAPE_PLANE = 8        /* params: normal x, y, z, distance */
};

enum kz_motion_type {
    KZ_MOTION_STATIC = 0,
    KZ_MOTION_KINEMATIC = 1,
    KZ_MOTION_DYNAMIC = 2
};

typedef struct kz_shape_desc {
    uint32_t type;
    float params[4];
    uint32_t hull_offset; /* first float3 in the hull point pool */
    uint32_t hull_count;  /* number of input points */
} kz_shape_desc;

typedef struct kz_body_desc {
    uint32_t shape;          /* index into the shape table */
    uint32_t motion_type;    /* kz_motion_type */
    uint32_t is_sensor;
    uint32_t allow_sleeping;
    float position[3];
    float rotation[4];       /* quaternion x, y, z, w */
    float linear_velocity[3];
    float angular_velocity[3];
    float friction;          /* BodySettings default 0.5 */
    float restitution;       /* 0 */
    float linear_damping;    /* 0.05 */
    float angular_damping;   /* 0.05 */
    float gravity_factor;    /* 1 */
    float max_linear_velocity;  /* 500 */
    float max_angular_velocity; /* 50 */
    float mass_override;     /* 0 = from shape and density */
    float density;           /* 1000 */
} kz_body_desc;

typedef struct kz_world_desc {
    float gravity[3];                /* (0, -9.81, 0) */
    uint32_t velocity_iterations;    /* 16 */
    uint32_t position_iterations;    /* 6 */
    float baumgarte;                 /* 0.2 */
    float penetration_slop;          /* 0.005 */
    float restitution_threshold;     /* 1.0 */
    float max_bias_velocity;         /* 3.0 */
    float warm_start_factor;         /* 1.0 */
    uint32_t use_simd_solver;        /* 1 */
    uint32_t use_velocity_bias;      /* 0 */
    uint32_t allow_sleeping;         /* 1 in the reference; every workload here uses 0 */
    uint32_t use_multithreading;     /* 1 */
    float fixed_timestep;            /* 1/60: the broadphase displacement uses it */
} kz_world_desc;

/* Post-step body state, 13 floats. */
typedef struct kz_body_state {
    float position[3];
    float rotation[4];
    float linear_velocity[3];
    float angular_velocity[3];
} kz_body_state;

/* One contact point as the narrowphase emits it (ContactPoint,
 * include/Kizuri/Collision/ContactManifold.h:20-31). */
typedef struct kz_contact_point {
    float point_on_a[3];
    float point_on_b[3];
    float normal[3];
    float separation;
    uint32_t feature_id;
} kz_contact_point;

#ifdef __cplusplus
}
#endif

#endif /* KZ_SCENE_H */
