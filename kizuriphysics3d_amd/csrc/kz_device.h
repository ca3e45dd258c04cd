// kz_device.h - device buffer layout and kernel entry points of the contact path.
//
// All per-body, per-pair and per-point data live in HBM as structure-of-arrays
// of float4/uint4 so that every gather of a kernel is one 16-byte load.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sweep_limits.h"

namespace kz {

// Shape types (the reference's ShapeType numbering, Shape.h:13-25).
enum : uint32_t { kSphere = 0, kBox = 1, kCapsule = 2, kHull = 4, kPlane = 8 };
// Narrowphase work lists.  An item is the pair index (bits 0-26), its
// (type A, type B) class (bits 27-30; sphere, box, capsule, hull -> 0..3)
// and the swap flag (bit 31).  The GJK kernel's list is sorted by class in
// chunks of kWorkChunk items, so a wave of that kernel mostly holds one kind
// of pair.
constexpr uint32_t kWorkClasses = 16;
constexpr uint32_t kWorkChunk = 1024;
constexpr uint32_t kWorkPairBits = 27;
constexpr uint32_t kWorkPairMask = (1u << kWorkPairBits) - 1u;
// Entries of a work buffer for `cap` pairs: the list, the class-major list,
// the list handed to the wave kernel, and the per-chunk class counts.
inline size_t work_entries(uint32_t cap) {
    return 3 * size_t(cap) + size_t(kWorkClasses) * ((cap + kWorkChunk - 1) / kWorkChunk);
}
__host__ __device__ inline uint32_t work_class(uint32_t ta, uint32_t tb) {
    return 4u * (ta == 4u ? 3u : ta) + (tb == 4u ? 3u : tb);
}

// Body metadata word: shape index (bits 0-23), motion type (24-25), sensor
// (26), asleep (27: the body is off the active list, ActivationState::
// Inactive), sleeping allowed (28: BodySettings::allowSleeping).
constexpr uint32_t kMetaAsleep = 1u << 27, kMetaMaySleep = 1u << 28;
__host__ __device__ inline uint32_t meta_shape(uint32_t m) { return m & 0x00FFFFFFu; }
__host__ __device__ inline uint32_t meta_motion(uint32_t m) { return (m >> 24) & 3u; }
__host__ __device__ inline bool meta_sensor(uint32_t m) { return (m >> 26) & 1u; }
__host__ __device__ inline bool meta_asleep(uint32_t m) { return (m & kMetaAsleep) != 0u; }
__host__ __device__ inline bool meta_may_sleep(uint32_t m) { return (m & kMetaMaySleep) != 0u; }

// Maximum manifold points kept per pair.  Every in-scope narrowphase path
// (plane, sphere, GJK/EPA, polyhedral clipping) reduces to at most four
// (src/Collision/CollideShape.cpp:216-224, 406-411).
constexpr int kPairPoints = 4;
// Hull limits: ConvexHullShape::Create caps input at 256 points; triangle
// vertex indices are packed as bytes.
constexpr int kMaxHullVerts = 256;
constexpr int kMaxHullTris = 512;

// Bytes of one GJK result handed from the lane kernel to the wave kernel.
constexpr size_t kGjkRecBytes = 256;

struct ShapeBuf {
    uint32_t* type;   // [S]
    float4* params;   // [S] sphere r | box hx hy hz | capsule hh r | plane nx ny nz d
    uint4* hull;      // [S] vert offset, vert count, tri offset, tri count
    float4* hverts;   // [V] local vertex xyz
    uint32_t* htris;  // [T] a | b << 8 | c << 16
    float4* hfn;      // [T] local face normal NormalizedOr((b-a)x(c-a), +Y)
};

struct BodyBuf {
    float4* x;        // position xyz, inverse mass
    float4* q;        // rotation
    float4* v;        // linear velocity
    float4* w;        // angular velocity
    float4* pv;       // pseudo linear velocity
    float4* pw;       // pseudo angular velocity
    float4* iw;       // inverse world inertia, 3 columns per body
    float4* il;       // inverse local inertia, 3 columns per body
    float4* mat;      // friction, restitution, linear damping, angular damping
    float4* lim;      // gravity factor, max linear velocity, max angular velocity, 0
    uint32_t* meta;   // shape | motion << 24 | sensor << 26 | asleep << 27 | may sleep << 28
    uint32_t* packed; // packed BodyID
};

// Per-pair manifold slots (kPairPoints points each).
struct ManifoldBuf {
    uint32_t* count;  // [pairs]
    float4* pa;       // [pairs*4] point on A xyz, separation
    float4* pb;       // [pairs*4] point on B xyz, feature id (bits)
    float4* n;        // [pairs*4] normal xyz
};

// Contact constraints (compacted in pair order) and their points.
struct ContactBuf {
    uint4* info;      // [C] body A, body B, first point, num points | flags << 8
    uint32_t* pair;   // [C] source pair index
    uint4* body;      // [P] body A, body B, flags (1 sensor, 2 dynA, 4 dynB), constraint
    float4* pa;       // [P] point on A, separation
    float4* pb;       // [P] point on B, feature id bits
    float4* n;        // [P] normal, normal mass
    float4* t1;       // [P] tangent 1, tangent mass 1
    float4* t2;       // [P] tangent 2, tangent mass 2
    float4* ra;       // [P] rA, restitution bias
    float4* rb;       // [P] rB, position bias
    float4* la;       // [P] local anchor A, combined friction
    float4* lb;       // [P] local anchor B, combined restitution
    float4* imp;      // [P] normal, tangent 1, tangent 2 impulse, bias impulse
    float* nvel;      // [P] normal velocity before solving
};

// Warm-start manifold cache: open addressing, key 0 = empty, never evicted
// (the semantics of PhysicsWorld::FindCached / GetOrCreateCached,
// src/World/PhysicsWorld.cpp:405-452).
struct CacheBuf {
    unsigned long long* key;  // [cap]
    uint32_t* used;           // number of occupied slots
    uint32_t* count;          // [cap]
    float4* p;                // [cap*4] position xyz, normal impulse
    float2* t;                // [cap*4] tangent impulses
    uint32_t cap;             // power of two
};

// Level schedule of one solve: per group a range of levels, per level a
// range of sweep records (records are stored in schedule order).
struct Schedule {
    const uint32_t* lvl_off;  // [levels + 1] first unit of each level
    const uint2* grp;         // [groups] first level, end level
    // [units] unit descriptors: body A, body B, flags | records << 8, first
    // record.  A unit's records belong to one body pair and are applied in
    // order by one thread.  nullptr: one record per unit (the descriptor is
    // the record's body entry).
    const uint4* unit;
};

// Sweep records: the constant per-point data of a sweep, laid out in the
// order the sweep visits the points so that a level reads contiguous memory.
// Velocity records (velocity schedule order):
struct VelRecBuf {
    uint4* bd;    // body A, body B, flags (1 sensor, 2 dynA, 4 dynB), point index
    float4* ra;   // rotate(qA, localAnchorA), inverse mass A
    float4* rb;   // rotate(qB, localAnchorB), inverse mass B
    float4* n;    // normal, normal mass
    float4* t1;   // tangent 1, tangent mass 1
    float4* t2;   // tangent 2, tangent mass 2
    float2* fb;   // combined friction, normal bias
    float4* imp;  // normal, tangent 1, tangent 2 impulse
};
// Warm-start / bias records (constraint order within each group):
struct WsRecBuf {
    uint4* bd;    // body A, body B, flags, point index
    float4* ra;   // stored rA (pointOnA - position A), inverse mass A
    float4* rb;   // stored rB, inverse mass B
    float4* qa;   // rotate(qA, localAnchorA), position bias
    float4* qb;   // rotate(qB, localAnchorB), 0
    float4* n;    // normal, normal mass
    float4* t1;   // tangent 1, 0
    float4* t2;   // tangent 2, 0
    float4* imp;  // normal, tangent 1, tangent 2 impulse, bias impulse
};

struct SolveParams {
    float dt;
    float warm_start_factor;
    uint32_t velocity_iterations;
    uint32_t position_iterations;
    uint32_t use_simd;
    uint32_t use_velocity_bias;
    // Which sweeps a launch runs (kPhase* bits): the warm start; the bias
    // sweeps on their own after it; the velocity iterations, with the bias
    // sweeps alongside unless they ran already.  Default: warm start, then
    // velocity with bias alongside.
    uint32_t phases = 3;
};
constexpr uint32_t kPhaseWarm = 1, kPhaseVelocity = 2, kPhaseBiasAlone = 4, kPhaseBiasDone = 8;
// The sweeps run on packed copies of the bodies: whether a sweep launch
// packs them first and unpacks them after (a warm-start-only launch packs,
// the velocity launch that follows it unpacks).
constexpr uint32_t kPackBodies = 1, kUnpackBodies = 2;

__host__ __device__ inline uint64_t pair_key(uint32_t a, uint32_t b) {
    return a < b ? ((uint64_t(a) << 32) | b) : ((uint64_t(b) << 32) | a);
}
__host__ __device__ inline uint32_t cache_hash(uint64_t k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return uint32_t(k);
}

}  // namespace kz
