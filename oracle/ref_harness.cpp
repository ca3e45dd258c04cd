// ref_harness.cpp - C-ABI harness around the REFERENCE engine, built from the
// reference's own sources where they lie (oracle/Makefile, target _ref).
//
// TEST INFRASTRUCTURE ONLY.  This file and the library it builds
// (oracle/_ref/libkzref.so) are the checker and the reference CPU baseline:
// they are loaded by tests/, by the fixture generator under tests/golden/ and
// by bench.py's cpu_baseline leg, never by the product library.
//
// The harness creates a kizuri::PhysicsWorld from the plain-data scene of
// include/kz_scene.h and exposes stepping, per-phase timing and dumps of the
// pair list, raw manifolds, contact constraints and body state.  To reach the
// per-phase methods of PhysicsWorld::StepFixed
// (src/World/PhysicsWorld.cpp:245-329) the reference headers are included with
// `private` redefined to `public`; no reference source is modified or copied.
#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <limits>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <type_traits>
#include <utility>
#include <unordered_map>
#include <vector>

#define private public
#include "Kizuri/Kizuri.h"
#undef private

#include "kz_scene.h"

using namespace kizuri;

namespace {

using Clock = std::chrono::steady_clock;
double MsSince(Clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
}

ShapeRef MakeShape(const kz_shape_desc& d, const float* hullPool) {
    switch (d.type) {
        case KZ_SHAPE_SPHERE: return MakeRef<SphereShape>(Real(d.params[0]));
        case KZ_SHAPE_BOX: return MakeRef<BoxShape>(Vec3(d.params[0], d.params[1], d.params[2]));
        case KZ_SHAPE_CAPSULE: return MakeRef<CapsuleShape>(Real(d.params[0]), Real(d.params[1]));
        case KZ_SHAPE_PLANE:
            return MakeRef<PlaneShape>(Plane(Vec3(d.params[0], d.params[1], d.params[2]), Real(d.params[3])));
        case KZ_SHAPE_CONVEX_HULL: {
            std::vector<Vec3> pts(d.hull_count);
            for (uint32_t i = 0; i < d.hull_count; ++i) {
                const float* p = hullPool + 3u * (d.hull_offset + i);
                pts[i] = Vec3(p[0], p[1], p[2]);
            }
            Ref<ConvexHullShape> h = ConvexHullShape::Create(ArrayView<const Vec3>(pts.data(), pts.size()));
            return ShapeRef(h.Get());
        }
        default: return ShapeRef();
    }
}

struct RefWorld {
    PhysicsWorld world;
    std::vector<ShapeRef> shapes;
    std::vector<BodyID> ids;
};

void FillPoint(float* o, const ContactPointConstraint& p) {
    auto v3 = [&](int at, const Vec3& v) { o[at] = v.x; o[at + 1] = v.y; o[at + 2] = v.z; };
    v3(0, p.pointOnA); v3(3, p.pointOnB); v3(6, p.rA); v3(9, p.rB);
    v3(12, p.localAnchorA); v3(15, p.localAnchorB);
    v3(18, p.normal); v3(21, p.tangent1); v3(24, p.tangent2);
    o[27] = p.normalMass; o[28] = p.tangentMass1; o[29] = p.tangentMass2;
    o[30] = p.normalImpulse; o[31] = p.tangentImpulse1; o[32] = p.tangentImpulse2;
    o[33] = p.separation; o[34] = p.restitutionBias; o[35] = p.positionBias;
    o[36] = p.biasImpulse; o[37] = p.combinedFriction; o[38] = p.combinedRestitution;
    o[39] = p.normalVelocity;
    std::memcpy(&o[40], &p.featureId, 4);
    o[41] = o[42] = o[43] = 0.0f;
}

void FillContact(kz_contact_point& o, const ContactPoint& p) {
    o.point_on_a[0] = p.pointOnA.x; o.point_on_a[1] = p.pointOnA.y; o.point_on_a[2] = p.pointOnA.z;
    o.point_on_b[0] = p.pointOnB.x; o.point_on_b[1] = p.pointOnB.y; o.point_on_b[2] = p.pointOnB.z;
    o.normal[0] = p.normal.x; o.normal[1] = p.normal.y; o.normal[2] = p.normal.z;
    o.separation = p.separation;
    o.feature_id = p.featureId;
}

// Restatement of the inline block of StepFixed that integrates the split-
// impulse pseudo velocities (src/World/PhysicsWorld.cpp:275-289); the
// rotation update is the file-local IntegrateRotation (:31-35).
void IntegratePseudo(PhysicsWorld& w, Real dt) {
    const auto& active = w.mBodies.GetActiveBodyList();
    for (u32 i = 0; i < active.Size(); ++i) {
        Body* body = w.mBodies.GetBody(active[i]);
        if (!body || body->IsStatic()) continue;
        MotionProperties& mp = body->GetMotionProperties();
        if (mp.pseudoLinearVelocity.LengthSq() > Real(0))
            body->SetPosition(body->GetPosition() + mp.pseudoLinearVelocity * dt);
        if (mp.pseudoAngularVelocity.LengthSq() > Real(0)) {
            const Vec3& av = mp.pseudoAngularVelocity;
            Quat q = body->GetRotation();
            Quat dq(Real(0.5) * dt * av.x, Real(0.5) * dt * av.y, Real(0.5) * dt * av.z, Real(0));
            body->SetRotation((q + dq * q).Normalized());
        }
        mp.pseudoLinearVelocity = Vec3::Zero();
        mp.pseudoAngularVelocity = Vec3::Zero();
    }
}

} // namespace

extern "C" {

// Start the reference JobSystem with `workers` worker threads (0 = hardware
// concurrency - 1, the reference default).  Returns the worker count.
uint32_t kzr_init_jobs(uint32_t workers) {
    if (!JobSystem::IsInitialized()) JobSystem::Init(workers);
    return JobSystem::Get().NumThreads();
}

void* kzr_world_create(const kz_world_desc* wd, const kz_shape_desc* shapes, uint32_t numShapes,
                       const float* hullPool, const kz_body_desc* bodies, uint32_t numBodies) {
    auto* rw = new RefWorld();
    WorldSettings& ws = rw->world.GetSettings();
    ws.gravity = Vec3(wd->gravity[0], wd->gravity[1], wd->gravity[2]);
    ws.solver.velocityIterations = wd->velocity_iterations;
    ws.solver.positionIterations = wd->position_iterations;
    ws.solver.baumgarte = wd->baumgarte;
    ws.solver.penetrationSlop = wd->penetration_slop;
    ws.solver.restitutionThreshold = wd->restitution_threshold;
    ws.solver.maxBiasVelocity = wd->max_bias_velocity;
    ws.solver.warmStartFactor = wd->warm_start_factor;
    ws.solver.useSIMDSolver = wd->use_simd_solver != 0;
    ws.solver.useVelocityBias = wd->use_velocity_bias != 0;
    ws.allowSleeping = wd->allow_sleeping != 0;
    ws.useMultithreading = wd->use_multithreading != 0;
    ws.fixedTimestep = wd->fixed_timestep;
    for (uint32_t i = 0; i < numShapes; ++i) {
        ShapeRef s = MakeShape(shapes[i], hullPool);
        if (!s) { delete rw; return nullptr; }
        rw->shapes.push_back(s);
    }
    for (uint32_t i = 0; i < numBodies; ++i) {
        const kz_body_desc& d = bodies[i];
        BodySettings s;
        s.shape = rw->shapes.at(d.shape);
        s.motionType = MotionType(d.motion_type);
        s.isSensor = d.is_sensor != 0;
        s.allowSleeping = d.allow_sleeping != 0;
        s.position = Vec3(d.position[0], d.position[1], d.position[2]);
        s.rotation = Quat(d.rotation[0], d.rotation[1], d.rotation[2], d.rotation[3]);
        s.linearVelocity = Vec3(d.linear_velocity[0], d.linear_velocity[1], d.linear_velocity[2]);
        s.angularVelocity = Vec3(d.angular_velocity[0], d.angular_velocity[1], d.angular_velocity[2]);
        s.friction = d.friction;
        s.restitution = d.restitution;
        s.linearDamping = d.linear_damping;
        s.angularDamping = d.angular_damping;
        s.gravityFactor = d.gravity_factor;
        s.maxLinearVelocity = d.max_linear_velocity;
        s.maxAngularVelocity = d.max_angular_velocity;
        s.massOverride = d.mass_override;
        s.density = d.density;
        BodyID id = rw->world.CreateBody(s);
        if (id.IsInvalid()) { delete rw; return nullptr; }
        rw->ids.push_back(id);
    }
    return rw;
}

void kzr_world_destroy(void* w) { delete static_cast<RefWorld*>(w); }

// SerializeWorld (src/Serialization/Serializer.cpp:182-205) into out (cap
// bytes); returns the record size, 0 on failure (out may be null to query).
uint64_t kzr_save_kzp1(void* wp, uint8_t* out, uint64_t cap) {
    Vector<u8> buf;
    if (!SerializeWorld(static_cast<RefWorld*>(wp)->world, buf)) return 0;
    if (out && cap >= buf.Size()) std::memcpy(out, buf.Data(), buf.Size());
    return uint64_t(buf.Size());
}

// A fresh world with settings wd filled by DeserializeWorld
// (src/Serialization/Serializer.cpp:207-234).
void* kzr_load_kzp1(const kz_world_desc* wd, const uint8_t* data, uint64_t size) {
    auto* rw = new RefWorld();
    WorldSettings& ws = rw->world.GetSettings();
    ws.gravity = Vec3(wd->gravity[0], wd->gravity[1], wd->gravity[2]);
    ws.solver.velocityIterations = wd->velocity_iterations;
    ws.solver.positionIterations = wd->position_iterations;
    ws.solver.baumgarte = wd->baumgarte;
    ws.solver.penetrationSlop = wd->penetration_slop;
    ws.solver.restitutionThreshold = wd->restitution_threshold;
    ws.solver.maxBiasVelocity = wd->max_bias_velocity;
    ws.solver.warmStartFactor = wd->warm_start_factor;
    ws.solver.useSIMDSolver = wd->use_simd_solver != 0;
    ws.solver.useVelocityBias = wd->use_velocity_bias != 0;
    ws.allowSleeping = wd->allow_sleeping != 0;
    ws.useMultithreading = wd->use_multithreading != 0;
    ws.fixedTimestep = wd->fixed_timestep;
    if (!DeserializeWorld(rw->world, data, size_t(size))) {
        delete rw;
        return nullptr;
    }
    const BodyManager& bm = rw->world.GetBodyManager();
    for (u32 i = 0; i < bm.GetBodySlotCount(); ++i) {
        const Body* b = bm.GetBodyBySlot(i);
        if (b) rw->ids.push_back(b->GetID());
    }
    return rw;
}

// One PhysicsWorld::StepFixed.
void kzr_step(void* w, float dt) { static_cast<RefWorld*>(w)->world.StepFixed(dt); }

// StepFixed's sequence, phase by phase, with a timer around each phase:
// ms[0] broadphase, [1] Collide, [2] BuildContactConstraints,
// [3] IntegrateVelocities, [4] SolveContacts, [5] the rest of the step.
// `stop_after` (0..5) ends the step early after that phase so that tests can
// read intermediate state; 5 runs the whole step.  Bit-identical to StepFixed
// (tests/test_reference_harness.py checks it).
void kzr_step_phases(void* wp, float dt, double* ms, int stop_after) {
    PhysicsWorld& w = static_cast<RefWorld*>(wp)->world;
    double local[6] = {0, 0, 0, 0, 0, 0};
    auto t = Clock::now();
    ++w.mFrame;
    if (w.mFrame == 0) w.mFrame = 1;
    w.UpdateBroadPhase();
    w.mBroadPhase.UpdatePairs();
    local[0] = MsSince(t);
    if (stop_after >= 1) { t = Clock::now(); w.Collide(); local[1] = MsSince(t); }
    if (stop_after >= 2) { t = Clock::now(); w.BuildContactConstraints(); local[2] = MsSince(t); }
    if (stop_after >= 3) { t = Clock::now(); w.IntegrateVelocities(dt); local[3] = MsSince(t); }
    if (stop_after >= 4) { t = Clock::now(); w.SolveContacts(dt); local[4] = MsSince(t); }
    if (stop_after >= 5) {
        t = Clock::now();
        w.IntegratePositions(dt);
        if (!w.mSettings.solver.useVelocityBias) IntegratePseudo(w, dt);
        w.UpdateBodyInertias();
        for (u32 i = 0; i < u32(w.mIslands.Size()); ++i) w.SolveIslandPositions(i, dt);
        w.UpdateSleeping(dt);
        w.mBodies.ForEachBody([](Body& b) { b.ClearForces(); });
        local[5] = MsSince(t);
    }
    if (ms) for (int i = 0; i < 6; ++i) ms[i] = local[i];
}

uint32_t kzr_num_bodies(void* w) { return uint32_t(static_cast<RefWorld*>(w)->ids.size()); }

void kzr_get_states(void* wp, kz_body_state* out) {
    RefWorld* rw = static_cast<RefWorld*>(wp);
    for (size_t i = 0; i < rw->ids.size(); ++i) {
        const Body* b = rw->world.GetBody(rw->ids[i]);
        kz_body_state& s = out[i];
        const Vec3& p = b->GetPosition(); const Quat& q = b->GetRotation();
        const Vec3& v = b->GetLinearVelocity(); const Vec3& a = b->GetAngularVelocity();
        s.position[0] = p.x; s.position[1] = p.y; s.position[2] = p.z;
        s.rotation[0] = q.x; s.rotation[1] = q.y; s.rotation[2] = q.z; s.rotation[3] = q.w;
        s.linear_velocity[0] = v.x; s.linear_velocity[1] = v.y; s.linear_velocity[2] = v.z;
        s.angular_velocity[0] = a.x; s.angular_velocity[1] = a.y; s.angular_velocity[2] = a.z;
    }
}

// Overwrite position, rotation and velocities of every body (a snapshot
// restart: the warm-start cache is left as it is).
void kzr_set_states(void* wp, const kz_body_state* in) {
    RefWorld* rw = static_cast<RefWorld*>(wp);
    for (size_t i = 0; i < rw->ids.size(); ++i) {
        Body* b = rw->world.GetBody(rw->ids[i]);
        const kz_body_state& s = in[i];
        b->SetPosition(Vec3(s.position[0], s.position[1], s.position[2]));
        b->SetRotation(Quat(s.rotation[0], s.rotation[1], s.rotation[2], s.rotation[3]));
        b->GetMotionProperties().linearVelocity = Vec3(s.linear_velocity[0], s.linear_velocity[1], s.linear_velocity[2]);
        b->GetMotionProperties().angularVelocity = Vec3(s.angular_velocity[0], s.angular_velocity[1], s.angular_velocity[2]);
        if (b->IsDynamic()) b->GetMotionProperties().UpdateWorldInertia(b->GetRotation());
    }
}

// Per body: inverse mass, inverse local inertia (9, column-major), inverse
// world inertia (9), 19 floats.
void kzr_get_mass(void* wp, float* out) {
    RefWorld* rw = static_cast<RefWorld*>(wp);
    for (size_t i = 0; i < rw->ids.size(); ++i) {
        const MotionProperties& mp = rw->world.GetBody(rw->ids[i])->GetMotionProperties();
        float* o = out + 19 * i;
        o[0] = mp.inverseMass;
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 3; ++r) {
                o[1 + 3 * c + r] = mp.inverseInertiaLocal.col[c][r];
                o[10 + 3 * c + r] = mp.inverseInertiaWorld.col[c][r];
            }
    }
}

uint32_t kzr_num_pairs(void* wp) {
    return uint32_t(static_cast<RefWorld*>(wp)->world.mBroadPhase.GetPairs().Size());
}

// Pairs of the last broadphase update as (packed id A, packed id B), in the
// order the narrowphase consumes them.
void kzr_get_pairs(void* wp, uint32_t* out) {
    PhysicsWorld& w = static_cast<RefWorld*>(wp)->world;
    auto pairs = w.mBroadPhase.GetPairs();
    for (u32 i = 0; i < pairs.Size(); ++i) {
        out[2 * i] = w.mBroadPhase.GetUserData(pairs[i].proxyA);
        out[2 * i + 1] = w.mBroadPhase.GetUserData(pairs[i].proxyB);
    }
}

// Raw per-pair manifolds of the last Collide(): counts[i], 8 points per pair,
// keys[i] (MakePairKey or 0).
void kzr_get_manifolds(void* wp, uint32_t* counts, kz_contact_point* points, uint64_t* keys) {
    PhysicsWorld& w = static_cast<RefWorld*>(wp)->world;
    for (u32 i = 0; i < w.mManifolds.Size(); ++i) {
        const Manifold& m = w.mManifolds[i];
        counts[i] = m.numPoints;
        keys[i] = w.mManifoldKeys[i];
        for (u32 k = 0; k < kMaxManifoldPoints; ++k) FillContact(points[8 * i + k], m.points[k]);
    }
}

uint32_t kzr_num_constraints(void* wp) {
    return uint32_t(static_cast<RefWorld*>(wp)->world.GetContactConstraints().Size());
}

// Contact constraints: hdr[4*c] = packed A, packed B, numPoints,
// flags (1 sensor, 2 dynamicA, 4 dynamicB); pts[(8*c+k)*44 ...] = point k.
void kzr_get_constraints(void* wp, uint32_t* hdr, float* pts) {
    PhysicsWorld& w = static_cast<RefWorld*>(wp)->world;
    auto cs = w.GetContactConstraints();
    for (u32 c = 0; c < cs.Size(); ++c) {
        const ContactConstraint& cc = cs[c];
        hdr[4 * c + 0] = cc.bodyA ? cc.bodyA->GetID().Packed() : 0xFFFFFFFFu;
        hdr[4 * c + 1] = cc.bodyB ? cc.bodyB->GetID().Packed() : 0xFFFFFFFFu;
        hdr[4 * c + 2] = cc.numPoints;
        hdr[4 * c + 3] = (cc.isSensor ? 1u : 0u) | (cc.dynamicA ? 2u : 0u) | (cc.dynamicB ? 4u : 0u);
        for (u32 k = 0; k < kMaxManifoldPoints; ++k) FillPoint(pts + (8 * c + k) * 44, cc.points[k]);
    }
}

// Island partition of the last solve: island count, then for each contact
// constraint its island (0xFFFFFFFF when unassigned).
uint32_t kzr_get_islands(void* wp, uint32_t* islandOfConstraint) {
    PhysicsWorld& w = static_cast<RefWorld*>(wp)->world;
    u32 n = u32(w.GetContactConstraints().Size());
    for (u32 c = 0; c < n; ++c) islandOfConstraint[c] = 0xFFFFFFFFu;
    for (u32 i = 0; i < w.mIslands.Size(); ++i)
        for (u32 c : w.mIslands[i].contacts) islandOfConstraint[c] = i;
    return u32(w.mIslands.Size());
}

// The active body list in order, as body slots (BodyManager::GetActiveBodyList);
// returns its length.
uint32_t kzr_get_active(void* wp, uint32_t* out) {
    RefWorld* rw = static_cast<RefWorld*>(wp);
    std::unordered_map<u32, uint32_t> slot;
    for (size_t i = 0; i < rw->ids.size(); ++i) slot[rw->ids[i].Packed()] = uint32_t(i);
    const auto& act = rw->world.mBodies.GetActiveBodyList();
    for (u32 i = 0; i < act.Size(); ++i) out[i] = slot.at(act[i].Packed());
    return u32(act.Size());
}

uint64_t kzr_state_hash(void* wp) { return static_cast<RefWorld*>(wp)->world.ComputeStateHash(); }

// Build a hull with ConvexHullShape::Create.  Returns 0 on success, -1 when
// Create rejects the input.  Capacities: 256 points, 512 triangles.
int kzr_hull_create(const float* pts, uint32_t n, float* outPts, uint32_t* outNumPts,
                    uint32_t* outFaces, uint32_t* outNumFaces) {
    std::vector<Vec3> v(n);
    for (uint32_t i = 0; i < n; ++i) v[i] = Vec3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
    Ref<ConvexHullShape> h = ConvexHullShape::Create(ArrayView<const Vec3>(v.data(), v.size()));
    if (!h) return -1;
    auto hp = h->GetPoints();
    auto hi = h->GetFaceIndices();
    if (hp.Size() > 256 || h->GetNumFaces() > 512) return -2;
    for (u32 i = 0; i < hp.Size(); ++i) { outPts[3 * i] = hp[i].x; outPts[3 * i + 1] = hp[i].y; outPts[3 * i + 2] = hp[i].z; }
    for (u32 i = 0; i < hi.Size(); ++i) outFaces[i] = hi[i];
    *outNumPts = u32(hp.Size());
    *outNumFaces = h->GetNumFaces();
    return 0;
}

// NarrowPhase::Collide over a batch of pairs of shapes from a shape table.
// xf = 7 floats per transform (quat x y z w, translation x y z).  When
// `parallel` is set the batch runs through JobSystem::ParallelFor with the
// world's batch size of 32 (src/World/PhysicsWorld.cpp:395-399).
// Returns the elapsed wall time in milliseconds, or -1 on a bad shape.
double kzr_collide_batch(const kz_shape_desc* shapes, uint32_t numShapes, const float* hullPool,
                         const uint32_t* shapeA, const uint32_t* shapeB,
                         const float* xfA, const float* xfB, uint32_t n, float maxSeparation,
                         uint32_t* outCounts, kz_contact_point* outPts, int parallel) {
    std::vector<ShapeRef> table;
    for (uint32_t i = 0; i < numShapes; ++i) {
        table.push_back(MakeShape(shapes[i], hullPool));
        if (!table.back()) return -1.0;
    }
    struct Ctx {
        std::vector<ShapeRef>* table; const uint32_t* sa; const uint32_t* sb;
        const float* xa; const float* xb; float maxSep; uint32_t* counts; kz_contact_point* pts;
    } ctx{ &table, shapeA, shapeB, xfA, xfB, maxSeparation, outCounts, outPts };
    auto one = [](void* data, u32 i) {
        Ctx* c = static_cast<Ctx*>(data);
        const float* a = c->xa + 7 * i; const float* b = c->xb + 7 * i;
        Transform ta(Quat(a[0], a[1], a[2], a[3]), Vec3(a[4], a[5], a[6]));
        Transform tb(Quat(b[0], b[1], b[2], b[3]), Vec3(b[4], b[5], b[6]));
        CollideSettings cs;
        cs.maxSeparation = c->maxSep;
        Manifold m;
        NarrowPhase::Collide(*(*c->table)[c->sa[i]], ta, *(*c->table)[c->sb[i]], tb, cs, m);
        c->counts[i] = m.numPoints;
        for (u32 k = 0; k < kMaxManifoldPoints; ++k) FillContact(c->pts[8 * size_t(i) + k], m.points[k]);
    };
    auto t0 = Clock::now();
    if (parallel && JobSystem::IsInitialized() && n > 64) JobSystem::Get().ParallelFor(0, n, 32, one, &ctx);
    else for (uint32_t i = 0; i < n; ++i) one(&ctx, i);
    return MsSince(t0);
}

} // extern "C"
