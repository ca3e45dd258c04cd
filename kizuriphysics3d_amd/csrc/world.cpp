// world.cpp - the host world: bodies, shapes, broadphase, and the per-step
// orchestration of PhysicsWorld::StepFixed (src/World/PhysicsWorld.cpp:245-329)
// with steps 2-6 on the device.  Flat C-ABI in include/kzw.h.
#include <chrono>
#include <cstring>
#include <string>
#include <vector>

#include "host_geometry.h"
#include "kzg.h"
#include "kzw.h"

using namespace kz;

struct kzw_world {
    kz_world_desc settings{};
    int device = -1;
    kzg_ctx* ctx = nullptr;
    std::string err;
    struct ShapeRec {
        kz_shape_desc desc{};
        Hull hull;
        Aabb local;
    };
    std::vector<ShapeRec> shapes;
    std::vector<kz_body_desc> bodies;
    std::vector<kz_body_state> states;
    std::vector<uint32_t> proxy;
    std::vector<float> mass;  // 10 per body
    BroadPhase bp;
    std::vector<uint32_t> pairs;  // slot A, slot B
    bool shapes_dirty = true, bodies_dirty = true;
    kzw_stats stats{};
};

namespace {

int wfail(kzw_world* w, int code, const std::string& m) {
    if (w) w->err = m;
    return code;
}

Aabb world_bounds(const kzw_world& w, uint32_t i) {
    const kz_body_state& s = w.states[i];
    const auto& sh = w.shapes[w.bodies[i].shape];
    return transform_aabb(sh.local, f4q{s.rotation[0], s.rotation[1], s.rotation[2], s.rotation[3]},
                          f3{s.position[0], s.position[1], s.position[2]});
}

int sync_device(kzw_world* w) {
    if (!w->ctx) return wfail(w, KZG_ERR_DEVICE, "host-only world: no device");
    if (w->shapes_dirty) {
        std::vector<kz_shape_desc> sd;
        std::vector<kzg_hull_range> hr;
        std::vector<float> hv;
        std::vector<uint32_t> ht;
        for (const auto& s : w->shapes) {
            sd.push_back(s.desc);
            kzg_hull_range r{0, 0, 0, 0};
            if (s.desc.type == KZ_SHAPE_CONVEX_HULL) {
                r.vert_offset = uint32_t(hv.size() / 3);
                r.vert_count = uint32_t(s.hull.verts.size());
                r.tri_offset = uint32_t(ht.size() / 3);
                r.tri_count = uint32_t(s.hull.tris.size() / 3);
                for (const f3& v : s.hull.verts) { hv.push_back(v.x); hv.push_back(v.y); hv.push_back(v.z); }
                ht.insert(ht.end(), s.hull.tris.begin(), s.hull.tris.end());
            }
            hr.push_back(r);
        }
        int rc = kzg_upload_shapes(w->ctx, sd.data(), uint32_t(sd.size()), hr.data(), hv.data(),
                                   uint32_t(hv.size() / 3), ht.data(), uint32_t(ht.size() / 3));
        if (rc) return wfail(w, rc, kzg_last_error(w->ctx));
        w->shapes_dirty = false;
        w->bodies_dirty = true;
    }
    if (w->bodies_dirty) {
        std::vector<kzg_body> b(w->bodies.size());
        for (size_t i = 0; i < b.size(); ++i) {
            const kz_body_desc& d = w->bodies[i];
            const kz_body_state& s = w->states[i];
            kzg_body& o = b[i];
            o.packed_id = uint32_t(i);
            o.shape = d.shape;
            o.motion_type = d.motion_type;
            o.is_sensor = d.is_sensor;
            std::memcpy(o.position, s.position, sizeof(o.position));
            std::memcpy(o.rotation, s.rotation, sizeof(o.rotation));
            std::memcpy(o.linear_velocity, s.linear_velocity, sizeof(o.linear_velocity));
            std::memcpy(o.angular_velocity, s.angular_velocity, sizeof(o.angular_velocity));
            o.inverse_mass = w->mass[10 * i];
            std::memcpy(o.inverse_inertia_local, &w->mass[10 * i + 1], sizeof(float) * 9);
            o.friction = d.friction;
            o.restitution = d.restitution;
            o.linear_damping = d.linear_damping;
            o.angular_damping = d.angular_damping;
            o.gravity_factor = d.gravity_factor;
            o.max_linear_velocity = d.max_linear_velocity;
            o.max_angular_velocity = d.max_angular_velocity;
        }
        int rc = kzg_upload_bodies(w->ctx, b.data(), uint32_t(b.size()));
        if (rc) return wfail(w, rc, kzg_last_error(w->ctx));
        w->bodies_dirty = false;
    }
    return KZG_OK;
}

}  // namespace

extern "C" {

int kzw_create(kzw_world** out, const kz_world_desc* ws, int device) {
    if (!out || !ws) return KZG_ERR_ARGUMENT;
    *out = nullptr;
    auto* w = new kzw_world();
    w->settings = *ws;
    w->device = device;
    if (ws->allow_sleeping)
        return delete w, KZG_ERR_UNSUPPORTED;  // the device path does not model sleeping
    if (device >= 0) {
        int rc = kzg_create(&w->ctx, ws, device);
        if (rc) {
            delete w;
            return rc;
        }
    }
    *out = w;
    return KZG_OK;
}

void kzw_destroy(kzw_world* w) {
    if (!w) return;
    if (w->ctx) kzg_destroy(w->ctx);
    delete w;
}

const char* kzw_last_error(const kzw_world* w) { return w ? w->err.c_str() : "null world"; }

int kzw_add_shapes(kzw_world* w, const kz_shape_desc* s, uint32_t n, const float* hp, uint32_t nhp) {
    if (!w || (n && !s)) return KZG_ERR_ARGUMENT;
    for (uint32_t i = 0; i < n; ++i) {
        kzw_world::ShapeRec r;
        r.desc = s[i];
        const uint32_t t = s[i].type;
        if (t != KZ_SHAPE_SPHERE && t != KZ_SHAPE_BOX && t != KZ_SHAPE_CAPSULE && t != KZ_SHAPE_CONVEX_HULL &&
            t != KZ_SHAPE_PLANE)
            return wfail(w, KZG_ERR_UNSUPPORTED, "shape type " + std::to_string(t) + " is outside the device path");
        if (t == KZ_SHAPE_CONVEX_HULL) {
            if (!hp || size_t(s[i].hull_offset) + s[i].hull_count > nhp)
                return wfail(w, KZG_ERR_ARGUMENT, "hull points out of range");
            if (!build_hull(hp + 3 * size_t(s[i].hull_offset), s[i].hull_count, r.hull))
                return wfail(w, KZG_ERR_ARGUMENT, "ConvexHullShape::Create rejects the points of shape " +
                                                      std::to_string(w->shapes.size()));
            if (r.hull.verts.size() > 256 || r.hull.tris.size() / 3 > 512)
                return wfail(w, KZG_ERR_UNSUPPORTED, "hull larger than 256 vertices / 512 triangles");
        }
        r.local = shape_local_bounds(t, s[i].params, &r.hull);
        w->shapes.push_back(std::move(r));
    }
    w->shapes_dirty = true;
    return KZG_OK;
}

int kzw_add_bodies(kzw_world* w, const kz_body_desc* b, uint32_t n) {
    if (!w || (n && !b)) return KZG_ERR_ARGUMENT;
    for (uint32_t i = 0; i < n; ++i) {
        const kz_body_desc& d = b[i];
        if (d.shape >= w->shapes.size()) return wfail(w, KZG_ERR_ARGUMENT, "bad shape index");
        if (d.motion_type > 2) return wfail(w, KZG_ERR_ARGUMENT, "bad motion type");
        const uint32_t slot = uint32_t(w->bodies.size());
        w->bodies.push_back(d);
        kz_body_state s;
        std::memcpy(s.position, d.position, sizeof(s.position));
        std::memcpy(s.rotation, d.rotation, sizeof(s.rotation));
        std::memcpy(s.linear_velocity, d.linear_velocity, sizeof(s.linear_velocity));
        std::memcpy(s.angular_velocity, d.angular_velocity, sizeof(s.angular_velocity));
        w->states.push_back(s);
        const auto& sh = w->shapes[d.shape];
        float im;
        m3 il;
        body_mass(d.motion_type, sh.desc.type, sh.desc.params, &sh.hull, d.density, d.mass_override, im, il);
        const float m[10] = {im, il.c0.x, il.c0.y, il.c0.z, il.c1.x, il.c1.y, il.c1.z, il.c2.x, il.c2.y, il.c2.z};
        w->mass.insert(w->mass.end(), m, m + 10);
        // PhysicsWorld::CreateBody: proxy from the world bounds, user data =
        // packed id, layer / mask from BodySettings defaults.
        w->proxy.push_back(w->bp.create_proxy(world_bounds(*w, slot), slot, 1u, 0xFFFFFFFFu));
    }
    w->bodies_dirty = true;
    return KZG_OK;
}

int kzw_update_broadphase(kzw_world* w) {
    if (!w) return KZG_ERR_ARGUMENT;
    // UpdateBroadPhase (PhysicsWorld.cpp:334-345): active non-static bodies,
    // displacement = v * fixedTimestep.
    for (uint32_t i = 0; i < w->bodies.size(); ++i) {
        if (w->bodies[i].motion_type == KZ_MOTION_STATIC) continue;
        const kz_body_state& s = w->states[i];
        const f3 disp = mul(f3{s.linear_velocity[0], s.linear_velocity[1], s.linear_velocity[2]},
                            w->settings.fixed_timestep);
        w->bp.move_proxy(w->proxy[i], world_bounds(*w, i), disp);
    }
    w->bp.update_pairs();
    const auto& pp = w->bp.pairs();
    w->pairs.resize(pp.size());
    for (size_t k = 0; k < pp.size(); ++k) w->pairs[k] = w->bp.user(pp[k]);
    return KZG_OK;
}

int kzw_step(kzw_world* w, float dt) {
    if (!w) return KZG_ERR_ARGUMENT;
    using clock = std::chrono::steady_clock;
    const auto t0 = clock::now();
    int rc = sync_device(w);
    if (rc) return rc;
    kzw_update_broadphase(w);
    const auto t1 = clock::now();
    const uint32_t np = uint32_t(w->pairs.size() / 2);
    if ((rc = kzg_narrowphase(w->ctx, w->pairs.data(), np)) || (rc = kzg_build_constraints(w->ctx, 0)) ||
        (rc = kzg_integrate_velocities(w->ctx, dt)) || (rc = kzg_solve(w->ctx, dt)))
        return wfail(w, rc, kzg_last_error(w->ctx));
    kzg_timings tm;
    if ((rc = kzg_get_timings(w->ctx, &tm))) return wfail(w, rc, kzg_last_error(w->ctx));  // synchronises
    const auto t2 = clock::now();
    if ((rc = kzg_integrate_positions(w->ctx, dt)) ||
        (rc = kzg_download_states(w->ctx, w->states.data(), uint32_t(w->states.size()))))
        return wfail(w, rc, kzg_last_error(w->ctx));
    const auto t3 = clock::now();
    uint32_t npairs = 0, nc = 0, npts = 0;
    kzg_query_sizes(w->ctx, &npairs, &nc, &npts);
    w->stats.num_bodies = uint32_t(w->bodies.size());
    w->stats.num_broadphase_pairs = np;
    w->stats.num_contact_constraints = nc;
    w->stats.num_contact_points = npts;
    w->stats.broadphase_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    w->stats.contact_path_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
    w->stats.step_ms = std::chrono::duration<double, std::milli>(t3 - t0).count();
    return KZG_OK;
}

int kzw_get_states(kzw_world* w, kz_body_state* out, uint32_t n) {
    if (!w || !out || n != w->states.size()) return KZG_ERR_ARGUMENT;
    std::memcpy(out, w->states.data(), sizeof(kz_body_state) * n);
    return KZG_OK;
}

int kzw_set_states(kzw_world* w, const kz_body_state* in, uint32_t n) {
    if (!w || !in || n != w->states.size()) return KZG_ERR_ARGUMENT;
    std::memcpy(w->states.data(), in, sizeof(kz_body_state) * n);
    if (w->ctx && !w->bodies_dirty && !w->shapes_dirty) {
        int rc = kzg_upload_states(w->ctx, in, n);
        if (rc) return wfail(w, rc, kzg_last_error(w->ctx));
    }
    return KZG_OK;
}

int kzw_get_pairs(kzw_world* w, uint32_t* out, uint32_t cap) {
    if (!w) return KZG_ERR_ARGUMENT;
    const uint32_t n = uint32_t(w->pairs.size() / 2);
    if (out) std::memcpy(out, w->pairs.data(), sizeof(uint32_t) * 2 * std::min(n, cap));
    return int(n);
}

int kzw_get_stats(kzw_world* w, kzw_stats* out) {
    if (!w || !out) return KZG_ERR_ARGUMENT;
    *out = w->stats;
    return KZG_OK;
}

int kzw_get_mass(kzw_world* w, float* out, uint32_t n) {
    if (!w || !out || n != w->bodies.size()) return KZG_ERR_ARGUMENT;
    std::memcpy(out, w->mass.data(), sizeof(float) * 10 * n);
    return KZG_OK;
}

int kzw_get_hull(kzw_world* w, uint32_t shape, float* verts, uint32_t* tris, uint32_t* nv, uint32_t* nt) {
    if (!w || shape >= w->shapes.size()) return KZG_ERR_ARGUMENT;
    const Hull& h = w->shapes[shape].hull;
    if (nv) *nv = uint32_t(h.verts.size());
    if (nt) *nt = uint32_t(h.tris.size() / 3);
    if (verts)
        for (size_t i = 0; i < h.verts.size(); ++i) {
            verts[3 * i] = h.verts[i].x;
            verts[3 * i + 1] = h.verts[i].y;
            verts[3 * i + 2] = h.verts[i].z;
        }
    if (tris) std::memcpy(tris, h.tris.data(), sizeof(uint32_t) * h.tris.size());
    return KZG_OK;
}

kzg_ctx* kzw_device(kzw_world* w) { return w ? w->ctx : nullptr; }

}  // extern "C"
