// world.cpp - the host world: bodies, shapes, broadphase, and the per-step
// orchestration of PhysicsWorld::StepFixed (src/World/PhysicsWorld.cpp:245-329)
// with steps 2-6 on the device.  Flat C-ABI in include/kzw.h.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
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
    // The pair query on the device (kzg_broadphase_pairs): the pairs then
    // stay on the device, and `pairs` is filled only when asked for.
    bool device_bp = true, pairs_on_device = false;
    bool host_moves = false;  // kzw_set_device_broadphase mode 2
    uint32_t reinserted = 0;
    uint32_t device_pairs = 0;
    std::vector<float> bp_image;
    bool shapes_dirty = true, bodies_dirty = true;
    kzw_stats stats{};
    std::vector<uint32_t> active;  // the context's active list (BodyManager's), read each step
    std::vector<uint8_t> moved;    // per active-list entry: left its fat box
    // The device keeps the tree image between steps (kzg_broadphase_move /
    // _update): current while dev_image; leaf tight boxes then live there
    // and the host's are stale (tight_stale) until synchronised.
    bool dev_image = false, dev_proxies = false, tight_stale = false;
    uint32_t dev_nodes = 0;
    std::vector<uint32_t> moved_rec, upd_ids;
    std::vector<float> upd_rec;
    std::unique_ptr<HostPool> pool;  // host threads for the per-body loops of the broadphase update
    int host_threads = 1;            // kzw_set_host_threads
    // Bodies the last loaded snapshot marks inactive: the list the body
    // upload sets is then every other body in slot order.
    std::vector<uint8_t> load_inactive;
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
            o.allow_sleeping = d.allow_sleeping;
            o.angular_damping = d.angular_damping;
            o.gravity_factor = d.gravity_factor;
            o.max_linear_velocity = d.max_linear_velocity;
            o.max_angular_velocity = d.max_angular_velocity;
        }
        int rc = kzg_upload_bodies(w->ctx, b.data(), uint32_t(b.size()));
        if (rc) return wfail(w, rc, kzg_last_error(w->ctx));
        w->dev_proxies = false;
        if (!w->load_inactive.empty()) {
            // Serializer.cpp:230-234: CreateBody pushes each body, then an
            // inactive one is removed again at once (it is the list's tail).
            std::vector<uint32_t> list;
            for (uint32_t i = 0; i < uint32_t(b.size()); ++i)
                if (i >= w->load_inactive.size() || !w->load_inactive[i]) list.push_back(i);
            rc = kzg_set_active_bodies(w->ctx, list.data(), uint32_t(list.size()));
            if (rc) return wfail(w, rc, kzg_last_error(w->ctx));
            w->load_inactive.clear();
        }
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

namespace {
HostPool& pool(kzw_world* w) {
    if (!w->pool) w->pool = std::make_unique<HostPool>(unsigned(std::max(w->host_threads, 1)) - 1u);
    return *w->pool;
}

// UpdateBroadPhase's proxy moves (PhysicsWorld.cpp:334-345): the active
// list's non-static bodies in list order, displacement = v * fixedTimestep.
// A move that stays inside the leaf's fat box only stores the tight box and
// touches nothing another move reads, so those run on host threads; the
// leaves that left their fat boxes are then re-inserted one after another
// in list order, which is the order the tree sees in the reference.
int move_proxies(kzw_world* w) {
    const uint32_t nb = uint32_t(w->bodies.size());
    w->active.resize(nb);
    uint32_t na = nb;
    if (w->ctx) {
        int rc = kzg_get_active_bodies(w->ctx, w->active.data(), nb, &na);
        if (rc) return wfail(w, rc, kzg_last_error(w->ctx));
    } else {
        for (uint32_t i = 0; i < nb; ++i) w->active[i] = i;
    }
    w->moved.resize(na);
    pool(w).run(na, [w](uint32_t b, uint32_t e) {
        for (uint32_t k = b; k < e; ++k) {
            const uint32_t i = w->active[k];
            if (w->bodies[i].motion_type == KZ_MOTION_STATIC) {
                w->moved[k] = 0;
                continue;
            }
            w->moved[k] = w->bp.set_tight(w->proxy[i], world_bounds(*w, i)) ? 0 : 1;
        }
    });
    for (uint32_t k = 0; k < na; ++k) {
        if (!w->moved[k]) continue;
        ++w->reinserted;
        const uint32_t i = w->active[k];
        const kz_body_state& s = w->states[i];
        const f3 disp = mul(f3{s.linear_velocity[0], s.linear_velocity[1], s.linear_velocity[2]},
                            w->settings.fixed_timestep);
        w->bp.reinsert(w->proxy[i], disp);
    }
    return KZG_OK;
}
// The host's leaf tight boxes from the device's image (after device moves).
int sync_tight(kzw_world* w) {
    if (!w->tight_stale) return KZG_OK;
    if (!w->ctx || w->dev_nodes == 0) return wfail(w, KZG_ERR_STATE, "no device tree image to read tight boxes from");
    w->bp_image.resize(16 * size_t(w->dev_nodes));
    int rc = kzg_broadphase_download(w->ctx, reinterpret_cast<kzg_bp_node*>(w->bp_image.data()), w->dev_nodes);
    if (rc) return wfail(w, rc, kzg_last_error(w->ctx));
    for (uint32_t i = 0; i < w->bodies.size(); ++i) {
        const float* r = w->bp_image.data() + 16 * size_t(w->proxy[i]);
        w->bp.set_tight(w->proxy[i], Aabb{f3{r[8], r[9], r[10]}, f3{r[12], r[13], r[14]}});
    }
    w->tight_stale = false;
    return KZG_OK;
}

// The proxy moves on the device (kzg_broadphase_move): the leaves that left
// their fat boxes come back with their tight boxes and are re-inserted here
// in active-list order; the nodes that changed go up with the query.
int device_moves(kzw_world* w) {
    const uint32_t nb = uint32_t(w->bodies.size());
    w->active.resize(nb);
    uint32_t na = nb, m = 0;
    int rc = kzg_get_active_bodies(w->ctx, w->active.data(), nb, &na);
    if (rc) return wfail(w, rc, kzg_last_error(w->ctx));
    w->moved_rec.resize(7 * size_t(std::max<uint32_t>(na, 1)));
    rc = kzg_broadphase_move(w->ctx, reinterpret_cast<float*>(w->moved_rec.data()), na, &m);
    if (rc) return wfail(w, rc, kzg_last_error(w->ctx));
    if (m > na) return wfail(w, KZG_ERR_DEVICE, "more moved proxies than active bodies");
    // MoveProxy in list order for the leaves that left their fat boxes.
    std::vector<uint32_t> order(m);
    for (uint32_t k = 0; k < m; ++k) order[k] = k;
    std::sort(order.begin(), order.end(),
              [&](uint32_t a, uint32_t b) { return w->moved_rec[7 * size_t(a)] < w->moved_rec[7 * size_t(b)]; });
    w->bp.clear_dirty();
    for (uint32_t k : order) {
        const uint32_t* r = w->moved_rec.data() + 7 * size_t(k);
        if (r[0] >= na) return wfail(w, KZG_ERR_DEVICE, "moved proxy outside the active list");
        const uint32_t i = w->active[r[0]];
        float t[6];
        std::memcpy(t, r + 1, sizeof(t));
        w->bp.set_tight(w->proxy[i], Aabb{f3{t[0], t[1], t[2]}, f3{t[3], t[4], t[5]}});
        const kz_body_state& s = w->states[i];
        w->bp.reinsert(w->proxy[i], mul(f3{s.linear_velocity[0], s.linear_velocity[1], s.linear_velocity[2]},
                                        w->settings.fixed_timestep));
    }
    w->reinserted = m;
    w->tight_stale = true;
    return KZG_OK;
}

// The pair query over the device's image with the changed nodes replaced;
// a tree that grew goes up whole (its leaves' tight boxes read back first).
int device_pairs_update(kzw_world* w) {
    const uint32_t nn = w->bp.num_nodes();
    if (nn != w->dev_nodes) {
        int rc = sync_tight(w);
        if (rc) return rc;
        w->dev_image = false;
        return KZG_OK;
    }
    const auto& dirty = w->bp.dirty();
    const uint32_t n = uint32_t(dirty.size());
    w->upd_ids.assign(dirty.begin(), dirty.end());
    w->upd_rec.resize(16 * size_t(std::max<uint32_t>(n, 1)));
    for (uint32_t k = 0; k < n; ++k) w->bp.image(w->upd_rec.data() + 16 * size_t(k), dirty[k], dirty[k] + 1);
    uint32_t np = 0;
    int rc = kzg_broadphase_pairs_update(w->ctx, w->upd_ids.data(), reinterpret_cast<const kzg_bp_node*>(w->upd_rec.data()),
                                         n, nn, w->bp.root(), &np);
    if (rc) return wfail(w, rc, kzg_last_error(w->ctx));
    w->bp.clear_dirty();
    w->device_pairs = np;
    w->pairs_on_device = true;
    return KZG_OK;
}

// The pair query on the device over the moved tree; the pairs stay there.
int device_pairs(kzw_world* w) {
    const uint32_t nn = w->bp.num_nodes();
    w->bp_image.resize(16 * size_t(nn));
    float* img = w->bp_image.data();
    pool(w).run(nn, [w, img](uint32_t b, uint32_t e) { w->bp.image(img + 16 * size_t(b), b, e); });
    const auto& lay = w->bp.layers();
    const auto& msk = w->bp.masks();
    uint32_t np = 0;
    int rc = kzg_broadphase_pairs(w->ctx, reinterpret_cast<const kzg_bp_node*>(w->bp_image.data()), nn,
                                  w->bp.root(), lay.data(), msk.data(), uint32_t(lay.size()), &np);
    if (rc) return wfail(w, rc, kzg_last_error(w->ctx));
    w->device_pairs = np;
    w->pairs_on_device = true;
    w->bp.clear_dirty();
    w->dev_image = true;
    w->dev_nodes = nn;
    if (!w->dev_proxies) {
        std::vector<float> local(6 * w->bodies.size());
        for (size_t i = 0; i < w->bodies.size(); ++i) {
            const Aabb& l = w->shapes[w->bodies[i].shape].local;
            const float v[6] = {l.mn.x, l.mn.y, l.mn.z, l.mx.x, l.mx.y, l.mx.z};
            std::memcpy(&local[6 * i], v, sizeof(v));
        }
        rc = kzg_broadphase_set_proxies(w->ctx, local.data(), w->proxy.data(), uint32_t(w->bodies.size()));
        if (rc) return wfail(w, rc, kzg_last_error(w->ctx));
        w->dev_proxies = true;
    }
    return KZG_OK;
}

// The pair query on host threads.
void host_pairs(kzw_world* w) {
    w->bp.update_pairs();
    const auto& pp = w->bp.pairs();
    w->pairs.resize(pp.size());
    for (size_t k = 0; k < pp.size(); ++k) w->pairs[k] = w->bp.user(pp[k]);
    w->pairs_on_device = false;
}
}  // namespace

int kzw_update_broadphase(kzw_world* w) {
    if (!w) return KZG_ERR_ARGUMENT;
    int rc = sync_tight(w);
    if (rc) return rc;
    w->dev_image = false;  // the host moves the tree from here on
    rc = move_proxies(w);
    if (rc) return rc;
    host_pairs(w);
    return KZG_OK;
}

int kzw_set_host_threads(kzw_world* w, int threads) {
    if (!w || threads < 1 || threads > 256) return KZG_ERR_ARGUMENT;
    w->host_threads = threads;
    w->pool.reset();
    return KZG_OK;
}

int kzw_set_device_broadphase(kzw_world* w, int on) {
    if (!w || on < 0 || on > 2) return KZG_ERR_ARGUMENT;
    w->device_bp = on != 0;
    w->host_moves = on == 2;
    return KZG_OK;
}

int kzw_step(kzw_world* w, float dt) {
    if (!w) return KZG_ERR_ARGUMENT;
    using clock = std::chrono::steady_clock;
    const auto t0 = clock::now();
    int rc = sync_device(w);
    if (rc) return rc;
    uint32_t np = 0;
    const bool dev_moves = w->device_bp && !w->host_moves && w->dev_image && w->dev_proxies;
    w->reinserted = 0;
    if (dev_moves) {
        if ((rc = device_moves(w))) return rc;
    } else {
        if ((rc = sync_tight(w))) return rc;
        if ((rc = move_proxies(w))) return rc;
        w->dev_image = false;  // host moves: the image goes up whole
    }
    const auto tmv = clock::now();
    if (w->device_bp) {
        if (dev_moves && (rc = device_pairs_update(w))) return rc;
        if (!w->dev_image && (rc = device_pairs(w))) return rc;
        np = w->device_pairs;
    } else {
        w->dev_image = false;
        host_pairs(w);
        np = uint32_t(w->pairs.size() / 2);
    }
    const auto t1 = clock::now();
    w->stats.broadphase_move_ms = std::chrono::duration<double, std::milli>(tmv - t0).count();
    w->stats.broadphase_query_ms = std::chrono::duration<double, std::milli>(t1 - tmv).count();
    if ((rc = w->device_bp ? kzg_narrowphase_resident(w->ctx) : kzg_narrowphase(w->ctx, w->pairs.data(), np)) ||
        (rc = kzg_build_constraints(w->ctx, 0)) || (rc = kzg_integrate_velocities(w->ctx, dt)) ||
        (rc = kzg_solve(w->ctx, dt)))
        return wfail(w, rc, kzg_last_error(w->ctx));
    kzg_timings tm;
    if ((rc = kzg_get_timings(w->ctx, &tm))) return wfail(w, rc, kzg_last_error(w->ctx));  // synchronises
    const auto t2 = clock::now();
    if ((rc = kzg_integrate_positions(w->ctx, dt)) || (rc = kzg_update_sleeping(w->ctx, dt)) ||
        (rc = kzg_download_states(w->ctx, w->states.data(), uint32_t(w->states.size()))))
        return wfail(w, rc, kzg_last_error(w->ctx));
    const auto t3 = clock::now();
    uint32_t npairs = 0, nc = 0, npts = 0;
    kzg_query_sizes(w->ctx, &npairs, &nc, &npts);
    w->stats.num_bodies = uint32_t(w->bodies.size());
    w->stats.num_broadphase_pairs = np;
    w->stats.num_contact_constraints = nc;
    w->stats.num_contact_points = npts;
    w->stats.broadphase_reinserted = w->reinserted;
    kzg_get_active_bodies(w->ctx, nullptr, 0, &w->stats.num_active_bodies);
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
    if (w->pairs_on_device) {
        w->pairs.resize(2 * size_t(w->device_pairs));
        const int rc = kzg_download_pairs(w->ctx, w->pairs.data(), w->device_pairs);
        if (rc) return wfail(w, rc, kzg_last_error(w->ctx));
        w->pairs_on_device = false;
    }
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

// ---------------------------------------------------------------------------
// Snapshots.  The file starts with the reference's KZP1 world record
// (src/Serialization/Serializer.cpp:182-205: magic, version, body count, then
// per body its shape, its BodySettings with the current state, and the
// active flag), so the reference's DeserializeWorld reads it as a plain
// world.  A KZP1 reader stops after the last body; what follows is this
// library's extension, which makes a resumed world step on bit-identically:
// the world settings, the shape table with the built hulls, the body
// descriptions and states, the broadphase tree, and the warm-start cache.
// Extension: "KZX1", version, then blocks (u32 tag, u64 length, payload)
// ending with "END ".
// ---------------------------------------------------------------------------
namespace {

constexpr uint32_t kKzp1Magic = 0x31505A4Bu;  // "KZP1"
// Smallest KZP1 body record (a sphere): shape type and radius, then the body
// fields read_kzp1_body reads (118 bytes).
constexpr size_t kMinKzp1Body = 1 + 4 + 12 + 16 + 12 + 12 + 1 + 1 + 4 + 7 * 4 + 4 + 4 + 1 + 1 + 8 + 4 + 4 + 1;
constexpr uint32_t kKzx1Magic = 0x31585A4Bu;  // "KZX1"
constexpr uint32_t tag4(const char (&s)[5]) {
    return uint32_t(uint8_t(s[0])) | uint32_t(uint8_t(s[1])) << 8 | uint32_t(uint8_t(s[2])) << 16 |
           uint32_t(uint8_t(s[3])) << 24;
}

struct Writer {
    std::vector<uint8_t> b;
    template <class T> void put(const T& v) {
        const auto* p = reinterpret_cast<const uint8_t*>(&v);
        b.insert(b.end(), p, p + sizeof(T));
    }
    void bytes(const void* p, size_t n) {
        const auto* q = static_cast<const uint8_t*>(p);
        b.insert(b.end(), q, q + n);
    }
    size_t begin_block(uint32_t tag) {
        put(tag);
        put(uint64_t(0));
        return b.size();
    }
    void end_block(size_t start) {
        const uint64_t len = b.size() - start;
        std::memcpy(b.data() + start - sizeof(uint64_t), &len, sizeof(len));
    }
};

struct Reader {
    const uint8_t* p;
    size_t n, off = 0;
    template <class T> bool get(T& v) {
        if (off + sizeof(T) > n) return false;
        std::memcpy(&v, p + off, sizeof(T));
        off += sizeof(T);
        return true;
    }
    bool bytes(void* dst, size_t k) {
        if (k > n - off) return false;
        std::memcpy(dst, p + off, k);
        off += k;
        return true;
    }
};

// One KZP1 body: the shape (ShapeType code, parameters, hull points) and the
// BodySettings fields in the reference's order.
void write_kzp1_body(Writer& w, const kzw_world& W, uint32_t i, bool active) {
    const kz_body_desc& d = W.bodies[i];
    const kz_body_state& s = W.states[i];
    const auto& sh = W.shapes[d.shape];
    w.put(uint8_t(sh.desc.type));
    switch (sh.desc.type) {
        case KZ_SHAPE_SPHERE: w.put(sh.desc.params[0]); break;
        case KZ_SHAPE_BOX: w.bytes(sh.desc.params, 3 * sizeof(float)); break;
        case KZ_SHAPE_CAPSULE: w.bytes(sh.desc.params, 2 * sizeof(float)); break;
        case KZ_SHAPE_PLANE: w.bytes(sh.desc.params, 4 * sizeof(float)); break;
        default: {  // convex hull: the built hull's points (ConvexHullShape::GetPoints)
            w.put(uint32_t(sh.hull.verts.size()));
            for (const f3& v : sh.hull.verts) {
                w.put(v.x);
                w.put(v.y);
                w.put(v.z);
            }
        }
    }
    w.bytes(s.position, sizeof(s.position));
    w.bytes(s.rotation, sizeof(s.rotation));
    w.bytes(s.linear_velocity, sizeof(s.linear_velocity));
    w.bytes(s.angular_velocity, sizeof(s.angular_velocity));
    w.put(uint8_t(d.motion_type));
    w.put(uint8_t(0));  // motionQuality: Discrete
    w.put(0.0f);        // ccdMotionThreshold
    for (float f : {d.friction, d.restitution, d.linear_damping, d.angular_damping, d.gravity_factor,
                    d.max_linear_velocity, d.max_angular_velocity})
        w.put(f);
    w.put(uint32_t(1));            // objectLayer
    w.put(uint32_t(0xFFFFFFFFu));  // collisionMask
    w.put(uint8_t(d.allow_sleeping ? 1 : 0));
    w.put(uint8_t(d.is_sensor ? 1 : 0));
    w.put(uint64_t(0));  // userData
    w.put(d.mass_override);
    w.put(d.density);
    w.put(uint8_t(active ? 1 : 0));  // Body::IsActive
}

struct Kzp1Body {
    kz_shape_desc shape{};
    std::vector<float> points;
    kz_body_desc body{};
    uint8_t active = 1;
};

bool read_kzp1_body(Reader& r, Kzp1Body& o) {
    uint8_t type = 0;
    if (!r.get(type)) return false;
    o.shape.type = type;
    switch (type) {
        case KZ_SHAPE_SPHERE:
            if (!r.bytes(o.shape.params, sizeof(float))) return false;
            break;
        case KZ_SHAPE_BOX:
            if (!r.bytes(o.shape.params, 3 * sizeof(float))) return false;
            break;
        case KZ_SHAPE_CAPSULE:
            if (!r.bytes(o.shape.params, 2 * sizeof(float))) return false;
            break;
        case KZ_SHAPE_PLANE:
            if (!r.bytes(o.shape.params, 4 * sizeof(float))) return false;
            break;
        case KZ_SHAPE_CONVEX_HULL: {
            uint32_t n = 0;
            if (!r.get(n) || n > 100000u || size_t(n) * 3 * sizeof(float) > r.n - r.off) return false;
            o.points.resize(3 * size_t(n));
            if (!r.bytes(o.points.data(), sizeof(float) * o.points.size())) return false;
            o.shape.hull_count = n;
            break;
        }
        default:
            return false;  // cylinder and other shapes are outside the device path
    }
    kz_body_desc& d = o.body;
    uint8_t motion = 0, quality = 0, sleeping = 0, sensor = 0, active = 1;
    float ccd = 0.0f;
    uint32_t layer = 0, mask = 0;
    uint64_t user = 0;
    if (!r.bytes(d.position, sizeof(d.position)) || !r.bytes(d.rotation, sizeof(d.rotation)) ||
        !r.bytes(d.linear_velocity, sizeof(d.linear_velocity)) ||
        !r.bytes(d.angular_velocity, sizeof(d.angular_velocity)) || !r.get(motion) || !r.get(quality) ||
        !r.get(ccd) || !r.get(d.friction) || !r.get(d.restitution) || !r.get(d.linear_damping) ||
        !r.get(d.angular_damping) || !r.get(d.gravity_factor) || !r.get(d.max_linear_velocity) ||
        !r.get(d.max_angular_velocity) || !r.get(layer) || !r.get(mask) || !r.get(sleeping) || !r.get(sensor) ||
        !r.get(user) || !r.get(d.mass_override) || !r.get(d.density) || !r.get(active))
        return false;
    d.motion_type = motion;
    d.allow_sleeping = sleeping;
    d.is_sensor = sensor;
    o.active = active;
    return true;
}

}  // namespace

extern "C" {

int kzw_save(kzw_world* w, uint8_t* out, uint64_t cap, uint64_t* size) {
    if (!w || !size) return KZG_ERR_ARGUMENT;
    Writer wr;
    const uint32_t nb = uint32_t(w->bodies.size());
    wr.put(kKzp1Magic);
    wr.put(uint32_t(1));
    wr.put(nb);
    std::vector<uint8_t> active(nb, 1);
    if (w->ctx) {
        std::vector<uint32_t> list(nb);
        uint32_t na = 0;
        int rc = kzg_get_active_bodies(w->ctx, list.data(), nb, &na);
        if (rc) return wfail(w, rc, kzg_last_error(w->ctx));
        active.assign(nb, 0);
        for (uint32_t k = 0; k < na && k < nb; ++k) active[list[k]] = 1;
    }
    for (uint32_t i = 0; i < nb; ++i) write_kzp1_body(wr, *w, i, active[i] != 0);
    wr.put(kKzx1Magic);
    wr.put(uint32_t(1));
    size_t blk = wr.begin_block(tag4("WSET"));
    wr.put(w->settings);
    wr.end_block(blk);
    blk = wr.begin_block(tag4("SHAP"));
    wr.put(uint32_t(w->shapes.size()));
    for (const auto& s : w->shapes) {
        wr.put(s.desc);
        wr.put(uint32_t(s.hull.verts.size()));
        wr.put(uint32_t(s.hull.tris.size()));
        wr.bytes(s.hull.verts.data(), sizeof(f3) * s.hull.verts.size());
        wr.bytes(s.hull.tris.data(), sizeof(uint32_t) * s.hull.tris.size());
        wr.put(s.hull.bounds);
        wr.put(s.hull.bounding_radius);
        wr.put(s.local);
    }
    wr.end_block(blk);
    blk = wr.begin_block(tag4("BODY"));
    wr.put(nb);
    wr.bytes(w->bodies.data(), sizeof(kz_body_desc) * nb);
    wr.bytes(w->states.data(), sizeof(kz_body_state) * nb);
    wr.end_block(blk);
    blk = wr.begin_block(tag4("BPHT"));
    std::vector<uint32_t> tree;
    int rc0 = sync_tight(w);  // leaf tight boxes as the device moved them
    if (rc0) return rc0;
    w->bp.save(tree);
    wr.put(uint64_t(tree.size()));
    wr.bytes(tree.data(), sizeof(uint32_t) * tree.size());
    wr.bytes(w->proxy.data(), sizeof(uint32_t) * nb);
    wr.end_block(blk);
    if (w->ctx) {
        uint32_t ccap = 0, used = 0;
        int rc = kzg_cache_info(w->ctx, &ccap, &used);
        if (rc) return wfail(w, rc, kzg_last_error(w->ctx));
        if (ccap) {
            blk = wr.begin_block(tag4("CACH"));
            wr.put(ccap);
            wr.put(used);
            const size_t at = wr.b.size();
            wr.b.resize(at + size_t(ccap) * KZG_CACHE_SLOT_BYTES);
            rc = kzg_download_cache(w->ctx, wr.b.data() + at, uint64_t(ccap) * KZG_CACHE_SLOT_BYTES);
            if (rc) return wfail(w, rc, kzg_last_error(w->ctx));
            wr.end_block(blk);
        }
    }
    wr.put(tag4("END "));
    wr.put(uint64_t(0));
    *size = wr.b.size();
    if (out) {
        if (cap < wr.b.size()) return wfail(w, KZG_ERR_ARGUMENT, "snapshot buffer too small");
        std::memcpy(out, wr.b.data(), wr.b.size());
    }
    return KZG_OK;
}

int kzw_load(kzw_world** out, const uint8_t* data, uint64_t size, const kz_world_desc* settings, int device) {
    if (!out || !data) return KZG_ERR_ARGUMENT;
    *out = nullptr;
    Reader r{data, size_t(size)};
    uint32_t magic = 0, version = 0, nb = 0;
    if (!r.get(magic) || magic != kKzp1Magic || !r.get(version) || version != 1 || !r.get(nb))
        return KZG_ERR_ARGUMENT;
    // Every KZP1 body record holds at least kMinKzp1Body bytes: a count
    // beyond what the record can hold is malformed (and must not size an
    // allocation).
    if (nb > (r.n - r.off) / kMinKzp1Body || nb >= (1u << 24)) return KZG_ERR_ARGUMENT;
    std::vector<Kzp1Body> kb(nb);
    for (auto& b : kb)
        if (!read_kzp1_body(r, b)) return KZG_ERR_UNSUPPORTED;
    // Extension blocks, if any.
    const uint8_t *wset = nullptr, *shap = nullptr, *body = nullptr, *bpht = nullptr, *cach = nullptr;
    size_t lshap = 0, lbody = 0, lbpht = 0, lcach = 0;
    uint32_t xm = 0, xv = 0;
    if (r.get(xm) && xm == kKzx1Magic && r.get(xv) && xv == 1) {
        for (;;) {
            uint32_t tag = 0;
            uint64_t len = 0;
            if (!r.get(tag) || !r.get(len) || len > r.n - r.off) return KZG_ERR_ARGUMENT;
            const uint8_t* p = r.p + r.off;
            if (tag == tag4("END ")) break;
            if (tag == tag4("WSET") && len == sizeof(kz_world_desc)) wset = p;
            else if (tag == tag4("SHAP")) { shap = p; lshap = len; }
            else if (tag == tag4("BODY")) { body = p; lbody = len; }
            else if (tag == tag4("BPHT")) { bpht = p; lbpht = len; }
            else if (tag == tag4("CACH")) { cach = p; lcach = len; }
            r.off += size_t(len);
        }
    }
    kz_world_desc ws{};
    if (wset) std::memcpy(&ws, wset, sizeof(ws));
    else if (settings) ws = *settings;
    else return KZG_ERR_ARGUMENT;  // a plain KZP1 record carries no world settings
    kzw_world* w = nullptr;
    int rc = kzw_create(&w, &ws, device);
    if (rc) return rc;
    auto bad = [&](int code) {
        kzw_destroy(w);
        return code;
    };
    std::vector<kz_body_desc> descs(nb);
    std::vector<kz_body_state> states(nb);
    if (shap && body) {
        // Exact tables: shapes with their built hulls, body descriptions and states.
        Reader s{shap, lshap};
        uint32_t ns = 0;
        if (!s.get(ns)) return bad(KZG_ERR_ARGUMENT);
        for (uint32_t i = 0; i < ns; ++i) {
            kzw_world::ShapeRec rec;
            uint32_t nv = 0, nt = 0;
            if (!s.get(rec.desc) || !s.get(nv) || !s.get(nt) || nv > 256 || nt > 3 * 512 || nt % 3)
                return bad(KZG_ERR_ARGUMENT);
            // The checks kzw_add_shapes applies: a type of the device path,
            // and a hull (only for hull shapes) of at least a tetrahedron.
            const uint32_t t = rec.desc.type;
            if (t != KZ_SHAPE_SPHERE && t != KZ_SHAPE_BOX && t != KZ_SHAPE_CAPSULE && t != KZ_SHAPE_CONVEX_HULL &&
                t != KZ_SHAPE_PLANE)
                return bad(KZG_ERR_UNSUPPORTED);
            if ((t == KZ_SHAPE_CONVEX_HULL) != (nv > 0) || (nv && (nv < 4 || nt < 3 * 4)) || (!nv && nt))
                return bad(KZG_ERR_ARGUMENT);
            rec.hull.verts.resize(nv);
            rec.hull.tris.resize(nt);
            if (!s.bytes(rec.hull.verts.data(), sizeof(f3) * nv) || !s.bytes(rec.hull.tris.data(), sizeof(uint32_t) * nt) ||
                !s.get(rec.hull.bounds) || !s.get(rec.hull.bounding_radius) || !s.get(rec.local))
                return bad(KZG_ERR_ARGUMENT);
            for (uint32_t tri : rec.hull.tris)
                if (tri >= nv) return bad(KZG_ERR_ARGUMENT);
            w->shapes.push_back(std::move(rec));
        }
        w->shapes_dirty = true;
        Reader bb{body, lbody};
        uint32_t nb2 = 0;
        if (!bb.get(nb2) || nb2 != nb || !bb.bytes(descs.data(), sizeof(kz_body_desc) * nb) ||
            !bb.bytes(states.data(), sizeof(kz_body_state) * nb))
            return bad(KZG_ERR_ARGUMENT);
    } else {
        // Plain KZP1: one shape per body, hulls rebuilt from their points as
        // ConvexHullShape::Create builds them.
        for (uint32_t i = 0; i < nb; ++i) {
            kz_shape_desc sd = kb[i].shape;
            sd.hull_offset = 0;
            rc = kzw_add_shapes(w, &sd, 1, kb[i].points.data(), sd.hull_count);
            if (rc) return bad(rc);
            descs[i] = kb[i].body;
            descs[i].shape = i;
            std::memcpy(states[i].position, descs[i].position, sizeof(states[i].position));
            std::memcpy(states[i].rotation, descs[i].rotation, sizeof(states[i].rotation));
            std::memcpy(states[i].linear_velocity, descs[i].linear_velocity, sizeof(states[i].linear_velocity));
            std::memcpy(states[i].angular_velocity, descs[i].angular_velocity, sizeof(states[i].angular_velocity));
        }
    }
    if ((rc = kzw_add_bodies(w, descs.data(), nb))) return bad(rc);
    std::memcpy(w->states.data(), states.data(), sizeof(kz_body_state) * nb);
    w->load_inactive.assign(nb, 0);
    bool any_inactive = false;
    for (uint32_t i = 0; i < nb; ++i) {
        w->load_inactive[i] = kb[i].active == 0;
        any_inactive |= kb[i].active == 0;
        // DeactivateBody zeroes the velocities (BodyManager.cpp:152-153).
        if (kb[i].active == 0)
            for (int k = 0; k < 3; ++k) w->states[i].linear_velocity[k] = w->states[i].angular_velocity[k] = 0.0f;
    }
    if (!any_inactive) w->load_inactive.clear();
    if (bpht) {
        Reader t{bpht, lbpht};
        uint64_t words = 0;
        if (!t.get(words) || words > (t.n - t.off) / 4 || t.n - t.off - words * 4 != size_t(nb) * 4)
            return bad(KZG_ERR_ARGUMENT);
        // The image is copied out: the record need not be 4-byte aligned.
        std::vector<uint32_t> tree(static_cast<size_t>(words)), proxy(nb);
        std::memcpy(tree.data(), bpht + t.off, size_t(words) * 4);
        std::memcpy(proxy.data(), bpht + t.off + words * 4, sizeof(uint32_t) * nb);
        if (!w->bp.restore(tree.data(), tree.size())) return bad(KZG_ERR_ARGUMENT);
        for (uint32_t i = 0; i < nb; ++i)
            if (!w->bp.is_proxy(proxy[i], i)) return bad(KZG_ERR_ARGUMENT);
        w->proxy = proxy;
    }
    if (w->ctx) {
        if ((rc = sync_device(w))) return bad(rc);
        if (cach) {
            Reader c{cach, lcach};
            uint32_t ccap = 0, used = 0;
            if (!c.get(ccap) || !c.get(used) || uint64_t(ccap) * KZG_CACHE_SLOT_BYTES != c.n - c.off)
                return bad(KZG_ERR_ARGUMENT);
            rc = kzg_upload_cache(w->ctx, cach + c.off, uint64_t(ccap) * KZG_CACHE_SLOT_BYTES, ccap, used);
            if (rc) return bad(rc);
        }
    }
    *out = w;
    return KZG_OK;
}

}  // extern "C"
