// host_geometry.cpp - convex hulls, mass properties, bounds and the dynamic
// AABB tree, on the host.  Each routine follows the reference function cited
// beside it, operation for operation (compiled with -ffp-contract=off).
#include "host_geometry.h"

#include <cstring>

#include <algorithm>
#include <atomic>
#include <thread>
#include <cmath>

namespace kz {

namespace {

inline f3 vmin(f3 a, f3 b) { return f3{fmin_(a.x, b.x), fmin_(a.y, b.y), fmin_(a.z, b.z)}; }
inline f3 vmax(f3 a, f3 b) { return f3{fmax_(a.x, b.x), fmax_(a.y, b.y), fmax_(a.z, b.z)}; }
inline void encapsulate(Aabb& a, f3 p) {
    a.mn = vmin(a.mn, p);
    a.mx = vmax(a.mx, p);
}
inline Aabb merge(const Aabb& a, const Aabb& b) { return Aabb{vmin(a.mn, b.mn), vmax(a.mx, b.mx)}; }
inline float surface(const Aabb& a) {
    f3 d = sub(a.mx, a.mn);
    return 2.0f * (d.x * d.y + d.y * d.z + d.z * d.x);
}
inline bool contains(const Aabb& a, const Aabb& o) {
    return a.mn.x <= o.mn.x && a.mx.x >= o.mx.x && a.mn.y <= o.mn.y && a.mx.y >= o.mx.y && a.mn.z <= o.mn.z &&
           a.mx.z >= o.mx.z;
}
inline bool overlaps(const Aabb& a, const Aabb& o) {
    return a.mn.x <= o.mx.x && a.mx.x >= o.mn.x && a.mn.y <= o.mx.y && a.mx.y >= o.mn.y && a.mn.z <= o.mx.z &&
           a.mx.z >= o.mn.z;
}

// Matrix helpers in the reference's association (Math.h Mat3).
inline m3 mscale(const m3& m, float s) { return m3{mul(m.c0, s), mul(m.c1, s), mul(m.c2, s)}; }
inline m3 madd(const m3& a, const m3& b) { return m3{add(a.c0, b.c0), add(a.c1, b.c1), add(a.c2, b.c2)}; }
inline m3 msub(const m3& a, const m3& b) { return m3{sub(a.c0, b.c0), sub(a.c1, b.c1), sub(a.c2, b.c2)}; }
inline m3 identity() { return m3{f3{1, 0, 0}, f3{0, 1, 0}, f3{0, 0, 1}}; }
inline m3 diag(float a, float b, float c) { return m3{f3{a, 0, 0}, f3{0, b, 0}, f3{0, 0, c}}; }
inline m3 outer(f3 a, f3 b) { return m3{mul(a, b.x), mul(a, b.y), mul(a, b.z)}; }
inline float determinant(const m3& m) { return dot(m.c0, cross(m.c1, m.c2)); }
// Mat3::Inversed (src/Core/Math.cpp:27-43).
inline m3 inverse(const m3& m) {
    f3 r0 = cross(m.c1, m.c2), r1 = cross(m.c2, m.c0), r2 = cross(m.c0, m.c1);
    float det = dot(m.c0, r0);
    float inv = det != 0.0f ? 1.0f / det : 0.0f;
    return m3{mul(f3{r0.x, r1.x, r2.x}, inv), mul(f3{r0.y, r1.y, r2.y}, inv), mul(f3{r0.z, r1.z, r2.z}, inv)};
}

constexpr float kPi = 3.14159265358979323846f;

struct HullFace {
    uint32_t a = 0, b = 0, c = 0;
    f3 n{0.0f, 1.0f, 0.0f};
    float dist = 0.0f;
    bool alive = true;
};

// MakeFace (src/Geometry/Shapes.cpp:315-328).
void make_face(HullFace& f, const std::vector<f3>& p, f3 interior) {
    f3 A = p[f.a], B = p[f.b], C = p[f.c];
    f3 n = cross(sub(B, A), sub(C, A));
    float l = len(n);
    if (l < 1.0e-12f) {
        f.alive = false;
        return;
    }
    n = divs(n, l);
    if (dot(n, sub(A, interior)) < 0.0f) {
        std::swap(f.b, f.c);
        n = neg(n);
    }
    f.n = n;
    f.dist = dot(n, A);
}

}  // namespace

bool build_hull(const float* points, uint32_t n, Hull& out, uint32_t max_points) {
    out = Hull{};
    if (n < 4) return false;
    std::vector<f3> pts;
    for (uint32_t i = 0; i < n; ++i) {
        f3 p{points[3 * i], points[3 * i + 1], points[3 * i + 2]};
        bool dup = false;
        for (const f3& q : pts)
            if (len2(sub(p, q)) < 1.0e-10f) {
                dup = true;
                break;
            }
        if (!dup) {
            pts.push_back(p);
            if (pts.size() >= max_points) break;
        }
    }
    if (pts.size() < 4) return false;
    const uint32_t np = uint32_t(pts.size());
    // Initial tetrahedron (Shapes.cpp:348-379).
    uint32_t i0 = 0;
    for (uint32_t i = 1; i < np; ++i)
        if (pts[i].x < pts[i0].x) i0 = i;
    uint32_t i1 = 0;
    float best = -1.0f;
    for (uint32_t i = 0; i < np; ++i) {
        if (i == i0) continue;
        float d = len2(sub(pts[i], pts[i0]));
        if (d > best) { best = d; i1 = i; }
    }
    uint32_t i2 = 0;
    best = -1.0f;
    const f3 axis = normalized_or(sub(pts[i1], pts[i0]), f3{0.0f, 0.0f, 0.0f});
    for (uint32_t i = 0; i < np; ++i) {
        if (i == i0 || i == i1) continue;
        f3 rel = sub(pts[i], pts[i0]);
        float d = len2(sub(rel, mul(axis, dot(rel, axis))));
        if (d > best) { best = d; i2 = i; }
    }
    if (best < 1.0e-12f) return false;
    const f3 pn = normalized_or(cross(sub(pts[i1], pts[i0]), sub(pts[i2], pts[i0])), f3{0.0f, 0.0f, 0.0f});
    uint32_t i3 = 0;
    best = -1.0f;
    for (uint32_t i = 0; i < np; ++i) {
        if (i == i0 || i == i1 || i == i2) continue;
        float d = fabs_(dot(pn, sub(pts[i], pts[i0])));
        if (d > best) { best = d; i3 = i; }
    }
    if (best < 1.0e-10f) return false;
    const f3 interior = mul(add(add(add(pts[i0], pts[i1]), pts[i2]), pts[i3]), 0.25f);
    std::vector<HullFace> faces;
    {
        const uint32_t idx[4] = {i0, i1, i2, i3};
        const int tri[4][3] = {{0, 1, 2}, {0, 2, 3}, {0, 3, 1}, {1, 3, 2}};
        for (auto& t : tri) {
            HullFace f;
            f.a = idx[t[0]]; f.b = idx[t[1]]; f.c = idx[t[2]];
            make_face(f, pts, interior);
            if (f.alive) faces.push_back(f);
        }
    }
    if (faces.size() < 4) return false;
    // Incremental insertion (Shapes.cpp:381-429).
    const float eps = 1.0e-7f;
    std::vector<uint32_t> visible;
    std::vector<std::pair<uint32_t, uint32_t>> horizon;
    for (uint32_t pi = 0; pi < np; ++pi) {
        if (pi == i0 || pi == i1 || pi == i2 || pi == i3) continue;
        const f3 p = pts[pi];
        visible.clear();
        for (size_t fi = 0; fi < faces.size(); ++fi)
            if (faces[fi].alive && dot(faces[fi].n, p) - faces[fi].dist > eps) visible.push_back(uint32_t(fi));
        if (visible.empty()) continue;
        horizon.clear();
        for (uint32_t fi : visible) {
            const HullFace& f = faces[fi];
            const uint32_t ev[3][2] = {{f.a, f.b}, {f.b, f.c}, {f.c, f.a}};
            for (auto& e : ev) {
                bool shared = false;
                for (uint32_t fj : visible) {
                    if (fj == fi) continue;
                    const HullFace& g = faces[fj];
                    const uint32_t gv[3][2] = {{g.a, g.b}, {g.b, g.c}, {g.c, g.a}};
                    for (auto& ge : gv)
                        if (ge[0] == e[1] && ge[1] == e[0]) { shared = true; break; }
                    if (shared) break;
                }
                if (!shared) horizon.push_back({e[0], e[1]});
            }
        }
        for (uint32_t fi : visible) faces[fi].alive = false;
        for (auto& e : horizon) {
            HullFace f;
            f.a = e.second; f.b = e.first; f.c = pi;
            make_face(f, pts, interior);
            if (f.alive) faces.push_back(f);
        }
    }
    // Compact used vertices in first-use order (Shapes.cpp:431-458).
    std::vector<int32_t> remap(np, -1);
    for (const HullFace& f : faces) {
        if (!f.alive) continue;
        const uint32_t t[3] = {f.a, f.b, f.c};
        for (uint32_t k = 0; k < 3; ++k)
            if (remap[t[k]] < 0) {
                remap[t[k]] = int32_t(out.verts.size());
                out.verts.push_back(pts[t[k]]);
            }
        out.tris.push_back(uint32_t(remap[f.a]));
        out.tris.push_back(uint32_t(remap[f.b]));
        out.tris.push_back(uint32_t(remap[f.c]));
    }
    if (out.verts.size() < 4 || out.tris.size() < 12) return false;
    // ComputeDerivedData: bounds and bounding radius (Shapes.cpp:475-492).
    for (const f3& p : out.verts) encapsulate(out.bounds, p);
    const f3 c = mul(add(out.bounds.mn, out.bounds.mx), 0.5f);
    out.bounding_radius = 0.0f;
    for (const f3& p : out.verts) out.bounding_radius = fmax_(out.bounding_radius, len(sub(p, c)));
    return true;
}

MassProps shape_mass(uint32_t type, const float* prm, const Hull* hull, float density) {
    MassProps mp;
    if (type == 0) {  // SphereShape::GetMassProperties (Shapes.cpp:27-33)
        const float r = prm[0];
        const float vol = float(4.0 / 3.0) * kPi * r * r * r;
        mp.mass = vol * density;
        mp.inertia = mscale(identity(), 0.4f * mp.mass * r * r);
    } else if (type == 1) {  // BoxShape (Shapes.cpp:53-65)
        const float hx = prm[0], hy = prm[1], hz = prm[2];
        mp.mass = 8.0f * hx * hy * hz * density;
        const float m = mp.mass;
        mp.inertia = diag(m / 3.0f * (hy * hy + hz * hz), m / 3.0f * (hx * hx + hz * hz), m / 3.0f * (hx * hx + hy * hy));
    } else if (type == 2) {  // CapsuleShape (Shapes.cpp:108-141)
        const float h = prm[0], r = prm[1];
        const float vcyl = kPi * r * r * (2.0f * h);
        const float vsph = float(4.0 / 3.0) * kPi * r * r * r;
        const float vol = vcyl + vsph;
        if (vol <= kEps) return mp;
        const float mcyl = vcyl * density, msph = vsph * density, m = vol * density;
        const float icyl_ax = mcyl * r * r * 0.5f;
        const float icyl_tr = mcyl * (3.0f * r * r + 4.0f * h * h) / 12.0f;
        const float isph = 0.4f * msph * r * r;
        const float dsph = h + 3.0f * r / 8.0f;
        const float isph_tr = isph + msph * dsph * dsph;
        const float ixx = icyl_tr + isph_tr, iyy = icyl_ax + isph;
        mp.mass = m;
        mp.inertia = diag(ixx, iyy, ixx);
    } else if (type == 4 && hull) {  // ConvexHullShape (Shapes.cpp:549-608)
        float total = 0.0f;
        f3 cacc{0.0f, 0.0f, 0.0f};
        m3 iacc{};
        const size_t nt = hull->tris.size() / 3;
        for (size_t f = 0; f < nt; ++f) {
            const f3 a = hull->verts[hull->tris[3 * f]], b = hull->verts[hull->tris[3 * f + 1]],
                     c = hull->verts[hull->tris[3 * f + 2]];
            const float v6 = dot(a, cross(b, c));
            const float v = v6 / 6.0f;
            if (fabs_(v) < 1.0e-12f) continue;
            const float tm = fabs_(v) * density;
            const float sg = v > 0.0f ? 1.0f : -1.0f;
            total += fabs_(v);
            cacc = add(cacc, mul(add(add(a, b), c), v / 4.0f));
            const float a2 = len2(a), b2 = len2(b), c2 = len2(c);
            const float ab = dot(a, b), ac = dot(a, c), bc = dot(b, c);
            const float tr = (a2 + b2 + c2) + (ab + ac + bc);
            const m3 oaa = outer(a, a), obb = outer(b, b), occ = outer(c, c);
            const m3 cab = madd(outer(a, b), outer(b, a)), cac = madd(outer(a, c), outer(c, a)),
                     cbc = madd(outer(b, c), outer(c, b));
            const m3 ti = msub(msub(mscale(identity(), tm / 10.0f * tr), mscale(madd(madd(oaa, obb), occ), tm / 10.0f)),
                               mscale(madd(madd(cab, cac), cbc), tm / 20.0f));
            iacc = madd(iacc, mscale(ti, sg));
        }
        if (total <= kEps) return MassProps{};
        mp.mass = total * density;
        const f3 com = divs(cacc, total);
        mp.com = com;
        const float cc = len2(com);
        const m3 shift = msub(mscale(identity(), mp.mass * cc), mscale(outer(com, com), mp.mass));
        mp.inertia = msub(iacc, shift);
    }
    return mp;
}

float shape_bounding_radius(uint32_t type, const float* prm, const Hull* hull) {
    switch (type) {
        case 0: return prm[0];
        case 1: return len(f3{prm[0], prm[1], prm[2]});
        case 2: return prm[0] + prm[1];
        case 4: return hull ? hull->bounding_radius : 0.0f;
        default: return 0.0f;
    }
}

Aabb shape_local_bounds(uint32_t type, const float* prm, const Hull* hull) {
    switch (type) {
        case 0: return Aabb{f3{-prm[0], -prm[0], -prm[0]}, f3{prm[0], prm[0], prm[0]}};
        case 1: return Aabb{f3{-prm[0], -prm[1], -prm[2]}, f3{prm[0], prm[1], prm[2]}};
        case 2: {
            const f3 e{prm[1], prm[0] + prm[1], prm[1]};
            return Aabb{neg(e), e};
        }
        case 4: return hull ? hull->bounds : Aabb{};
        case 8: {  // PlaneShape::GetLocalBounds (Shape.h:508-515)
            const float big = 1.0e6f;
            const f3 n{prm[0], prm[1], prm[2]};
            const f3 ext = sub(f3{big, big, big}, mul(vabs(n), big));
            const f3 c = mul(n, prm[3]);
            return Aabb{sub(c, ext), add(c, ext)};
        }
        default: return Aabb{};
    }
}

Aabb transform_aabb(const Aabb& a, f4q q, f3 t) {
    const f3 c = mul(add(a.mn, a.mx), 0.5f);
    const f3 e = mul(sub(a.mx, a.mn), 0.5f);
    const m3 m = to_mat3(q);
    const f3 nc = add(rotate(q, c), t);
    const f3 ne{fabs_(m.c0.x) * e.x + fabs_(m.c1.x) * e.y + fabs_(m.c2.x) * e.z,
                fabs_(m.c0.y) * e.x + fabs_(m.c1.y) * e.y + fabs_(m.c2.y) * e.z,
                fabs_(m.c0.z) * e.x + fabs_(m.c1.z) * e.y + fabs_(m.c2.z) * e.z};
    return Aabb{sub(nc, ne), add(nc, ne)};
}

void body_mass(uint32_t motion, uint32_t type, const float* prm, const Hull* hull, float density, float mass_override,
               float& inv_mass, m3& inv_il) {
    inv_mass = 0.0f;
    inv_il = m3{};
    if (motion != 2) return;
    const MassProps props = shape_mass(type, prm, hull, density);
    const float mass = mass_override > 0.0f ? mass_override : props.mass;
    if (mass <= kEps) return;
    inv_mass = 1.0f / mass;
    m3 inertia = props.inertia;
    if (fabs_(determinant(inertia)) < 1.0e-12f) {
        const float r = shape_bounding_radius(type, prm, hull);
        inertia = mscale(identity(), 0.4f * mass * r * r);
    }
    const f3 com = props.com;
    inertia = madd(inertia, mscale(msub(mscale(identity(), len2(com)), outer(com, com)), mass));
    inv_il = inverse(inertia);
}

// ---------------------------------------------------------------------------
// Dynamic AABB tree (src/BroadPhase/BroadPhase.cpp).
// ---------------------------------------------------------------------------
uint32_t BroadPhase::allocate() {
    if (free_ != kNull) {
        uint32_t id = free_;
        free_ = nodes_[id].next;
        nodes_[id] = Node{};
        nodes_[id].next = kNull;
        touch(id);
        return id;
    }
    nodes_.push_back(Node{});
    touch(uint32_t(nodes_.size() - 1));
    return uint32_t(nodes_.size() - 1);
}

void BroadPhase::free_node(uint32_t id) {
    nodes_[id] = Node{};
    nodes_[id].next = free_;
    nodes_[id].height = -1;
    free_ = id;
    touch(id);
}

void BroadPhase::touch(uint32_t id) {
    if (id == kNull) return;
    if (id >= mark_.size()) mark_.resize(std::max<size_t>(nodes_.size(), size_t(id) + 1), 0);
    if (!mark_[id]) {
        mark_[id] = 1;
        dirty_.push_back(id);
    }
}

void BroadPhase::clear_dirty() {
    for (uint32_t id : dirty_) mark_[id] = 0;
    dirty_.clear();
}

namespace {
constexpr size_t kNodeWords = 18;  // fat (6), tight (6), parent, child1, child2, user, next, height
inline uint32_t fbits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}
inline float bitsf(uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}
}  // namespace

void BroadPhase::save(std::vector<uint32_t>& out) const {
    out.clear();
    out.push_back(uint32_t(nodes_.size()));
    out.push_back(root_);
    out.push_back(free_);
    out.push_back(uint32_t(layers_.size()));
    for (const Node& n : nodes_) {
        for (const Aabb* b : {&n.fat, &n.tight}) {
            for (float f : {b->mn.x, b->mn.y, b->mn.z, b->mx.x, b->mx.y, b->mx.z}) out.push_back(fbits(f));
        }
        out.push_back(n.parent);
        out.push_back(n.child1);
        out.push_back(n.child2);
        out.push_back(n.user);
        out.push_back(n.next);
        out.push_back(uint32_t(n.height));
    }
    out.insert(out.end(), layers_.begin(), layers_.end());
    out.insert(out.end(), masks_.begin(), masks_.end());
}

bool BroadPhase::restore(const uint32_t* in, size_t n) {
    if (n < 4) return false;
    const size_t nn = in[0], nl = in[3];
    if (n != 4 + nn * kNodeWords + 2 * nl) return false;
    std::vector<Node> nodes(nn);
    const uint32_t* p = in + 4;
    for (size_t i = 0; i < nn; ++i, p += kNodeWords) {
        Node& d = nodes[i];
        d.fat = Aabb{f3{bitsf(p[0]), bitsf(p[1]), bitsf(p[2])}, f3{bitsf(p[3]), bitsf(p[4]), bitsf(p[5])}};
        d.tight = Aabb{f3{bitsf(p[6]), bitsf(p[7]), bitsf(p[8])}, f3{bitsf(p[9]), bitsf(p[10]), bitsf(p[11])}};
        d.parent = p[12];
        d.child1 = p[13];
        d.child2 = p[14];
        d.user = p[15];
        d.next = p[16];
        d.height = int32_t(p[17]);
        for (uint32_t link : {d.parent, d.child1, d.child2, d.next})
            if (link != kNull && link >= nn) return false;
    }
    if ((in[1] != kNull && in[1] >= nn) || (in[2] != kNull && in[2] >= nn) || nl > nn) return false;
    // The tree under the root: every node reached once, internal nodes with
    // two children whose parent link points back, leaves inside the proxy
    // layer table (the pair query reads it by leaf id).
    if (in[1] != kNull) {
        std::vector<uint8_t> seen(nn, 0);
        std::vector<uint32_t> todo{in[1]};
        while (!todo.empty()) {
            const uint32_t id = todo.back();
            todo.pop_back();
            if (seen[id]) return false;
            seen[id] = 1;
            const Node& d = nodes[id];
            if (d.leaf()) {
                if (d.child2 != kNull || id >= nl) return false;
                continue;
            }
            if (d.child2 == kNull || nodes[d.child1].parent != id || nodes[d.child2].parent != id) return false;
            todo.push_back(d.child1);
            todo.push_back(d.child2);
        }
    }
    nodes_ = std::move(nodes);
    root_ = in[1];
    free_ = in[2];
    layers_.assign(p, p + nl);
    masks_.assign(p + nl, p + 2 * nl);
    pairs_.clear();
    return true;
}

void BroadPhase::image(float* out) const { image(out, 0, uint32_t(nodes_.size())); }

void BroadPhase::image(float* out, uint32_t first, uint32_t end) const {
    for (uint32_t k = first; k < end; ++k) {
        const Node& n = nodes_[k];
        out[0] = n.fat.mn.x; out[1] = n.fat.mn.y; out[2] = n.fat.mn.z; out[3] = bitsf(n.child1);
        out[4] = n.fat.mx.x; out[5] = n.fat.mx.y; out[6] = n.fat.mx.z; out[7] = bitsf(n.child2);
        out[8] = n.tight.mn.x; out[9] = n.tight.mn.y; out[10] = n.tight.mn.z; out[11] = bitsf(n.user);
        out[12] = n.tight.mx.x; out[13] = n.tight.mx.y; out[14] = n.tight.mx.z; out[15] = bitsf(uint32_t(n.height));
        out += 16;
    }
}

bool BroadPhase::is_proxy(uint32_t id, uint32_t user) const {
    return id < nodes_.size() && id < layers_.size() && nodes_[id].leaf() && nodes_[id].height == 0 &&
           nodes_[id].user == user;
}

uint32_t BroadPhase::create_proxy(const Aabb& tight, uint32_t user, uint32_t layer, uint32_t mask) {
    const uint32_t id = allocate();
    const f3 m{0.1f, 0.1f, 0.1f};
    nodes_[id].fat = Aabb{sub(tight.mn, m), add(tight.mx, m)};
    nodes_[id].tight = tight;
    nodes_[id].user = user;
    nodes_[id].height = 0;
    nodes_[id].child1 = nodes_[id].child2 = nodes_[id].parent = kNull;
    insert_leaf(id);
    if (id >= layers_.size()) {
        layers_.resize(id + 1, 0);
        masks_.resize(id + 1, 0);
    }
    layers_[id] = layer;
    masks_[id] = mask;
    return id;
}

bool BroadPhase::move_proxy(uint32_t id, const Aabb& tight, f3 disp) {
    if (set_tight(id, tight)) return false;
    reinsert(id, disp);
    return true;
}

bool BroadPhase::set_tight(uint32_t id, const Aabb& tight) {
    nodes_[id].tight = tight;
    return contains(nodes_[id].fat, tight);
}

void BroadPhase::reinsert(uint32_t id, f3 disp) {
    touch(id);
    remove_leaf(id);
    const Aabb tight = nodes_[id].tight;
    const f3 m{0.1f, 0.1f, 0.1f};
    const f3 d = vabs(disp);
    nodes_[id].fat = Aabb{sub(sub(tight.mn, m), d), add(add(tight.mx, m), d)};
    insert_leaf(id);
}

HostPool::HostPool(unsigned workers) {
    for (unsigned k = 0; k < workers; ++k) threads_.emplace_back([this, k] { loop(k + 1); });
}

HostPool::~HostPool() {
    {
        std::lock_guard<std::mutex> lk(m_);
        stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : threads_) t.join();
}

void HostPool::dispatch(std::function<void(uint32_t)>* job, uint32_t workers) {
    {
        std::lock_guard<std::mutex> lk(m_);
        job_ = job;
        pending_ = workers;
        ++gen_;
    }
    cv_.notify_all();
}

void HostPool::wait() {
    std::unique_lock<std::mutex> lk(m_);
    done_.wait(lk, [this] { return pending_ == 0; });
    job_ = nullptr;
}

// Every worker runs every job once: run() returns only when all have, so a
// worker never misses a generation.
void HostPool::loop(uint32_t k) {
    uint64_t seen = 0;
    for (;;) {
        std::function<void(uint32_t)>* job;
        {
            std::unique_lock<std::mutex> lk(m_);
            cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
            if (stop_) return;
            seen = gen_;
            job = job_;
        }
        (*job)(k);
        {
            std::lock_guard<std::mutex> lk(m_);
            if (--pending_ == 0) done_.notify_one();
        }
    }
}

void BroadPhase::insert_leaf(uint32_t leaf) {
    if (root_ == kNull) {
        root_ = leaf;
        nodes_[leaf].parent = kNull;
        return;
    }
    const Aabb la = nodes_[leaf].fat;
    uint32_t index = root_;
    while (!nodes_[index].leaf()) {
        const uint32_t c1 = nodes_[index].child1, c2 = nodes_[index].child2;
        const float area = surface(nodes_[index].fat);
        const float combined = surface(merge(nodes_[index].fat, la));
        const float cost = 2.0f * combined;
        const float inherit = 2.0f * (combined - area);
        auto child_cost = [&](uint32_t ch) {
            const Aabb mg = merge(la, nodes_[ch].fat);
            if (nodes_[ch].leaf()) return surface(mg) + inherit;
            return (surface(mg) - surface(nodes_[ch].fat)) + inherit;
        };
        const float cost1 = child_cost(c1), cost2 = child_cost(c2);
        if (cost < cost1 && cost < cost2) break;
        index = cost1 < cost2 ? c1 : c2;
    }
    const uint32_t sibling = index;
    const uint32_t old_parent = nodes_[sibling].parent;
    const uint32_t np = allocate();
    touch(old_parent);
    nodes_[np].parent = old_parent;
    nodes_[np].fat = merge(la, nodes_[sibling].fat);
    nodes_[np].tight = nodes_[np].fat;
    nodes_[np].height = nodes_[sibling].height + 1;
    nodes_[np].user = 0;
    if (old_parent != kNull) {
        if (nodes_[old_parent].child1 == sibling) nodes_[old_parent].child1 = np;
        else nodes_[old_parent].child2 = np;
    } else {
        root_ = np;
    }
    nodes_[np].child1 = sibling;
    nodes_[np].child2 = leaf;
    nodes_[sibling].parent = np;
    nodes_[leaf].parent = np;
    index = nodes_[leaf].parent;
    while (index != kNull) {
        index = balance(index);
        touch(index);
        const uint32_t c1 = nodes_[index].child1, c2 = nodes_[index].child2;
        nodes_[index].height = 1 + std::max(nodes_[c1].height, nodes_[c2].height);
        nodes_[index].fat = merge(nodes_[c1].fat, nodes_[c2].fat);
        index = nodes_[index].parent;
    }
}

void BroadPhase::remove_leaf(uint32_t leaf) {
    if (leaf == root_) {
        root_ = kNull;
        return;
    }
    const uint32_t parent = nodes_[leaf].parent;
    const uint32_t grand = nodes_[parent].parent;
    const uint32_t sibling = nodes_[parent].child1 == leaf ? nodes_[parent].child2 : nodes_[parent].child1;
    if (grand != kNull) {
        touch(grand);
        if (nodes_[grand].child1 == parent) nodes_[grand].child1 = sibling;
        else nodes_[grand].child2 = sibling;
        nodes_[sibling].parent = grand;
        free_node(parent);
        uint32_t index = grand;
        while (index != kNull) {
            index = balance(index);
            touch(index);
            const uint32_t c1 = nodes_[index].child1, c2 = nodes_[index].child2;
            nodes_[index].fat = merge(nodes_[c1].fat, nodes_[c2].fat);
            nodes_[index].height = 1 + std::max(nodes_[c1].height, nodes_[c2].height);
            index = nodes_[index].parent;
        }
    } else {
        root_ = sibling;
        nodes_[sibling].parent = kNull;
        free_node(parent);
    }
}

// DynamicAABBTree::Balance (BroadPhase.cpp:155-248): AVL-style rotations.
uint32_t BroadPhase::balance(uint32_t ia) {
    Node& A = nodes_[ia];
    if (A.leaf() || A.height < 2) return ia;
    const uint32_t ib = A.child1, ic = A.child2;
    Node& B = nodes_[ib];
    Node& C = nodes_[ic];
    const int32_t bal = C.height - B.height;
    // A rotation rewrites A, the child that takes its place (internal: it is
    // two levels taller than its sibling) and the parent's child pointer;
    // the other child and the grandchildren only get new parents, which no
    // image record holds.  (A leaf must not be marked: its tight box may be
    // newer elsewhere than here.)
    if (bal > 1) {
        touch(ia);
        touch(ic);
        touch(A.parent);
    } else if (bal < -1) {
        touch(ia);
        touch(ib);
        touch(A.parent);
    }
    if (bal > 1) {
        const uint32_t i_f = C.child1, i_g = C.child2;
        Node& F = nodes_[i_f];
        Node& G = nodes_[i_g];
        C.child1 = ia;
        C.parent = A.parent;
        A.parent = ic;
        if (C.parent != kNull) {
            if (nodes_[C.parent].child1 == ia) nodes_[C.parent].child1 = ic;
            else nodes_[C.parent].child2 = ic;
        } else {
            root_ = ic;
        }
        if (F.height > G.height) {
            C.child2 = i_f;
            A.child2 = i_g;
            G.parent = ia;
            A.fat = merge(B.fat, G.fat);
            C.fat = merge(A.fat, F.fat);
            A.height = 1 + std::max(B.height, G.height);
            C.height = 1 + std::max(A.height, F.height);
        } else {
            C.child2 = i_g;
            A.child2 = i_f;
            F.parent = ia;
            A.fat = merge(B.fat, F.fat);
            C.fat = merge(A.fat, G.fat);
            A.height = 1 + std::max(B.height, F.height);
            C.height = 1 + std::max(A.height, G.height);
        }
        return ic;
    }
    if (bal < -1) {
        const uint32_t i_d = B.child1, i_e = B.child2;
        Node& D = nodes_[i_d];
        Node& E = nodes_[i_e];
        B.child1 = ia;
        B.parent = A.parent;
        A.parent = ib;
        if (B.parent != kNull) {
            if (nodes_[B.parent].child1 == ia) nodes_[B.parent].child1 = ib;
            else nodes_[B.parent].child2 = ib;
        } else {
            root_ = ib;
        }
        if (D.height > E.height) {
            B.child2 = i_d;
            A.child1 = i_e;
            E.parent = ia;
            A.fat = merge(C.fat, E.fat);
            B.fat = merge(A.fat, D.fat);
            A.height = 1 + std::max(C.height, E.height);
            B.height = 1 + std::max(A.height, D.height);
        } else {
            B.child2 = i_e;
            A.child1 = i_d;
            D.parent = ia;
            A.fat = merge(C.fat, D.fat);
            B.fat = merge(A.fat, E.fat);
            A.height = 1 + std::max(C.height, D.height);
            B.height = 1 + std::max(A.height, E.height);
        }
        return ib;
    }
    return ia;
}

// QueryPairs (BroadPhase.h:61-104) + UpdatePairs (BroadPhase.cpp:368-381).
// One step of the reference's self-query (DynamicAABBTree::QueryPairs,
// include/Kizuri/BroadPhase/BroadPhase.h:62-104, with UpdatePairs' filters,
// src/BroadPhase/BroadPhase.cpp:368-381): pop the top entry, emit a leaf
// pair or push the entries to visit.
void BroadPhase::query_step(std::vector<Entry>& stack, std::vector<uint32_t>& out) const {
    const Entry e = stack.back();
    stack.pop_back();
    const uint32_t n1 = e.n1, n2 = e.n2;
    const Node& a = nodes_[n1];
    if (n1 == n2) {
        if (a.leaf()) return;
        stack.push_back(Entry{a.child1, a.child1});
        stack.push_back(Entry{a.child2, a.child2});
        stack.push_back(Entry{a.child1, a.child2});
        return;
    }
    const Node& b = nodes_[n2];
    if (!overlaps(a.fat, b.fat)) return;
    if (a.leaf() && b.leaf()) {
        if (n1 < n2) {
            if ((layers_[n1] & masks_[n2]) == 0 || (layers_[n2] & masks_[n1]) == 0) return;
            if (!overlaps(nodes_[n1].tight, nodes_[n2].tight)) return;
            out.push_back(n1);
            out.push_back(n2);
        }
        return;
    }
    if (a.leaf() || (!b.leaf() && a.height < b.height)) {
        stack.push_back(Entry{n1, b.child1});
        stack.push_back(Entry{n1, b.child2});
    } else {
        stack.push_back(Entry{a.child1, n2});
        stack.push_back(Entry{a.child2, n2});
    }
}

// The pairs in the reference's order.  The traversal is a depth-first walk
// with a LIFO stack, so the output is that of a pre-order walk of the
// traversal tree whose children are visited in reverse push order.  Large
// trees are split breadth-first while keeping that order: a list of items
// (an emitted pair, or an entry still to walk) is refined by replacing each
// entry with what popping it yields, in visiting order, until the list holds
// enough entries; the entries' subtrees are then walked on host threads and
// everything is concatenated in list order.
void BroadPhase::update_pairs() {
    pairs_.clear();
    if (root_ == kNull) return;
    stack_.clear();
    stack_.push_back(Entry{root_, root_});
    const size_t leaves = (nodes_.size() + 1) / 2;
    const unsigned hw = std::thread::hardware_concurrency();
    const unsigned threads = std::min(hw ? hw : 1u, 16u);
    if (leaves < kParallelLeaves || threads < 2) {
        while (!stack_.empty()) query_step(stack_, pairs_);
        return;
    }
    // Items: an entry to walk (n1, n2), or an emitted pair (pair flag set).
    struct Item {
        Entry e;
        bool pair;
    };
    std::vector<Item> list{{Entry{root_, root_}, false}}, next;
    std::vector<Entry> st;
    std::vector<uint32_t> out;
    size_t pending = 1;
    for (int round = 0; round < 64 && pending > 0 && pending < kSplitEntries; ++round) {
        next.clear();
        pending = 0;
        for (const Item& it : list) {
            if (it.pair) {
                next.push_back(it);
                continue;
            }
            // One pop of this entry: its output, then the pushed entries in
            // the order the stack will pop them.
            st.assign(1, it.e);
            out.clear();
            query_step(st, out);
            for (size_t k = 0; k < out.size(); k += 2) next.push_back(Item{Entry{out[k], out[k + 1]}, true});
            for (size_t k = st.size(); k-- > 0;) next.push_back(Item{st[k], false});
            pending += st.size();
        }
        list.swap(next);
    }
    std::vector<size_t> todo;
    for (size_t k = 0; k < list.size(); ++k)
        if (!list[k].pair) todo.push_back(k);
    std::vector<std::vector<uint32_t>> outs(list.size());
    std::atomic<size_t> cursor{0};
    auto worker = [&] {
        std::vector<Entry> ws;
        for (size_t i; (i = cursor.fetch_add(1)) < todo.size();) {
            ws.assign(1, list[todo[i]].e);
            while (!ws.empty()) query_step(ws, outs[todo[i]]);
        }
    };
    std::vector<std::thread> pool;
    for (unsigned t = 1; t < threads; ++t) pool.emplace_back(worker);
    worker();
    for (auto& th : pool) th.join();
    for (size_t k = 0; k < list.size(); ++k) {
        if (list[k].pair) {
            pairs_.push_back(list[k].e.n1);
            pairs_.push_back(list[k].e.n2);
        } else {
            pairs_.insert(pairs_.end(), outs[k].begin(), outs[k].end());
        }
    }
    stack_.clear();
}

}  // namespace kz
