// kzoracle.cpp - CPU oracle: clean-room restatement of the reference engine's
// per-step contact path, templated on the scalar type.
//
// TEST INFRASTRUCTURE ONLY (the checker).  Loaded by tests/, by
// __graft_entry__.smoke() and by bench.py's cpu_baseline leg; never by the
// product.  Instantiated with float it must reproduce the reference bit for
// bit; instantiated with double it is the FP64 re-run the reference itself
// cannot build (its SIMD translation unit does not compile in double).
//
// Bit-exactness rules followed throughout: IEEE + - * / sqrt only, no
// contraction (built with -ffp-contract=off), the reference's operation order
// and association, `Vec3 / s` as multiplication by (1 / s)
// (include/Kizuri/Core/Math.h:39,45), strict comparisons and first-index tie
// breaks where the reference has them.
//
// Parity pins (tests/test_oracle.py): every function below is checked
// bit-for-bit against golden fixtures dumped from the reference build
// (tests/golden/) and, independently of the reference, against closed forms
// (sphere-sphere, box-on-plane corners, one box resting on a plane: lambda_n
// -> m g dt, momentum conservation of a free pair).
#include "kzoracle.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <thread>
#include <type_traits>
#include <unordered_map>
#include <vector>

namespace kzo {

using u32 = uint32_t;
using u64 = uint64_t;

// ---------------------------------------------------------------------------
// Math (include/Kizuri/Core/Math.h, src/Core/Math.cpp)
// ---------------------------------------------------------------------------
template <class R> struct V3 {
    R x = 0, y = 0, z = 0;
    V3() = default;
    V3(R a, R b, R c) : x(a), y(b), z(c) {}
    R operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    R& at(int i) { return i == 0 ? x : (i == 1 ? y : z); }
};
template <class R> inline V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> inline V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> inline V3<R> operator-(V3<R> a) { return {-a.x, -a.y, -a.z}; }
template <class R> inline V3<R> operator*(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
// Math.h:39 - division by a scalar is a multiplication by its reciprocal.
template <class R> inline V3<R> divs(V3<R> a, R s) { R inv = R(1) / s; return {a.x * inv, a.y * inv, a.z * inv}; }
template <class R> inline R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> inline V3<R> cross(V3<R> a, V3<R> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class R> inline R len2(V3<R> a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
template <class R> inline R len(V3<R> a) { return std::sqrt(len2(a)); }
template <class R> constexpr R kEps = R(1.1920928955078125e-7);
template <class R> inline V3<R> normalized_or(V3<R> a, V3<R> fb) {
    R l = len(a);
    return l >= kEps<R> ? divs(a, l) : fb;
}
template <class R> inline V3<R> vabs(V3<R> a) {
    auto ab = [](R v) { return v < R(0) ? -v : v; };
    return {ab(a.x), ab(a.y), ab(a.z)};
}
template <class R> inline int max_index(V3<R> a) { return a.x >= a.y ? (a.x >= a.z ? 0 : 2) : (a.y >= a.z ? 1 : 2); }
template <class R> inline R rmin(R a, R b) { return a < b ? a : b; }
template <class R> inline R rmax(R a, R b) { return a > b ? a : b; }
template <class R> inline R rclamp(R v, R lo, R hi) { return v < lo ? lo : (v > hi ? hi : v); }
template <class R> inline R rabs(R v) { return v < R(0) ? -v : v; }
// Vec3::GetNormalizedPerpendicular (Math.h:101-110): true divisions here.
template <class R> inline V3<R> perpendicular(V3<R> n) {
    if (rabs(n.x) > rabs(n.y)) {
        R l = std::sqrt(n.x * n.x + n.z * n.z);
        return {-n.z / l, R(0), n.x / l};
    }
    R l = std::sqrt(n.y * n.y + n.z * n.z);
    return {R(0), n.z / l, -n.y / l};
}

template <class R> struct Q4 {
    R x = 0, y = 0, z = 0, w = 1;
};
template <class R> inline Q4<R> qmul(Q4<R> a, Q4<R> q) {
    return {a.w * q.x + a.x * q.w + a.y * q.z - a.z * q.y,
            a.w * q.y - a.x * q.z + a.y * q.w + a.z * q.x,
            a.w * q.z + a.x * q.y - a.y * q.x + a.z * q.w,
            a.w * q.w - a.x * q.x - a.y * q.y - a.z * q.z};
}
template <class R> inline Q4<R> qnormalized(Q4<R> q) {
    R l = std::sqrt(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
    if (!(l >= kEps<R>)) return Q4<R>{};
    R s = R(1) / l;
    return {q.x * s, q.y * s, q.z * s, q.w * s};
}
// Quat::Rotate / InverseRotate (Math.h:296-307): t = 2 (q x v), v + w t + q x t.
template <class R> inline V3<R> rotate(Q4<R> q, V3<R> v) {
    V3<R> qv{q.x, q.y, q.z};
    V3<R> t = cross(qv, v) * R(2);
    return v + t * q.w + cross(qv, t);
}
template <class R> inline V3<R> inv_rotate(Q4<R> q, V3<R> v) {
    V3<R> qv{-q.x, -q.y, -q.z};
    V3<R> t = cross(qv, v) * R(2);
    return v + t * q.w + cross(qv, t);
}
// q' = (q + (0.5 dt w, 0) q) normalised  (src/World/PhysicsWorld.cpp:31-35).
template <class R> inline Q4<R> integrate_rotation(Q4<R> q, V3<R> w, R dt) {
    Q4<R> dq{R(0.5) * dt * w.x, R(0.5) * dt * w.y, R(0.5) * dt * w.z, R(0)};
    Q4<R> p = qmul(dq, q);
    return qnormalized(Q4<R>{q.x + p.x, q.y + p.y, q.z + p.z, q.w + p.w});
}

template <class R> struct Xf {
    Q4<R> q;
    V3<R> t;
    V3<R> apply(V3<R> p) const { return rotate(q, p) + t; }
};

// Column-major 3x3 (Math.h Mat3).
template <class R> struct M3 {
    V3<R> c[3];
    V3<R> mul(V3<R> v) const { return c[0] * v.x + c[1] * v.y + c[2] * v.z; }
    M3 mul(const M3& o) const { return M3{{mul(o.c[0]), mul(o.c[1]), mul(o.c[2])}}; }
    M3 transposed() const {
        return M3{{V3<R>{c[0].x, c[1].x, c[2].x}, V3<R>{c[0].y, c[1].y, c[2].y}, V3<R>{c[0].z, c[1].z, c[2].z}}};
    }
};
// Quat::ToMat3 (Math.h:316-324).
template <class R> inline M3<R> to_mat3(Q4<R> q) {
    R xx = q.x * q.x, yy = q.y * q.y, zz = q.z * q.z;
    R xy = q.x * q.y, xz = q.x * q.z, yz = q.y * q.z;
    R wx = q.w * q.x, wy = q.w * q.y, wz = q.w * q.z;
    return M3<R>{{V3<R>{R(1) - R(2) * (yy + zz), R(2) * (xy + wz), R(2) * (xz - wy)},
                  V3<R>{R(2) * (xy - wz), R(1) - R(2) * (xx + zz), R(2) * (yz + wx)},
                  V3<R>{R(2) * (xz + wy), R(2) * (yz - wx), R(1) - R(2) * (xx + yy)}}};
}
// MotionProperties::UpdateWorldInertia (include/Kizuri/Body/Body.h:122-125).
template <class R> inline M3<R> world_inertia(Q4<R> q, const M3<R>& local) {
    M3<R> r = to_mat3(q);
    return r.mul(local).mul(r.transposed());
}

// ---------------------------------------------------------------------------
// Shapes and support mappings (include/Kizuri/Geometry/Shape.h,
// src/Geometry/Shapes.cpp:538-547)
// ---------------------------------------------------------------------------
enum { SPHERE = 0, BOX = 1, CAPSULE = 2, HULL = 4, PLANE = 8 };

template <class R> struct Shape {
    int type = SPHERE;
    R radius = 0, half_height = 0;
    V3<R> half;              // box
    V3<R> pn; R pd = 0;      // plane
    std::vector<V3<R>> verts;
    std::vector<u32> tris;   // 3 per triangle
};

template <class R> V3<R> support(const Shape<R>& s, V3<R> d) {
    switch (s.type) {
        case SPHERE: return normalized_or(d, V3<R>{0, 1, 0}) * s.radius;
        case BOX: return {d.x >= R(0) ? s.half.x : -s.half.x, d.y >= R(0) ? s.half.y : -s.half.y,
                          d.z >= R(0) ? s.half.z : -s.half.z};
        case CAPSULE: {
            R h = d.y >= R(0) ? s.half_height : -s.half_height;
            return V3<R>{R(0), h, R(0)} + normalized_or(d, V3<R>{0, 1, 0}) * s.radius;
        }
        case HULL: {
            R best = -std::numeric_limits<R>::infinity();
            V3<R> bp = s.verts[0];
            for (const V3<R>& p : s.verts) {
                R v = dot(p, d);
                if (v > best) { best = v; bp = p; }
            }
            return bp;
        }
        case PLANE: return s.pn * s.pd + d * R(1.0e6);
    }
    return {};
}

template <class R> struct SP {  // SupportPoint (GJK.h:12-22)
    V3<R> a, b, p;
};
template <class R> struct SupportMap {  // GJK.h:25-50
    const Shape<R>* A; const Shape<R>* B; Xf<R> ta, tb;
    SP<R> operator()(V3<R> d) const {
        V3<R> la = inv_rotate(ta.q, d);
        V3<R> lb = inv_rotate(tb.q, -d);
        SP<R> r;
        r.a = ta.apply(support(*A, la));
        r.b = tb.apply(support(*B, lb));
        r.p = r.a - r.b;
        return r;
    }
};

// ---------------------------------------------------------------------------
// GJK (src/Collision/GJK.cpp)
// ---------------------------------------------------------------------------
template <class R> V3<R> closest_on_triangle(V3<R> p, V3<R> a, V3<R> b, V3<R> c, R& u, R& v, R& w) {
    V3<R> ab = b - a, ac = c - a, ap = p - a;
    R d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= R(0) && d2 <= R(0)) { u = 1; v = 0; w = 0; return a; }
    V3<R> bp = p - b;
    R d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= R(0) && d4 <= d3) { u = 0; v = 1; w = 0; return b; }
    R vc = d1 * d4 - d3 * d2;
    if (vc <= R(0) && d1 >= R(0) && d3 <= R(0)) {
        R t = d1 / (d1 - d3);
        u = 1 - t; v = t; w = 0;
        return a + ab * t;
    }
    V3<R> cp = p - c;
    R d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= R(0) && d5 <= d6) { u = 0; v = 0; w = 1; return c; }
    R vb = d5 * d2 - d1 * d6;
    if (vb <= R(0) && d2 >= R(0) && d6 <= R(0)) {
        R t = d2 / (d2 - d6);
        u = 1 - t; v = 0; w = t;
        return a + ac * t;
    }
    R va = d3 * d6 - d5 * d4;
    if (va <= R(0) && (d4 - d3) >= R(0) && (d5 - d6) >= R(0)) {
        R t = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        u = 0; v = 1 - t; w = t;
        return b + (c - b) * t;
    }
    R den = R(1) / (va + vb + vc);
    v = vb * den;
    w = vc * den;
    u = R(1) - v - w;
    return a + ab * v + ac * w;
}

template <class R> bool reduce_simplex(SP<R>* s, u32& n, V3<R>& closest) {
    const V3<R> O{};
    if (n == 1) {
        closest = s[0].p;
    } else if (n == 2) {
        V3<R> a = s[0].p, b = s[1].p, ab = b - a;
        R t = -dot(a, ab);
        if (t <= R(0)) { closest = a; n = 1; }
        else if (t >= len2(ab)) { s[0] = s[1]; closest = b; n = 1; }
        else closest = a + ab * (t / len2(ab));
    } else if (n == 3) {
        R u, v, w;
        closest = closest_on_triangle(O, s[0].p, s[1].p, s[2].p, u, v, w);
        SP<R> keep[3];
        u32 k = 0;
        if (u > R(0)) keep[k++] = s[0];
        if (v > R(0)) keep[k++] = s[1];
        if (w > R(0)) keep[k++] = s[2];
        if (k == 0) k = 1;
        for (u32 i = 0; i < k; ++i) s[i] = keep[i];
        n = k;
    } else if (n == 4) {
        auto outside = [&](V3<R> p0, V3<R> p1, V3<R> p2, V3<R> opp) {
            V3<R> nn = cross(p1 - p0, p2 - p0);
            if (dot(nn, opp - p0) > R(0)) nn = -nn;
            return dot(nn, O - p0) > R(0);
        };
        bool oABC = outside(s[0].p, s[1].p, s[2].p, s[3].p);
        bool oACD = outside(s[0].p, s[2].p, s[3].p, s[1].p);
        bool oADB = outside(s[0].p, s[3].p, s[1].p, s[2].p);
        bool oBDC = outside(s[1].p, s[3].p, s[2].p, s[0].p);
        if (!oABC && !oACD && !oADB && !oBDC) return true;
        R best = std::numeric_limits<R>::infinity();
        SP<R> tri[3];
        u32 cnt = 0;
        auto face = [&](u32 i0, u32 i1, u32 i2, bool out) {
            if (!out) return;
            R u, v, w;
            V3<R> p = closest_on_triangle(O, s[i0].p, s[i1].p, s[i2].p, u, v, w);
            R d2 = len2(p);
            if (d2 < best) {
                best = d2;
                u32 k = 0;
                if (u > R(0)) tri[k++] = s[i0];
                if (v > R(0)) tri[k++] = s[i1];
                if (w > R(0)) tri[k++] = s[i2];
                cnt = k;
                closest = p;
            }
        };
        face(0, 1, 2, oABC);
        face(0, 2, 3, oACD);
        face(0, 3, 1, oADB);
        face(1, 3, 2, oBDC);
        for (u32 i = 0; i < cnt; ++i) s[i] = tri[i];
        n = cnt;
    }
    return false;
}

template <class R> struct GjkOut {
    bool intersect = false;
    R distance = 0;
    V3<R> pa, pb, normal;
    u32 n = 0;
    SP<R> simplex[4];
};

template <class R> GjkOut<R> gjk(const SupportMap<R>& map, R tol, u32 max_iter = 32) {
    GjkOut<R> res;
    SP<R> s[4];
    u32 n = 0;
    s[n++] = map(V3<R>{1, 0, 0});
    V3<R> closest = s[0].p;
    const R tol2 = tol * tol;
    for (u32 it = 0; it < max_iter; ++it) {
        if (len2(closest) <= tol2) { res.intersect = true; res.distance = 0; break; }
        V3<R> d = -closest;
        SP<R> w = map(d);
        R progress = len2(closest) - dot(closest, w.p);
        if (progress <= tol * len2(closest)) break;
        bool dup = false;
        for (u32 i = 0; i < n; ++i)
            if (len2(s[i].p - w.p) < tol2) { dup = true; break; }
        if (dup) break;
        s[n++] = w;
        if (reduce_simplex(s, n, closest)) { res.intersect = true; res.distance = 0; break; }
    }
    res.n = n;
    for (u32 i = 0; i < n; ++i) res.simplex[i] = s[i];
    if (!res.intersect) {
        R dist = len(closest);
        res.distance = dist;
        if (dist < kEps<R>) res.intersect = true;
        else res.normal = divs(-closest, dist);
        if (n == 1) {
            res.pa = s[0].a; res.pb = s[0].b;
        } else if (n == 2) {
            V3<R> a = s[0].p, b = s[1].p, ab = b - a;
            R t = rclamp(-dot(a, ab) / rmax(len2(ab), kEps<R>), R(0), R(1));
            res.pa = s[0].a + (s[1].a - s[0].a) * t;
            res.pb = s[0].b + (s[1].b - s[0].b) * t;
        } else if (n == 3) {
            R u, v, w;
            closest_on_triangle(V3<R>{}, s[0].p, s[1].p, s[2].p, u, v, w);
            res.pa = s[0].a * u + s[1].a * v + s[2].a * w;
            res.pb = s[0].b * u + s[1].b * v + s[2].b * w;
        } else {
            res.pa = s[0].a; res.pb = s[0].b;
        }
    }
    return res;
}

// ---------------------------------------------------------------------------
// EPA (src/Collision/EPA.cpp)
// ---------------------------------------------------------------------------
template <class R> struct Face {
    u32 a = 0, b = 0, c = 0;
    V3<R> n{0, 1, 0};
    R dist = 0;
    bool alive = true;
};
template <class R> bool build_face(Face<R>& f, const SP<R>* v) {
    V3<R> A = v[f.a].p, B = v[f.b].p, C = v[f.c].p;
    V3<R> n = cross(B - A, C - A);
    R l = len(n);
    if (l < R(1.0e-12)) return false;
    n = divs(n, l);
    R d = dot(n, A);
    if (d < R(0)) { n = -n; d = -d; std::swap(f.b, f.c); }
    f.n = n;
    f.dist = d;
    return true;
}
template <class R> V3<R> perp(V3<R> v) {
    if (rabs(v.x) <= rabs(v.y) && rabs(v.x) <= rabs(v.z)) return {R(0), -v.z, v.y};
    if (rabs(v.y) <= rabs(v.z)) return {-v.z, R(0), v.x};
    return {-v.y, v.x, R(0)};
}
template <class R> bool origin_in_tetra(const SP<R>* v) {
    const V3<R> O{};
    auto ok = [&](V3<R> a, V3<R> b, V3<R> c, V3<R> d) {
        V3<R> n = cross(b - a, c - a);
        if (dot(n, d - a) > R(0)) n = -n;
        return dot(n, O - a) < R(-1.0e-9);
    };
    return ok(v[0].p, v[1].p, v[2].p, v[3].p) && ok(v[0].p, v[2].p, v[3].p, v[1].p) &&
           ok(v[0].p, v[3].p, v[1].p, v[2].p) && ok(v[1].p, v[3].p, v[2].p, v[0].p);
}
template <class R> bool initial_tetra(const SupportMap<R>& map, const GjkOut<R>& g, SP<R>* out) {
    const R tiny = R(1.0e-12);
    u32 n = 0;
    for (u32 i = 0; i < g.n && n < 4; ++i) {
        bool dup = false;
        for (u32 j = 0; j < n; ++j)
            if (len2(out[j].p - g.simplex[i].p) < tiny) dup = true;
        if (!dup) out[n++] = g.simplex[i];
    }
    if (n == 0) out[n++] = map(V3<R>{1, 0, 0});
    if (n == 4 && !origin_in_tetra(out)) n = 3;
    if (n == 1) {
        out[1] = map(-out[0].p);
        if (len2(out[1].p - out[0].p) > tiny) ++n;
    }
    if (n == 2) {
        V3<R> ab = out[1].p - out[0].p;
        if (len2(ab) > tiny) {
            V3<R> d1 = normalized_or(perp(ab), V3<R>{});
            V3<R> d2 = normalized_or(cross(ab, d1), V3<R>{});
            SP<R> k0 = out[0], k1 = out[1];
            out[0] = map(d1); out[1] = map(-d1); out[2] = map(d2); out[3] = map(-d2);
            if (origin_in_tetra(out)) return true;
            out[0] = k0; out[1] = k1;
        }
        SP<R> sp = map(perp(ab));
        if (len2(sp.p - out[0].p) > tiny && len2(sp.p - out[1].p) > tiny) out[n++] = sp;
    }
    if (n == 3) {
        V3<R> tn = cross(out[1].p - out[0].p, out[2].p - out[0].p);
        R l = len(tn);
        if (l > tiny) {
            tn = divs(tn, l);
            R sd = dot(tn, V3<R>{} - out[0].p);
            SP<R> a = map(sd > R(0) ? -tn : tn);
            SP<R> b = map(sd > R(0) ? tn : -tn);
            out[3] = a;
            if (!origin_in_tetra(out)) out[3] = b;
            if (origin_in_tetra(out)) return true;
        }
        n = 3;
    }
    if (n == 4 && origin_in_tetra(out)) return true;
    // Exhaustive fallback over the GJK simplex and 14 fixed directions
    // (EPA.cpp:116-146), first enclosing tetrahedron in i<j<k<l order.
    SP<R> cand[16];
    u32 nc = 0;
    auto add = [&](const SP<R>& sp) {
        for (u32 i = 0; i < nc; ++i)
            if (len2(cand[i].p - sp.p) < tiny) return;
        if (nc < 16) cand[nc++] = sp;
    };
    for (u32 i = 0; i < g.n && i < 4; ++i) add(g.simplex[i]);
    const R s = R(0.5773502692);
    const V3<R> dirs[14] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1},
                            {s, s, s}, {s, s, -s}, {s, -s, s}, {-s, s, s},
                            {-s, -s, s}, {-s, s, -s}, {s, -s, -s}, {-s, -s, -s}};
    for (int i = 0; i < 14; ++i) add(map(dirs[i]));
    for (u32 i = 0; i + 3 < nc; ++i)
        for (u32 j = i + 1; j + 2 < nc; ++j)
            for (u32 k = j + 1; k + 1 < nc; ++k)
                for (u32 l = k + 1; l < nc; ++l) {
                    SP<R> t[4] = {cand[i], cand[j], cand[k], cand[l]};
                    R vol = dot(cross(t[1].p - t[0].p, t[2].p - t[0].p), t[3].p - t[0].p);
                    if (rabs(vol) < tiny) continue;
                    if (origin_in_tetra(t)) {
                        for (int m = 0; m < 4; ++m) out[m] = t[m];
                        return true;
                    }
                }
    return false;
}

template <class R> struct EpaOut {
    bool valid = false;
    R depth = 0;
    V3<R> normal{0, 1, 0}, pa, pb;
};
template <class R> void finish_face(const Face<R>& f, const SP<R>* v, EpaOut<R>& r) {
    V3<R> p = f.n * f.dist;
    V3<R> a = v[f.a].p, b = v[f.b].p, c = v[f.c].p;
    V3<R> v0 = b - a, v1 = c - a, v2 = p - a;
    R d00 = dot(v0, v0), d01 = dot(v0, v1), d11 = dot(v1, v1);
    R d20 = dot(v2, v0), d21 = dot(v2, v1);
    R den = d00 * d11 - d01 * d01;
    R u, vv, w;
    if (rabs(den) < R(1.0e-18)) { u = 1; vv = 0; w = 0; }
    else {
        vv = (d11 * d20 - d01 * d21) / den;
        w = (d00 * d21 - d01 * d20) / den;
        u = R(1) - vv - w;
    }
    r.valid = true;
    r.depth = f.dist;
    r.normal = -f.n;
    r.pa = v[f.a].a * u + v[f.b].a * vv + v[f.c].a * w;
    r.pb = v[f.a].b * u + v[f.b].b * vv + v[f.c].b * w;
}

template <class R> EpaOut<R> epa(const SupportMap<R>& map, const GjkOut<R>& g, R tol, u32 max_iter = 64) {
    EpaOut<R> res;
    std::vector<SP<R>> verts;
    verts.reserve(256);
    SP<R> tet[4];
    if (!initial_tetra(map, g, tet)) return res;
    for (int i = 0; i < 4; ++i) verts.push_back(tet[i]);
    std::vector<Face<R>> faces;
    faces.reserve(512);
    static const u32 kT[4][3] = {{0, 1, 2}, {0, 2, 3}, {0, 3, 1}, {1, 3, 2}};
    for (auto& t : kT) {
        Face<R> f;
        f.a = t[0]; f.b = t[1]; f.c = t[2];
        if (build_face(f, verts.data())) faces.push_back(f);
    }
    if (faces.size() < 4) return res;
    const u32 kMaxV = 256, kMaxF = 512;
    for (u32 it = 0; it < max_iter; ++it) {
        R best = std::numeric_limits<R>::infinity();
        u32 bi = 0xFFFFFFFFu;
        for (u32 i = 0; i < faces.size(); ++i)
            if (faces[i].alive && faces[i].dist < best) { best = faces[i].dist; bi = i; }
        if (bi == 0xFFFFFFFFu) break;
        Face<R> face = faces[bi];
        SP<R> w = map(face.n);
        R d = dot(w.p, face.n);
        if (d - face.dist < tol || verts.size() >= kMaxV) {
            finish_face(face, verts.data(), res);
            return res;
        }
        u32 nv = u32(verts.size());
        verts.push_back(w);
        std::vector<u32> vis;
        for (u32 i = 0; i < faces.size(); ++i)
            if (faces[i].alive && dot(faces[i].n, w.p) - faces[i].dist > tol) {
                faces[i].alive = false;
                vis.push_back(i);
            }
        std::vector<std::pair<u32, u32>> horizon;
        for (u32 vi = 0; vi < vis.size(); ++vi) {
            const Face<R>& f = faces[vis[vi]];
            const u32 e[3][2] = {{f.a, f.b}, {f.b, f.c}, {f.c, f.a}};
            for (auto& ed : e) {
                bool shared = false;
                for (u32 vj = 0; vj < vis.size() && !shared; ++vj) {
                    if (vis[vj] == vis[vi]) continue;
                    const Face<R>& g2 = faces[vis[vj]];
                    const u32 ge[3][2] = {{g2.a, g2.b}, {g2.b, g2.c}, {g2.c, g2.a}};
                    for (auto& x : ge)
                        if (x[0] == ed[1] && x[1] == ed[0]) { shared = true; break; }
                }
                if (!shared && horizon.size() < 256) horizon.push_back({ed[0], ed[1]});
            }
        }
        std::vector<Face<R>> kept;
        kept.reserve(512);
        for (auto& f : faces)
            if (f.alive) kept.push_back(f);
        faces.swap(kept);
        for (auto& h : horizon) {
            if (faces.size() >= kMaxF) break;
            Face<R> f;
            f.a = h.second; f.b = h.first; f.c = nv;
            if (build_face(f, verts.data())) faces.push_back(f);
        }
    }
    R best = std::numeric_limits<R>::infinity();
    u32 bi = 0xFFFFFFFFu;
    for (u32 i = 0; i < faces.size(); ++i)
        if (faces[i].alive && faces[i].dist < best) { best = faces[i].dist; bi = i; }
    if (bi != 0xFFFFFFFFu) finish_face(faces[bi], verts.data(), res);
    return res;
}

// ---------------------------------------------------------------------------
// Manifolds and the narrowphase dispatch (src/Collision/CollideShape.cpp)
// ---------------------------------------------------------------------------
template <class R> struct CP {
    V3<R> pa, pb, n{0, 1, 0};
    R sep = 0;
    u32 feature = 0;
};
template <class R> struct Manifold {
    CP<R> pt[8];
    u32 n = 0;
    void add(const CP<R>& c) { if (n < 8) pt[n++] = c; }
};
template <class R> void swap_manifold(Manifold<R>& m) {
    for (u32 i = 0; i < m.n; ++i) { std::swap(m.pt[i].pa, m.pt[i].pb); m.pt[i].n = -m.pt[i].n; }
}
// Keep points[i*(n-1)/3], i = 0..3, when more than four survive
// (CollideShape.cpp:216-224, 406-411).
template <class R> u32 reduce_to_four(Manifold<R>& m, u32 count) {
    if (count > 4) {
        CP<R> kept[4];
        for (u32 i = 0; i < 4; ++i) kept[i] = m.pt[i * (count - 1) / 3];
        for (u32 i = 0; i < 4; ++i) m.pt[i] = kept[i];
        count = 4;
    }
    return count;
}

constexpr u32 kMaxFaceVerts = 16;

// GetShapeFace (CollideShape.cpp:30-98).
template <class R> bool shape_face(const Shape<R>& s, const Xf<R>& t, V3<R> dir, V3<R>* out, u32& cnt, V3<R>& nw) {
    if (s.type == BOX) {
        V3<R> ld = inv_rotate(t.q, dir);
        int ax = max_index(vabs(ld));
        R sg = ld[ax] >= R(0) ? R(1) : R(-1);
        V3<R> he = s.half;
        V3<R> n{};
        n.at(ax) = sg;
        V3<R> c{he.x * n.x, he.y * n.y, he.z * n.z};
        int a0 = (ax + 1) % 3, a1 = (ax + 2) % 3;
        R s0 = he[a0], s1 = he[a1];
        V3<R> k[4] = {c, c, c, c};
        k[0].at(a0) = s0; k[0].at(a1) = s1;
        k[1].at(a0) = -s0; k[1].at(a1) = s1;
        k[2].at(a0) = -s0; k[2].at(a1) = -s1;
        k[3].at(a0) = s0; k[3].at(a1) = -s1;
        for (int i = 0; i < 4; ++i) out[i] = t.apply(k[i]);
        cnt = 4;
        nw = rotate(t.q, n);
        return true;
    }
    if (s.type == HULL) {
        V3<R> ld = inv_rotate(t.q, dir);
        R best = -std::numeric_limits<R>::infinity();
        u32 bf = 0xFFFFFFFFu;
        u32 nf = u32(s.tris.size() / 3);
        for (u32 f = 0; f < nf; ++f) {
            V3<R> a = s.verts[s.tris[3 * f]], b = s.verts[s.tris[3 * f + 1]], c = s.verts[s.tris[3 * f + 2]];
            V3<R> fn = normalized_or(cross(b - a, c - a), V3<R>{0, 1, 0});
            R d = dot(fn, ld);
            if (d > best) { best = d; bf = f; }
        }
        if (bf == 0xFFFFFFFFu) return false;
        for (u32 i = 0; i < 3; ++i) out[i] = t.apply(s.verts[s.tris[3 * bf + i]]);
        cnt = 3;
        V3<R> a = s.verts[s.tris[3 * bf]], b = s.verts[s.tris[3 * bf + 1]], c = s.verts[s.tris[3 * bf + 2]];
        nw = rotate(t.q, normalized_or(cross(b - a, c - a), V3<R>{0, 1, 0}));
        return true;
    }
    return false;
}

template <class R> struct Poly {
    V3<R> v[2 * kMaxFaceVerts];
    u32 n = 0;
    void add(V3<R> p) { if (n < 2 * kMaxFaceVerts) v[n++] = p; }
};
// Sutherland-Hodgman, keep d <= 0 (CollideShape.cpp:111-126).
template <class R> void clip_plane(const Poly<R>& in, V3<R> pn, R pd, Poly<R>& out) {
    out.n = 0;
    if (in.n == 0) return;
    for (u32 i = 0; i < in.n; ++i) {
        V3<R> cur = in.v[i], nxt = in.v[(i + 1) % in.n];
        R dc = dot(pn, cur) - pd, dn = dot(pn, nxt) - pd;
        if (dc <= R(0)) out.add(cur);
        if ((dc <= R(0)) != (dn <= R(0))) {
            R t = dc / (dc - dn);
            out.add(cur + (nxt - cur) * t);
        }
    }
}

// CollidePolyhedral (CollideShape.cpp:131-228).
template <class R> u32 collide_polyhedral(const Shape<R>& A, const Xf<R>& ta, const Shape<R>& B, const Xf<R>& tb,
                                          R max_sep, Manifold<R>& m, V3<R> normal) {
    V3<R> fa[kMaxFaceVerts], fb[kMaxFaceVerts], na, nb;
    u32 ca = 0, cb = 0;
    bool ha = shape_face(A, ta, -normal, fa, ca, na);
    bool hb = shape_face(B, tb, normal, fb, cb, nb);
    if (!ha || !hb) return 0;
    R align_a = dot(na, -normal), align_b = dot(nb, normal);
    bool ref_a = align_a >= align_b;
    V3<R>* rv = ref_a ? fa : fb; u32 rc = ref_a ? ca : cb;
    V3<R>* iv = ref_a ? fb : fa; u32 ic = ref_a ? cb : ca;
    V3<R> rn = ref_a ? na : nb;
    R rd = dot(rn, rv[0]);
    Poly<R> poly;
    for (u32 i = 0; i < ic; ++i) poly.add(iv[i]);
    V3<R> cen{};
    for (u32 i = 0; i < rc; ++i) cen = cen + rv[i];
    cen = divs(cen, R(rc));
    Poly<R> tmp;
    for (u32 i = 0; i < rc; ++i) {
        V3<R> v0 = rv[i], v1 = rv[(i + 1) % rc];
        V3<R> sn = cross(v1 - v0, rn);
        R l = len(sn);
        if (l < R(1.0e-9)) continue;
        sn = divs(sn, l);
        if (dot(sn, v0 - cen) < R(0)) sn = -sn;
        clip_plane(poly, sn, dot(sn, v0), tmp);
        poly = tmp;
        if (poly.n == 0) break;
    }
    u32 k = 0;
    for (u32 i = 0; i < poly.n && k < 8; ++i) {
        V3<R> p = poly.v[i];
        R sep = dot(rn, p) - rd;
        if (sep > max_sep) continue;
        CP<R> c;
        V3<R> proj = p - rn * sep;
        if (ref_a) { c.pa = proj; c.pb = p; } else { c.pa = p; c.pb = proj; }
        c.n = normal;
        c.sep = sep;
        c.feature = i;
        m.pt[k++] = c;
    }
    k = reduce_to_four(m, k);
    m.n = k;
    return k;
}

template <class R> inline bool polyhedral(const Shape<R>& s) { return s.type == BOX || s.type == HULL; }

// CollideConvexGeneric (CollideShape.cpp:233-299).
template <class R> u32 collide_generic(const Shape<R>& A, const Xf<R>& ta, const Shape<R>& B, const Xf<R>& tb,
                                       R max_sep, R tol, Manifold<R>& m) {
    SupportMap<R> map{&A, &B, ta, tb};
    GjkOut<R> g = gjk(map, tol);
    if (!g.intersect) {
        if (g.distance > max_sep) return 0;
        if (polyhedral(A) && polyhedral(B) && len2(g.normal) > R(0)) {
            if (collide_polyhedral(A, ta, B, tb, max_sep, m, g.normal) > 0) return m.n;
        }
        CP<R> c;
        c.pa = g.pa; c.pb = g.pb; c.n = g.normal; c.sep = g.distance;
        m.add(c);
        return m.n;
    }
    if (polyhedral(A) && polyhedral(B)) {
        EpaOut<R> e = epa(map, g, tol);
        if (e.valid) {
            if (collide_polyhedral(A, ta, B, tb, max_sep, m, e.normal) > 0) return m.n;
            CP<R> c;
            c.pa = e.pa; c.pb = e.pb; c.n = e.normal; c.sep = -e.depth;
            m.add(c);
            return m.n;
        }
    }
    EpaOut<R> e = epa(map, g, tol);
    if (e.valid) {
        CP<R> c;
        c.pa = e.pa; c.pb = e.pb; c.n = e.normal; c.sep = -e.depth;
        m.add(c);
        return m.n;
    }
    CP<R> c;
    c.pa = g.pa; c.pb = g.pb;
    c.n = len2(g.normal) > R(0) ? g.normal : V3<R>{0, 1, 0};
    c.sep = R(0);
    m.add(c);
    return m.n;
}

// CollideSphereConvex (CollideShape.cpp:304-344): GJK against a 1e-4 proxy sphere.
template <class R> u32 collide_sphere_convex(const Shape<R>& S, const Xf<R>& ts, const Shape<R>& B, const Xf<R>& tb,
                                             R max_sep, R tol, Manifold<R>& m) {
    Shape<R> pt;
    pt.type = SPHERE;
    pt.radius = R(1.0e-4);
    SupportMap<R> map{&pt, &B, ts, tb};
    GjkOut<R> g = gjk(map, tol);
    R r = S.radius;
    V3<R> c = ts.t;
    if (g.intersect) {
        EpaOut<R> e = epa(map, g, tol);
        if (e.valid) {
            CP<R> cp;
            cp.n = -e.normal;
            cp.pb = e.pb;
            cp.pa = c + cp.n * r;
            cp.sep = -(e.depth + r);
            m.add(cp);
            return m.n;
        }
        return 0;
    }
    R dist = g.distance;
    R sep = dist - r;
    if (sep > max_sep) return 0;
    V3<R> n = dist > kEps<R> ? divs(c - g.pb, dist) : V3<R>{0, 1, 0};
    CP<R> cp;
    cp.n = n;
    cp.pa = c - n * r;
    cp.pb = g.pb;
    cp.sep = sep;
    m.add(cp);
    return m.n;
}

// CollideSphereSphere (CollideShape.cpp:349-365).
template <class R> u32 collide_sphere_sphere(const Shape<R>& A, const Xf<R>& ta, const Shape<R>& B, const Xf<R>& tb,
                                             R max_sep, Manifold<R>& m) {
    V3<R> d = ta.t - tb.t;
    R dist = len(d);
    R sep = dist - (A.radius + B.radius);
    if (sep > max_sep) return 0;
    V3<R> n = dist > kEps<R> ? divs(d, dist) : V3<R>{0, 1, 0};
    CP<R> c;
    c.n = n;
    c.pa = ta.t - n * A.radius;
    c.pb = tb.t + n * B.radius;
    c.sep = sep;
    m.add(c);
    return m.n;
}

// CollidePlaneConvex (CollideShape.cpp:370-430); A of the result = the convex.
template <class R> u32 collide_plane_convex(const Shape<R>& P, const Xf<R>& tp, const Shape<R>& B, const Xf<R>& tb,
                                            R max_sep, Manifold<R>& m) {
    V3<R> pn = rotate(tp.q, P.pn);
    R pd = P.pd + dot(pn, tp.t);
    if (polyhedral(B)) {
        V3<R> v[kMaxFaceVerts];
        u32 nv = 0;
        if (B.type == BOX) {
            for (int i = 0; i < 8; ++i) {
                V3<R> c{(i & 1) ? B.half.x : -B.half.x, (i & 2) ? B.half.y : -B.half.y, (i & 4) ? B.half.z : -B.half.z};
                v[nv++] = tb.apply(c);
            }
        } else {
            for (size_t i = 0; i < B.verts.size() && nv < kMaxFaceVerts; ++i) v[nv++] = tb.apply(B.verts[i]);
        }
        u32 k = 0;
        for (u32 i = 0; i < nv && k < 8; ++i) {
            R d = dot(pn, v[i]) - pd;
            if (d > max_sep) continue;
            CP<R> c;
            c.n = pn; c.pa = v[i]; c.pb = v[i] - pn * d; c.sep = d;
            m.pt[k++] = c;
        }
        k = reduce_to_four(m, k);
        m.n = k;
        return k;
    }
    V3<R> ld = inv_rotate(tb.q, -pn);
    V3<R> deep = tb.apply(support(B, ld));
    R d = dot(pn, deep) - pd;
    if (d > max_sep) return 0;
    CP<R> c;
    c.n = pn; c.pa = deep; c.pb = deep - pn * d; c.sep = d;
    m.add(c);
    return m.n;
}

// NarrowPhase::Collide dispatch (CollideShape.cpp:480-592), convex shapes only.
template <class R> u32 collide(const Shape<R>& A, const Xf<R>& ta, const Shape<R>& B, const Xf<R>& tb,
                               R max_sep, Manifold<R>& m) {
    const R tol = R(1.0e-4);
    m.n = 0;
    if (A.type == PLANE) {
        u32 n = collide_plane_convex(A, ta, B, tb, max_sep, m);
        if (n > 0) swap_manifold(m);
        return n;
    }
    if (B.type == PLANE) return collide_plane_convex(B, tb, A, ta, max_sep, m);
    if (A.type == SPHERE && B.type == SPHERE) return collide_sphere_sphere(A, ta, B, tb, max_sep, m);
    if (A.type == SPHERE) return collide_sphere_convex(A, ta, B, tb, max_sep, tol, m);
    if (B.type == SPHERE) {
        u32 n = collide_sphere_convex(B, tb, A, ta, max_sep, tol, m);
        if (n > 0) swap_manifold(m);
        return n;
    }
    return collide_generic(A, ta, B, tb, max_sep, tol, m);
}

// ---------------------------------------------------------------------------
// World: constraint build, solve, integration (src/World/PhysicsWorld.cpp,
// src/Solver/ContactSolver.cpp, src/Solver/ContactSolverSIMD.cpp)
// ---------------------------------------------------------------------------
template <class R> struct Body {
    int shape = 0, motion = 2;
    bool sensor = false;
    Xf<R> xf;
    V3<R> v, w, pv, pw, force, torque;
    R inv_mass = 0;
    M3<R> inv_i_local{}, inv_i_world{};
    R friction = 0, restitution = 0, lin_damp = 0, ang_damp = 0, gravity_factor = 1, max_lin = 0, max_ang = 0;
    bool dynamic() const { return motion == 2; }
    bool is_static() const { return motion == 0; }
};

template <class R> struct PointC {  // ContactPointConstraint (ContactSolver.h:13-43)
    V3<R> pa, pb, ra, rb, la, lb, n{0, 1, 0}, t1{1, 0, 0}, t2{0, 0, 1};
    R nm = 0, tm1 = 0, tm2 = 0, ln = 0, lt1 = 0, lt2 = 0, sep = 0, rbias = 0, pbias = 0, bimp = 0;
    R cf = 0, cr = 0, nvel = 0;
    u32 feature = 0;
};
template <class R> struct Constraint {
    int a = -1, b = -1;
    u32 n = 0;
    PointC<R> p[8];
    R friction = R(0.5), restitution = 0;
    bool sensor = false, dyn_a = true, dyn_b = true;
};
template <class R> struct CacheEntry {
    u32 n = 0;
    V3<R> pos[8];
    R ln[8], lt1[8], lt2[8];
};

template <class R> struct World {
    kz_world_desc s{};
    std::vector<Shape<R>> shapes;
    std::vector<Body<R>> bodies;
    std::vector<Manifold<R>> manifolds;
    std::vector<u64> keys;
    std::vector<std::pair<u32, u32>> pairs;
    std::vector<Constraint<R>> cons;
    std::unordered_map<u64, CacheEntry<R>> cache;
    std::vector<double> residuals;
    int warm_mode = 0;

    static u64 pair_key(u32 a, u32 b) {
        if (a > b) std::swap(a, b);
        return (u64(a) << 32) | u64(b);
    }

    // Collide (PhysicsWorld.cpp:350-400), sequential (per-pair slots make
    // the result independent of the thread schedule).
    void narrowphase() {
        manifolds.assign(pairs.size(), Manifold<R>{});
        keys.assign(pairs.size(), 0);
        for (size_t i = 0; i < pairs.size(); ++i) {
            u32 ia = pairs[i].first, ib = pairs[i].second;
            const Body<R>& A = bodies[ia];
            const Body<R>& B = bodies[ib];
            if (!A.dynamic() && !B.dynamic()) continue;
            bool swapped = ia > ib;
            const Body<R>& f = swapped ? B : A;
            const Body<R>& g = swapped ? A : B;
            keys[i] = pair_key(ia, ib);
            collide(shapes[f.shape], f.xf, shapes[g.shape], g.xf, R(0.02), manifolds[i]);
            if (swapped) swap_manifold(manifolds[i]);
        }
    }

    // BuildContactConstraints (PhysicsWorld.cpp:454-535).  Sleeping is not
    // modelled: every body is active, so the wake-up block is a no-op.
    void build() {
        cons.clear();
        for (size_t i = 0; i < manifolds.size(); ++i) {
            const Manifold<R>& m = manifolds[i];
            if (m.n == 0 || keys[i] == 0) continue;
            u32 ia = pairs[i].first, ib = pairs[i].second;
            if (ia > ib) std::swap(ia, ib);
            const Body<R>& A = bodies[ia];
            const Body<R>& B = bodies[ib];
            const CacheEntry<R>* cached = nullptr;
            if (warm_mode == 0) {
                auto it = cache.find(keys[i]);
                if (it != cache.end()) cached = &it->second;
            }
            Constraint<R> c;
            c.a = int(ia); c.b = int(ib);
            c.sensor = A.sensor || B.sensor;
            c.friction = std::sqrt(A.friction * B.friction);
            c.restitution = rmax(A.restitution, B.restitution);
            c.n = m.n;
            for (u32 k = 0; k < m.n; ++k) {
                PointC<R>& p = c.p[k];
                const CP<R>& src = m.pt[k];
                p.pa = src.pa; p.pb = src.pb; p.n = src.n; p.sep = src.sep; p.feature = src.feature;
                p.cf = c.friction; p.cr = c.restitution;
                p.ln = p.lt1 = p.lt2 = 0;
                if (cached) {
                    R bd = R(0.10) * R(0.10);
                    u32 bi = 0xFFFFFFFFu;
                    for (u32 j = 0; j < cached->n; ++j) {
                        R d = len2(cached->pos[j] - src.pa);
                        if (d < bd) { bd = d; bi = j; }
                    }
                    if (bi != 0xFFFFFFFFu) { p.ln = cached->ln[bi]; p.lt1 = cached->lt1[bi]; p.lt2 = cached->lt2[bi]; }
                }
            }
            cons.push_back(c);
        }
    }

    // IntegrateVelocities (PhysicsWorld.cpp:757-783).
    void integrate_velocities(R dt) {
        V3<R> g{R(s.gravity[0]), R(s.gravity[1]), R(s.gravity[2])};
        for (Body<R>& b : bodies) {
            if (!b.dynamic()) continue;
            V3<R> acc = g * b.gravity_factor + b.force * b.inv_mass;
            b.v = b.v + acc * dt;
            b.w = b.w + b.inv_i_world.mul(b.torque) * dt;
            b.v = b.v * (R(1) / (R(1) + b.lin_damp * dt));
            b.w = b.w * (R(1) / (R(1) + b.ang_damp * dt));
            auto clamp_len = [](V3<R> v, R mx) {
                R l = len(v);
                return (l > mx && l >= kEps<R>) ? v * (mx / l) : v;
            };
            b.v = clamp_len(b.v, b.max_lin);
            b.w = clamp_len(b.w, b.max_ang);
            auto fin = [](V3<R> v) { return std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z); };
            if (!fin(b.v) || !fin(b.w)) { b.v = V3<R>{}; b.w = V3<R>{}; }
        }
    }

    // EffectiveMass (ContactSolver.cpp:18-28).
    R eff_mass(const Body<R>& A, const Body<R>& B, V3<R> ra, V3<R> rb, V3<R> d) const {
        V3<R> xa = cross(ra, d), xb = cross(rb, d);
        R k = A.inv_mass + B.inv_mass;
        k = k + dot(xa, A.inv_i_world.mul(xa));
        k = k + dot(xb, B.inv_i_world.mul(xb));
        return k > kEps<R> ? R(1) / k : R(0);
    }

    // PrepareConstraint (ContactSolver.cpp:56-112).
    void prepare(Constraint<R>& c, R dt) {
        const Body<R>& A = bodies[c.a];
        const Body<R>& B = bodies[c.b];
        c.dyn_a = A.dynamic();
        c.dyn_b = B.dynamic();
        for (u32 k = 0; k < c.n; ++k) {
            PointC<R>& p = c.p[k];
            p.ra = p.pa - A.xf.t;
            p.rb = p.pb - B.xf.t;
            p.la = inv_rotate(A.xf.q, p.ra);
            p.lb = inv_rotate(B.xf.q, p.rb);
            V3<R> n = normalized_or(p.n, V3<R>{0, 1, 0});
            p.n = n;
            p.t1 = perpendicular(n);
            p.t2 = cross(n, p.t1);
            p.nm = eff_mass(A, B, p.ra, p.rb, n);
            p.tm1 = eff_mass(A, B, p.ra, p.rb, p.t1);
            p.tm2 = eff_mass(A, B, p.ra, p.rb, p.t2);
            V3<R> va = A.v + cross(A.w, p.pa - A.xf.t);
            V3<R> vb = B.v + cross(B.w, p.pb - B.xf.t);
            p.nvel = dot(va - vb, n);
            p.rbias = 0;
            if (p.nvel < -R(s.restitution_threshold)) p.rbias = -p.cr * p.nvel;
            p.pbias = 0;
            if (p.sep < -R(s.penetration_slop)) {
                p.pbias = R(s.baumgarte) / dt * (-p.sep - R(s.penetration_slop));
                p.pbias = rmin(p.pbias, R(s.max_bias_velocity));
            }
            p.bimp = 0;
            R mf = p.cf * p.ln;
            p.lt1 = rclamp(p.lt1, -mf, mf);
            p.lt2 = rclamp(p.lt2, -mf, mf);
            p.ln = rmax(p.ln, R(0));
        }
    }

    void apply_pair(Constraint<R>& c, V3<R> ra, V3<R> rb, V3<R> j) {
        Body<R>& A = bodies[c.a];
        Body<R>& B = bodies[c.b];
        if (c.dyn_a) { A.v = A.v + j * A.inv_mass; A.w = A.w + A.inv_i_world.mul(cross(ra, j)); }
        if (c.dyn_b) { B.v = B.v - j * B.inv_mass; B.w = B.w - B.inv_i_world.mul(cross(rb, j)); }
    }

    // WarmStart (ContactSolver.cpp:117-132): uses the stored rA/rB.
    void warm_start(Constraint<R>& c) {
        for (u32 k = 0; k < c.n; ++k) {
            PointC<R>& p = c.p[k];
            V3<R> j = p.n * p.ln + p.t1 * p.lt1 + p.t2 * p.lt2;
            j = j * R(s.warm_start_factor);
            apply_pair(c, p.ra, p.rb, j);
        }
    }

    double res_acc = 0;
    u32 res_cnt = 0;

    // One lane of SolveVelocityBatchSIMD (ContactSolverSIMD.cpp:93-256): the
    // lanes of a batch share no body, so a lane-by-lane evaluation with the
    // same operation trees is bit-identical to the 4-wide SSE evaluation.
    // _mm_max_ps / _mm_min_ps are (a > b ? a : b) / (a < b ? a : b).
    void solve_lane(Constraint<R>& c, u32 k) {
        PointC<R>& p = c.p[k];
        Body<R>& A = bodies[c.a];
        Body<R>& B = bodies[c.b];
        const V3<R> ra = rotate(A.xf.q, p.la), rb = rotate(B.xf.q, p.lb);
        V3<R> va = A.v, wa = A.w, vb = B.v, wb = B.w;
        const R ima = A.inv_mass, imb = B.inv_mass;
        const M3<R> ia = A.inv_i_world, ib = B.inv_i_world;
        const R nbias = p.rbias + (s.use_velocity_bias ? p.pbias : R(0));
        R acc_n = p.ln, acc_t[2] = {p.lt1, p.lt2};
        const R zero = R(0);
        auto rel = [&]() {
            V3<R> xa = cross(wa, ra), xb = cross(wb, rb);
            return V3<R>{(va.x + xa.x) - (vb.x + xb.x), (va.y + xa.y) - (vb.y + xb.y), (va.z + xa.z) - (vb.z + xb.z)};
        };
        auto apply = [&](V3<R> j) {
            va = V3<R>{va.x + j.x * ima, va.y + j.y * ima, va.z + j.z * ima};
            wa = wa + ia.mul(cross(ra, j));
            vb = V3<R>{vb.x - j.x * imb, vb.y - j.y * imb, vb.z - j.z * imb};
            wb = wb - ib.mul(cross(rb, j));
        };
        double dsq = 0;
        for (int t = 0; t < 2; ++t) {
            const V3<R> tv = t == 0 ? p.t1 : p.t2;
            const R mass = t == 0 ? p.tm1 : p.tm2;
            const R vt = dot(rel(), tv);
            const R lam = (zero - mass) * vt;
            const R mf = p.cf * acc_n;
            R na = acc_t[t] + lam;
            const R lo = zero - mf;
            na = na > lo ? na : lo;
            na = na < mf ? na : mf;
            const R dl = na - acc_t[t];
            acc_t[t] = na;
            dsq += double(dl) * double(dl);
            apply(tv * dl);
        }
        {
            const R vn = dot(rel(), p.n);
            const R lam = (zero - p.nm) * (vn - nbias);
            R na = acc_n + lam;
            na = na > zero ? na : zero;
            const R dl = na - acc_n;
            acc_n = na;
            dsq += double(dl) * double(dl);
            apply(p.n * dl);
        }
        if (c.dyn_a) { A.v = va; A.w = wa; }
        if (c.dyn_b) { B.v = vb; B.w = wb; }
        p.ln = acc_n; p.lt1 = acc_t[0]; p.lt2 = acc_t[1];
        res_acc += dsq;
        res_cnt += 1;
    }

    // SolveContactVelocity, the scalar path (ContactSolver.cpp:137-181): one
    // relative velocity for both friction axes.
    void solve_scalar(Constraint<R>& c, PointC<R>& p) {
        Body<R>& A = bodies[c.a];
        Body<R>& B = bodies[c.b];
        V3<R> ra = rotate(A.xf.q, p.la), rb = rotate(B.xf.q, p.lb);
        V3<R> vr = (A.v + cross(A.w, ra)) - (B.v + cross(B.w, rb));
        double dsq = 0;
        for (int t = 0; t < 2; ++t) {
            V3<R> tv = t == 0 ? p.t1 : p.t2;
            R mass = t == 0 ? p.tm1 : p.tm2;
            R& acc = t == 0 ? p.lt1 : p.lt2;
            R lam = -mass * dot(vr, tv);
            R mf = p.cf * p.ln;
            R ni = rclamp(acc + lam, -mf, mf);
            lam = ni - acc;
            acc = ni;
            dsq += double(lam) * double(lam);
            apply_pair(c, ra, rb, tv * lam);
        }
        vr = (A.v + cross(A.w, ra)) - (B.v + cross(B.w, rb));
        R bias = p.rbias;
        if (s.use_velocity_bias) bias = bias + p.pbias;
        R lam = -p.nm * (dot(vr, p.n) - bias);
        R ni = rmax(p.ln + lam, R(0));
        lam = ni - p.ln;
        p.ln = ni;
        dsq += double(lam) * double(lam);
        apply_pair(c, ra, rb, p.n * lam);
        res_acc += dsq;
        res_cnt += 1;
    }

    // SolveBiasConstraint (ContactSolver.cpp:257-290).
    void solve_bias(Constraint<R>& c) {
        Body<R>& A = bodies[c.a];
        Body<R>& B = bodies[c.b];
        if (A.inv_mass <= R(0) && B.inv_mass <= R(0)) return;
        for (u32 k = 0; k < c.n; ++k) {
            PointC<R>& p = c.p[k];
            if (p.pbias <= R(0)) continue;
            V3<R> ra = rotate(A.xf.q, p.la), rb = rotate(B.xf.q, p.lb);
            V3<R> vr = (A.pv + cross(A.pw, ra)) - (B.pv + cross(B.pw, rb));
            R lam = -p.nm * (dot(vr, p.n) - p.pbias);
            R ni = rmax(p.bimp + lam, R(0));
            lam = ni - p.bimp;
            p.bimp = ni;
            V3<R> j = p.n * lam;
            if (c.dyn_a) { A.pv = A.pv + j * A.inv_mass; A.pw = A.pw + A.inv_i_world.mul(cross(ra, j)); }
            if (c.dyn_b) { B.pv = B.pv - j * B.inv_mass; B.pw = B.pw - B.inv_i_world.mul(cross(rb, j)); }
        }
    }

    // BuildIslands (PhysicsWorld.cpp:587-664): union by rank, path halving,
    // islands numbered by first root appearance in body order.
    std::vector<std::vector<u32>> islands() {
        const u32 nb = u32(bodies.size());
        std::vector<u32> idx(nb, 0xFFFFFFFFu);
        u32 nd = 0;
        for (u32 i = 0; i < nb; ++i) if (bodies[i].dynamic()) idx[i] = nd++;
        std::vector<std::vector<u32>> out;
        if (nd == 0) return out;
        std::vector<u32> parent(nd), rank(nd, 0);
        for (u32 i = 0; i < nd; ++i) parent[i] = i;
        auto find = [&](u32 x) {
            while (parent[x] != x) { parent[x] = parent[parent[x]]; x = parent[x]; }
            return x;
        };
        for (const Constraint<R>& c : cons) {
            if (!bodies[c.a].dynamic() || !bodies[c.b].dynamic()) continue;
            u32 a = find(idx[c.a]), b = find(idx[c.b]);
            if (a == b) continue;
            if (rank[a] < rank[b]) std::swap(a, b);
            parent[b] = a;
            if (rank[a] == rank[b]) ++rank[a];
        }
        std::vector<u32> root_island(nd, 0xFFFFFFFFu);
        u32 ni = 0;
        for (u32 i = 0; i < nd; ++i) {
            u32 r = find(i);
            if (root_island[r] == 0xFFFFFFFFu) root_island[r] = ni++;
        }
        out.resize(ni);
        auto island_of = [&](int b) -> u32 {
            if (!bodies[b].dynamic()) return 0xFFFFFFFFu;
            return root_island[find(idx[b])];
        };
        for (u32 ci = 0; ci < cons.size(); ++ci) {
            u32 is = island_of(cons[ci].a);
            if (is == 0xFFFFFFFFu) is = island_of(cons[ci].b);
            if (is != 0xFFFFFFFFu) out[is].push_back(ci);
        }
        return out;
    }

    // SolveIsland (PhysicsWorld.cpp:666-743).  The residual bookkeeping
    // (RMS of the impulse increments of each velocity sweep) is the oracle's
    // own addition; it does not change any result.
    void solve_island(const std::vector<u32>& isl, std::vector<double>& sweep_sq, std::vector<u32>& sweep_n) {
        for (u32 ci : isl) warm_start(cons[ci]);
        if (s.use_simd_solver) {
            struct Batch { u32 c[4], k[4]; int a[4], b[4]; u32 n = 0; };
            std::vector<Batch> batches;
            for (u32 ci : isl) {
                Constraint<R>& c = cons[ci];
                if (c.sensor) continue;
                for (u32 k = 0; k < c.n; ++k) {
                    bool placed = false;
                    for (size_t bi = 0; bi < batches.size() && !placed; ++bi) {
                        Batch& b = batches[bi];
                        if (b.n >= 4) continue;
                        bool conflict = false;
                        for (u32 j = 0; j < b.n; ++j)
                            if (b.a[j] == c.a || b.a[j] == c.b || b.b[j] == c.a || b.b[j] == c.b) { conflict = true; break; }
                        if (conflict) continue;
                        b.c[b.n] = ci; b.k[b.n] = k; b.a[b.n] = c.a; b.b[b.n] = c.b; ++b.n;
                        placed = true;
                    }
                    if (!placed) {
                        Batch nb;
                        nb.c[0] = ci; nb.k[0] = k; nb.a[0] = c.a; nb.b[0] = c.b; nb.n = 1;
                        batches.push_back(nb);
                    }
                }
            }
            for (u32 it = 0; it < s.velocity_iterations; ++it) {
                res_acc = 0; res_cnt = 0;
                for (Batch& b : batches)
                    for (u32 j = 0; j < b.n; ++j) solve_lane(cons[b.c[j]], b.k[j]);
                sweep_sq[it] += res_acc; sweep_n[it] += res_cnt;
            }
        } else {
            for (u32 it = 0; it < s.velocity_iterations; ++it) {
                res_acc = 0; res_cnt = 0;
                // The scalar island path has no sensor test (PhysicsWorld.cpp:724-728).
                for (u32 ci : isl)
                    for (u32 k = 0; k < cons[ci].n; ++k) solve_scalar(cons[ci], cons[ci].p[k]);
                sweep_sq[it] += res_acc; sweep_n[it] += res_cnt;
            }
        }
        if (!s.use_velocity_bias)
            for (u32 it = 0; it < s.position_iterations; ++it)
                for (u32 ci : isl) solve_bias(cons[ci]);
    }

    // SolveContacts (PhysicsWorld.cpp:540-582).
    void solve(R dt) {
        for (Constraint<R>& c : cons) prepare(c, dt);
        auto isl = islands();
        std::vector<double> sq(s.velocity_iterations, 0.0);
        std::vector<u32> cnt(s.velocity_iterations, 0);
        for (auto& is : isl) solve_island(is, sq, cnt);
        residuals.assign(s.velocity_iterations, 0.0);
        for (u32 i = 0; i < s.velocity_iterations; ++i) residuals[i] = cnt[i] ? std::sqrt(sq[i] / cnt[i]) : 0.0;
        for (const Constraint<R>& c : cons) {
            CacheEntry<R>& e = cache[pair_key(u32(c.a), u32(c.b))];
            e.n = c.n;
            for (u32 k = 0; k < c.n; ++k) {
                e.pos[k] = c.p[k].pa; e.ln[k] = c.p[k].ln; e.lt1[k] = c.p[k].lt1; e.lt2[k] = c.p[k].lt2;
            }
        }
    }

    // IntegratePositions (:785-802), pseudo-velocity integration (:275-289),
    // UpdateBodyInertias (:854-861), force clear (:312).
    void finish(R dt) {
        for (Body<R>& b : bodies) {
            if (b.is_static()) continue;
            b.xf.t = b.xf.t + b.v * dt;
            if (len2(b.w) > R(0)) b.xf.q = integrate_rotation(b.xf.q, b.w, dt);
        }
        if (!s.use_velocity_bias) {
            for (Body<R>& b : bodies) {
                if (b.is_static()) continue;
                if (len2(b.pv) > R(0)) b.xf.t = b.xf.t + b.pv * dt;
                if (len2(b.pw) > R(0)) b.xf.q = integrate_rotation(b.xf.q, b.pw, dt);
                b.pv = V3<R>{}; b.pw = V3<R>{};
            }
        }
        for (Body<R>& b : bodies)
            if (b.dynamic()) b.inv_i_world = world_inertia(b.xf.q, b.inv_i_local);
        for (Body<R>& b : bodies) { b.force = V3<R>{}; b.torque = V3<R>{}; }
    }

    void step(R dt, int stop_after) {
        narrowphase();
        if (stop_after >= 2) build();
        if (stop_after >= 3) integrate_velocities(dt);
        if (stop_after >= 4) solve(dt);
        if (stop_after >= 5) finish(dt);
    }
};

template <class R> Shape<R> make_shape(const kz_shape_desc& d, const kzo_hull_ref* hulls, u32 index,
                                       const float* verts, const uint32_t* tris) {
    Shape<R> s;
    s.type = int(d.type);
    if (d.type == SPHERE) s.radius = d.params[0];
    if (d.type == BOX) s.half = V3<R>{d.params[0], d.params[1], d.params[2]};
    if (d.type == CAPSULE) { s.half_height = d.params[0]; s.radius = d.params[1]; }
    if (d.type == PLANE) { s.pn = V3<R>{d.params[0], d.params[1], d.params[2]}; s.pd = d.params[3]; }
    if (d.type == HULL) {
        const kzo_hull_ref& h = hulls[index];
        for (u32 i = 0; i < h.vert_count; ++i) {
            const float* p = verts + 3 * (h.vert_offset + i);
            s.verts.push_back(V3<R>{p[0], p[1], p[2]});
        }
        s.tris.assign(tris + h.tri_offset * 3, tris + (h.tri_offset + h.tri_count) * 3);
    }
    return s;
}

template <class R> void to_contact(kz_contact_point& o, const CP<R>& p) {
    o.point_on_a[0] = float(p.pa.x); o.point_on_a[1] = float(p.pa.y); o.point_on_a[2] = float(p.pa.z);
    o.point_on_b[0] = float(p.pb.x); o.point_on_b[1] = float(p.pb.y); o.point_on_b[2] = float(p.pb.z);
    o.normal[0] = float(p.n.x); o.normal[1] = float(p.n.y); o.normal[2] = float(p.n.z);
    o.separation = float(p.sep);
    o.feature_id = p.feature;
}

template <class R> Xf<R> load_xf(const float* x) {
    Xf<R> t;
    t.q = Q4<R>{x[0], x[1], x[2], x[3]};
    t.t = V3<R>{x[4], x[5], x[6]};
    return t;
}

struct Handle {
    int precision = 0;
    World<float>* f = nullptr;
    World<double>* d = nullptr;
};

template <class R> World<R>* make_world(const kz_world_desc* wd, const kz_shape_desc* shapes, u32 ns,
                                        const kzo_hull_ref* hulls, const float* hv, const uint32_t* ht,
                                        const kz_body_desc* bodies, u32 nb, const float* mass) {
    auto* w = new World<R>();
    w->s = *wd;
    for (u32 i = 0; i < ns; ++i) w->shapes.push_back(make_shape<R>(shapes[i], hulls, i, hv, ht));
    for (u32 i = 0; i < nb; ++i) {
        const kz_body_desc& d = bodies[i];
        Body<R> b;
        b.shape = int(d.shape);
        b.motion = int(d.motion_type);
        b.sensor = d.is_sensor != 0;
        b.xf.t = V3<R>{d.position[0], d.position[1], d.position[2]};
        b.xf.q = Q4<R>{d.rotation[0], d.rotation[1], d.rotation[2], d.rotation[3]};
        b.v = V3<R>{d.linear_velocity[0], d.linear_velocity[1], d.linear_velocity[2]};
        b.w = V3<R>{d.angular_velocity[0], d.angular_velocity[1], d.angular_velocity[2]};
        b.friction = d.friction; b.restitution = d.restitution;
        b.lin_damp = d.linear_damping; b.ang_damp = d.angular_damping; b.gravity_factor = d.gravity_factor;
        b.max_lin = d.max_linear_velocity; b.max_ang = d.max_angular_velocity;
        const float* m = mass + 10 * i;
        b.inv_mass = m[0];
        for (int c = 0; c < 3; ++c) b.inv_i_local.c[c] = V3<R>{m[1 + 3 * c], m[2 + 3 * c], m[3 + 3 * c]};
        b.inv_i_world = b.dynamic() ? world_inertia(b.xf.q, b.inv_i_local) : M3<R>{};
        w->bodies.push_back(b);
    }
    return w;
}

template <class R> void dump_constraints(const World<R>& w, uint32_t* hdr, float* pts) {
    for (size_t c = 0; c < w.cons.size(); ++c) {
        const Constraint<R>& cc = w.cons[c];
        hdr[4 * c] = u32(cc.a); hdr[4 * c + 1] = u32(cc.b); hdr[4 * c + 2] = cc.n;
        hdr[4 * c + 3] = (cc.sensor ? 1u : 0u) | (cc.dyn_a ? 2u : 0u) | (cc.dyn_b ? 4u : 0u);
        for (u32 k = 0; k < 8; ++k) {
            float* o = pts + (8 * c + k) * 44;
            const PointC<R>& p = cc.p[k];
            auto v3 = [&](int at, V3<R> v) { o[at] = float(v.x); o[at + 1] = float(v.y); o[at + 2] = float(v.z); };
            v3(0, p.pa); v3(3, p.pb); v3(6, p.ra); v3(9, p.rb); v3(12, p.la); v3(15, p.lb);
            v3(18, p.n); v3(21, p.t1); v3(24, p.t2);
            o[27] = float(p.nm); o[28] = float(p.tm1); o[29] = float(p.tm2);
            o[30] = float(p.ln); o[31] = float(p.lt1); o[32] = float(p.lt2);
            o[33] = float(p.sep); o[34] = float(p.rbias); o[35] = float(p.pbias);
            o[36] = float(p.bimp); o[37] = float(p.cf); o[38] = float(p.cr); o[39] = float(p.nvel);
            std::memcpy(&o[40], &p.feature, 4);
            o[41] = o[42] = o[43] = 0.0f;
        }
    }
}

template <class R> void collide_range(const std::vector<Shape<R>>& table, const uint32_t* sa, const uint32_t* sb,
                                      const float* xa, const float* xb, u32 lo, u32 hi, R max_sep,
                                      uint32_t* counts, kz_contact_point* out) {
    for (u32 i = lo; i < hi; ++i) {
        Manifold<R> m;
        collide(table[sa[i]], load_xf<R>(xa + 7 * size_t(i)), table[sb[i]], load_xf<R>(xb + 7 * size_t(i)), max_sep, m);
        counts[i] = m.n;
        for (u32 k = 0; k < 8; ++k) to_contact(out[8 * size_t(i) + k], m.pt[k]);
    }
}

template <class R> void collide_batch(const kz_shape_desc* shapes, u32 ns, const kzo_hull_ref* hulls, const float* hv,
                                      const uint32_t* ht, const uint32_t* sa, const uint32_t* sb, const float* xa,
                                      const float* xb, u32 n, float max_sep, uint32_t* counts,
                                      kz_contact_point* out, int threads) {
    std::vector<Shape<R>> table;
    for (u32 i = 0; i < ns; ++i) table.push_back(make_shape<R>(shapes[i], hulls, i, hv, ht));
    if (threads <= 1) {
        collide_range(table, sa, sb, xa, xb, 0, n, R(max_sep), counts, out);
        return;
    }
    std::vector<std::thread> pool;
    u32 per = (n + u32(threads) - 1) / u32(threads);
    for (int t = 0; t < threads; ++t) {
        u32 lo = u32(t) * per, hi = std::min(n, lo + per);
        if (lo >= hi) break;
        pool.emplace_back([&, lo, hi] { collide_range(table, sa, sb, xa, xb, lo, hi, R(max_sep), counts, out); });
    }
    for (auto& th : pool) th.join();
}

} // namespace kzo

using namespace kzo;

extern "C" {

void* kzo_world_create(int precision, const kz_world_desc* world, const kz_shape_desc* shapes, uint32_t num_shapes,
                       const kzo_hull_ref* hulls, const float* hull_verts, const uint32_t* hull_tris,
                       const kz_body_desc* bodies, uint32_t num_bodies, const float* mass) {
    auto* h = new Handle();
    h->precision = precision;
    if (precision == 0)
        h->f = make_world<float>(world, shapes, num_shapes, hulls, hull_verts, hull_tris, bodies, num_bodies, mass);
    else
        h->d = make_world<double>(world, shapes, num_shapes, hulls, hull_verts, hull_tris, bodies, num_bodies, mass);
    return h;
}

void kzo_world_destroy(void* p) {
    auto* h = static_cast<Handle*>(p);
    delete h->f;
    delete h->d;
    delete h;
}

void kzo_set_warm_start(void* p, int mode) {
    auto* h = static_cast<Handle*>(p);
    if (h->f) h->f->warm_mode = mode;
    if (h->d) h->d->warm_mode = mode;
}

void kzo_step(void* p, float dt, const uint32_t* pairs, uint32_t num_pairs, int stop_after) {
    auto* h = static_cast<Handle*>(p);
    auto run = [&](auto* w, auto rdt) {
        w->pairs.resize(num_pairs);
        for (uint32_t i = 0; i < num_pairs; ++i) w->pairs[i] = {pairs[2 * i], pairs[2 * i + 1]};
        w->step(rdt, stop_after);
    };
    if (h->f) run(h->f, float(dt));
    else run(h->d, double(dt));
}

uint32_t kzo_num_constraints(void* p) {
    auto* h = static_cast<Handle*>(p);
    return uint32_t(h->f ? h->f->cons.size() : h->d->cons.size());
}

void kzo_get_states(void* p, kz_body_state* out) {
    auto* h = static_cast<Handle*>(p);
    auto dump = [&](auto* w) {
        for (size_t i = 0; i < w->bodies.size(); ++i) {
            const auto& b = w->bodies[i];
            kz_body_state& s = out[i];
            s.position[0] = float(b.xf.t.x); s.position[1] = float(b.xf.t.y); s.position[2] = float(b.xf.t.z);
            s.rotation[0] = float(b.xf.q.x); s.rotation[1] = float(b.xf.q.y);
            s.rotation[2] = float(b.xf.q.z); s.rotation[3] = float(b.xf.q.w);
            s.linear_velocity[0] = float(b.v.x); s.linear_velocity[1] = float(b.v.y); s.linear_velocity[2] = float(b.v.z);
            s.angular_velocity[0] = float(b.w.x); s.angular_velocity[1] = float(b.w.y); s.angular_velocity[2] = float(b.w.z);
        }
    };
    if (h->f) dump(h->f); else dump(h->d);
}

void kzo_get_states_f64(void* p, double* out) {
    auto* h = static_cast<Handle*>(p);
    auto dump = [&](auto* w) {
        for (size_t i = 0; i < w->bodies.size(); ++i) {
            const auto& b = w->bodies[i];
            double* o = out + 13 * i;
            o[0] = b.xf.t.x; o[1] = b.xf.t.y; o[2] = b.xf.t.z;
            o[3] = b.xf.q.x; o[4] = b.xf.q.y; o[5] = b.xf.q.z; o[6] = b.xf.q.w;
            o[7] = b.v.x; o[8] = b.v.y; o[9] = b.v.z; o[10] = b.w.x; o[11] = b.w.y; o[12] = b.w.z;
        }
    };
    if (h->f) dump(h->f); else dump(h->d);
}

void kzo_set_states(void* p, const kz_body_state* in) {
    auto* h = static_cast<Handle*>(p);
    auto load = [&](auto* w) {
        using R = typename std::remove_reference_t<decltype(w->bodies[0].inv_mass)>;
        for (size_t i = 0; i < w->bodies.size(); ++i) {
            auto& b = w->bodies[i];
            const kz_body_state& s = in[i];
            b.xf.t = V3<R>{s.position[0], s.position[1], s.position[2]};
            b.xf.q = Q4<R>{s.rotation[0], s.rotation[1], s.rotation[2], s.rotation[3]};
            b.v = V3<R>{s.linear_velocity[0], s.linear_velocity[1], s.linear_velocity[2]};
            b.w = V3<R>{s.angular_velocity[0], s.angular_velocity[1], s.angular_velocity[2]};
            if (b.dynamic()) b.inv_i_world = world_inertia(b.xf.q, b.inv_i_local);
        }
    };
    if (h->f) load(h->f); else load(h->d);
}

void kzo_get_manifolds(void* p, uint32_t* counts, kz_contact_point* points, uint64_t* keys) {
    auto* h = static_cast<Handle*>(p);
    auto dump = [&](auto* w) {
        for (size_t i = 0; i < w->manifolds.size(); ++i) {
            counts[i] = w->manifolds[i].n;
            keys[i] = w->keys[i];
            for (u32 k = 0; k < 8; ++k) to_contact(points[8 * i + k], w->manifolds[i].pt[k]);
        }
    };
    if (h->f) dump(h->f); else dump(h->d);
}

void kzo_get_constraints(void* p, uint32_t* hdr, float* pts) {
    auto* h = static_cast<Handle*>(p);
    if (h->f) dump_constraints(*h->f, hdr, pts); else dump_constraints(*h->d, hdr, pts);
}

uint32_t kzo_get_residuals(void* p, double* out, uint32_t cap) {
    auto* h = static_cast<Handle*>(p);
    const std::vector<double>& r = h->f ? h->f->residuals : h->d->residuals;
    uint32_t n = uint32_t(std::min<size_t>(cap, r.size()));
    for (uint32_t i = 0; i < n; ++i) out[i] = r[i];
    return uint32_t(r.size());
}

void kzo_collide_batch(int precision, const kz_shape_desc* shapes, uint32_t num_shapes, const kzo_hull_ref* hulls,
                       const float* hull_verts, const uint32_t* hull_tris, const uint32_t* shape_a,
                       const uint32_t* shape_b, const float* xf_a, const float* xf_b, uint32_t n,
                       float max_separation, uint32_t* counts, kz_contact_point* points, int threads) {
    if (precision == 0)
        collide_batch<float>(shapes, num_shapes, hulls, hull_verts, hull_tris, shape_a, shape_b, xf_a, xf_b, n,
                             max_separation, counts, points, threads);
    else
        collide_batch<double>(shapes, num_shapes, hulls, hull_verts, hull_tris, shape_a, shape_b, xf_a, xf_b, n,
                              max_separation, counts, points, threads);
}

} // extern "C"
