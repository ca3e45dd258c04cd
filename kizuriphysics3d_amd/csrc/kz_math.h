// kz_math.h - float vector/quaternion/matrix math shared by the HIP kernels and
// the host library.
//
// Every operation keeps the reference engine's operation order and
// association (include/Kizuri/Core/Math.h) so that results are bit-identical
// to the reference's FP32 build: the library is compiled with
// -ffp-contract=off (no a*b+c fusion), correctly rounded division and square
// root, and no fast-math.  Notable conventions kept on purpose:
//   * vector / scalar is a multiplication by the reciprocal (Math.h:39,45);
//   * dot products sum left to right, x*x' + y*y' + z*z';
//   * Quat::Rotate is v + w t + q x t with t = 2 (q x v) (Math.h:296-307);
//   * Mat3 is column-major and M v = c0 x + c1 y + c2 z (Math.h:216).
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#define KZ_HD __host__ __device__ __forceinline__

namespace kz {

struct f3 {
    float x, y, z;
};
struct f4q {  // quaternion x, y, z, w
    float x, y, z, w;
};
struct m3 {  // column-major
    f3 c0, c1, c2;
};

constexpr float kEps = 1.1920928955078125e-7f;  // FLT_EPSILON (Types.h:43)
constexpr float kInf = __builtin_huge_valf();

KZ_HD f3 mk(float x, float y, float z) { return f3{x, y, z}; }
KZ_HD f3 add(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
KZ_HD f3 sub(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
KZ_HD f3 neg(f3 a) { return f3{-a.x, -a.y, -a.z}; }
KZ_HD f3 mul(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
KZ_HD f3 divs(f3 a, float s) {
    float inv = 1.0f / s;
    return f3{a.x * inv, a.y * inv, a.z * inv};
}
KZ_HD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
KZ_HD f3 cross(f3 a, f3 b) { return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
KZ_HD float len2(f3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
KZ_HD float len(f3 a) { return sqrtf(len2(a)); }
KZ_HD f3 normalized_or(f3 a, f3 fb) {
    float l = len(a);
    return l >= kEps ? divs(a, l) : fb;
}
KZ_HD float fabs_(float v) { return v < 0.0f ? -v : v; }
KZ_HD float fmin_(float a, float b) { return a < b ? a : b; }  // math::Min
KZ_HD float fmax_(float a, float b) { return a > b ? a : b; }  // math::Max
KZ_HD float fclamp(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }
KZ_HD f3 vabs(f3 a) { return f3{fabs_(a.x), fabs_(a.y), fabs_(a.z)}; }
KZ_HD int max_index(f3 a) { return a.x >= a.y ? (a.x >= a.z ? 0 : 2) : (a.y >= a.z ? 1 : 2); }
KZ_HD float comp(f3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
KZ_HD void set_comp(f3& a, int i, float v) {
    if (i == 0) a.x = v;
    else if (i == 1) a.y = v;
    else a.z = v;
}
// Vec3::GetNormalizedPerpendicular (Math.h:101-110), true divisions.
KZ_HD f3 perpendicular(f3 n) {
    if (fabs_(n.x) > fabs_(n.y)) {
        float l = sqrtf(n.x * n.x + n.z * n.z);
        return f3{-n.z / l, 0.0f, n.x / l};
    }
    float l = sqrtf(n.y * n.y + n.z * n.z);
    return f3{0.0f, n.z / l, -n.y / l};
}

KZ_HD f4q qmul(f4q a, f4q q) {
    return f4q{a.w * q.x + a.x * q.w + a.y * q.z - a.z * q.y, a.w * q.y - a.x * q.z + a.y * q.w + a.z * q.x,
               a.w * q.z + a.x * q.y - a.y * q.x + a.z * q.w, a.w * q.w - a.x * q.x - a.y * q.y - a.z * q.z};
}
KZ_HD f4q qnormalized(f4q q) {
    float l = sqrtf(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
    if (!(l >= kEps)) return f4q{0.0f, 0.0f, 0.0f, 1.0f};
    float s = 1.0f / l;
    return f4q{q.x * s, q.y * s, q.z * s, q.w * s};
}
KZ_HD f3 rotate(f4q q, f3 v) {
    f3 qv{q.x, q.y, q.z};
    f3 t = mul(cross(qv, v), 2.0f);
    return add(add(v, mul(t, q.w)), cross(qv, t));
}
KZ_HD f3 inv_rotate(f4q q, f3 v) {
    f3 qv{-q.x, -q.y, -q.z};
    f3 t = mul(cross(qv, v), 2.0f);
    return add(add(v, mul(t, q.w)), cross(qv, t));
}
// q' = (q + (0.5 dt w, 0) q) normalised (src/World/PhysicsWorld.cpp:31-35).
KZ_HD f4q integrate_rotation(f4q q, f3 w, float dt) {
    f4q dq{0.5f * dt * w.x, 0.5f * dt * w.y, 0.5f * dt * w.z, 0.0f};
    f4q p = qmul(dq, q);
    return qnormalized(f4q{q.x + p.x, q.y + p.y, q.z + p.z, q.w + p.w});
}
KZ_HD f3 xform(f4q q, f3 t, f3 p) { return add(rotate(q, p), t); }

KZ_HD f3 mmul(const m3& m, f3 v) { return add(add(mul(m.c0, v.x), mul(m.c1, v.y)), mul(m.c2, v.z)); }
KZ_HD m3 mmul(const m3& a, const m3& b) { return m3{mmul(a, b.c0), mmul(a, b.c1), mmul(a, b.c2)}; }
KZ_HD m3 transposed(const m3& m) {
    return m3{f3{m.c0.x, m.c1.x, m.c2.x}, f3{m.c0.y, m.c1.y, m.c2.y}, f3{m.c0.z, m.c1.z, m.c2.z}};
}
// Quat::ToMat3 (Math.h:316-324).
KZ_HD m3 to_mat3(f4q q) {
    float xx = q.x * q.x, yy = q.y * q.y, zz = q.z * q.z;
    float xy = q.x * q.y, xz = q.x * q.z, yz = q.y * q.z;
    float wx = q.w * q.x, wy = q.w * q.y, wz = q.w * q.z;
    return m3{f3{1.0f - 2.0f * (yy + zz), 2.0f * (xy + wz), 2.0f * (xz - wy)},
              f3{2.0f * (xy - wz), 1.0f - 2.0f * (xx + zz), 2.0f * (yz + wx)},
              f3{2.0f * (xz + wy), 2.0f * (yz - wx), 1.0f - 2.0f * (xx + yy)}};
}
// MotionProperties::UpdateWorldInertia (include/Kizuri/Body/Body.h:122-125).
KZ_HD m3 world_inertia(f4q q, const m3& local) {
    m3 r = to_mat3(q);
    return mmul(mmul(r, local), transposed(r));
}

}  // namespace kz
