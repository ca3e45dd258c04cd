// solver.hip - constraint build, velocity integration, Prepare, warm start,
// projected Gauss-Seidel velocity and split-impulse bias sweeps, cache store
// and position integration, hand-written for gfx950.
//
// Bit-exactness argument.  The reference solves each island sequentially:
// warm start over the island's contacts in constraint order, then
// `velocityIterations` passes over the first-fit SIMD batches, then
// `positionIterations` bias passes in constraint order
// (src/World/PhysicsWorld.cpp:666-743).  The per-point arithmetic reads and
// writes only the two bodies of that point; static and kinematic bodies are
// never written (src/Solver/ContactSolver.cpp:32-43,
// ContactSolverSIMD.cpp:241-256).  Hence any schedule that applies, for every
// DYNAMIC body, the point updates touching it in the reference's order gives
// bit-identical results.  The host turns each sequence into levels
// (level = 1 + max over the point's dynamic bodies of their last level); the
// points of one level share no dynamic body and run in parallel, levels run
// in order.  Islands are independent, so the islands of one group share
// level numbers and one persistent workgroup sweeps a group.
#include <type_traits>

#include "kz_device.h"
#include "kz_math.h"

namespace kz {

__device__ __forceinline__ f3 ld3(const float4& v) { return f3{v.x, v.y, v.z}; }
__device__ __forceinline__ f4q ldq(const float4& v) { return f4q{v.x, v.y, v.z, v.w}; }
__device__ __forceinline__ float4 st3(f3 a, float w) { return make_float4(a.x, a.y, a.z, w); }
__device__ __forceinline__ m3 ldm(const float4* c) { return m3{ld3(c[0]), ld3(c[1]), ld3(c[2])}; }

// ---------------------------------------------------------------------------
// IntegrateVelocities (PhysicsWorld.cpp:757-783), one thread per body.
// ---------------------------------------------------------------------------
__global__ void k_integrate_velocities(BodyBuf b, uint32_t n, float dt, float gx, float gy, float gz) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t m = b.meta[i];
    if (meta_motion(m) != 2u || meta_asleep(m)) return;  // the active list's dynamic bodies
    const float4 x = b.x[i], lim = b.lim[i], mat = b.mat[i];
    f3 v = ld3(b.v[i]), w = ld3(b.w[i]);
    const m3 iw = ldm(b.iw + 3 * i);
    const f3 zero{0.0f, 0.0f, 0.0f};  // accumulated force / torque: none on this path
    f3 acc = add(mul(f3{gx, gy, gz}, lim.x), mul(zero, x.w));
    v = add(v, mul(acc, dt));
    w = add(w, mul(mmul(iw, zero), dt));
    v = mul(v, 1.0f / (1.0f + mat.z * dt));
    w = mul(w, 1.0f / (1.0f + mat.w * dt));
    auto clamp_len = [](f3 a, float mx) {
        float l = len(a);
        return (l > mx && l >= kEps) ? mul(a, mx / l) : a;
    };
    v = clamp_len(v, lim.y);
    w = clamp_len(w, lim.z);
    if (!(isfinite(v.x) && isfinite(v.y) && isfinite(v.z)) || !(isfinite(w.x) && isfinite(w.y) && isfinite(w.z))) {
        v = zero;
        w = zero;
    }
    b.v[i] = st3(v, 0.0f);
    b.w[i] = st3(w, 0.0f);
}

// ---------------------------------------------------------------------------
// BuildContactConstraints (PhysicsWorld.cpp:454-535), one thread per
// constraint.  The host has already compacted the non-empty manifolds in pair
// order and ordered each constraint's bodies by packed id (info).
// ---------------------------------------------------------------------------
__global__ void k_build(ContactBuf c, ManifoldBuf m, BodyBuf b, CacheBuf cache, uint32_t nc, const uint32_t* nc_dev,
                        int warm) {
    uint32_t ci = blockIdx.x * blockDim.x + threadIdx.x;
    if (nc_dev) nc = *nc_dev;
    if (ci >= nc) return;
    const uint4 info = c.info[ci];
    const uint32_t pair = c.pair[ci];
    const uint32_t first = info.z, np = info.w & 0xFFu, flags = info.w >> 8;
    const float4 ma = b.mat[info.x], mb = b.mat[info.y];
    const float friction = sqrtf(ma.x * mb.x);
    const float restitution = fmax_(ma.y, mb.y);
    int slot = -1;
    if (warm == 0) {
        const unsigned long long key = pair_key(b.packed[info.x], b.packed[info.y]);
        uint32_t mask = cache.cap - 1u, s = cache_hash(key) & mask;
        for (uint32_t probe = 0; probe < cache.cap; ++probe) {
            unsigned long long k = cache.key[s];
            if (k == key) { slot = int(s); break; }
            if (k == 0ull) break;
            s = (s + 1u) & mask;
        }
    }
    uint32_t ncached = slot >= 0 ? cache.count[slot] : 0u;
    for (uint32_t k = 0; k < np; ++k) {
        const uint32_t p = first + k;
        const float4 pa = m.pa[pair * kPairPoints + k];
        const float4 pb = m.pb[pair * kPairPoints + k];
        const float4 nn = m.n[pair * kPairPoints + k];
        float4 imp = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (slot >= 0) {
            // Nearest cached point within 0.1 m, strict <, first wins.
            float best = 0.10f * 0.10f;
            int bi = -1;
            const f3 src = ld3(pa);
            for (uint32_t j = 0; j < ncached; ++j) {
                const float4 cp = cache.p[slot * kPairPoints + j];
                float d = len2(sub(ld3(cp), src));
                if (d < best) { best = d; bi = int(j); }
            }
            if (bi >= 0) {
                // w = 1 marks a cache hit for Prepare's feature row; Prepare
                // resets it to the bias impulse 0.
                const float4 cp = cache.p[slot * kPairPoints + bi];
                const float2 ct = cache.t[slot * kPairPoints + bi];
                imp = make_float4(cp.w, ct.x, ct.y, 1.0f);
            }
        }
        c.body[p] = make_uint4(info.x, info.y, flags, ci);
        c.pa[p] = pa;
        c.pb[p] = pb;
        c.n[p] = make_float4(nn.x, nn.y, nn.z, 0.0f);
        c.la[p] = make_float4(0.0f, 0.0f, 0.0f, friction);
        c.lb[p] = make_float4(0.0f, 0.0f, 0.0f, restitution);
        c.imp[p] = imp;
    }
}

// ---------------------------------------------------------------------------
// PrepareConstraint (ContactSolver.cpp:56-112), one thread per point.
// ---------------------------------------------------------------------------
struct PrepParams {
    float dt, restitution_threshold, baumgarte, slop, max_bias;
    int use_simd, use_velocity_bias;
    float gdt;      // |gravity| * dt (or dt without gravity): the impulse scale per unit normal mass
    float4* feat;   // optional: learned warm-start feature rows, 4 float4 per point
    float4* base;   // with feat: the clamped cache impulses (zero on a miss) and the impulse scale, 1 float4 per point
    const unsigned long long* load;  // optional, with feat: per-body loads (k_body_load), else features 14-15 are 0
};

// Load features of the learned warm start: for every contact point, its
// cached normal impulse clamped at 0 (as Prepare clamps it), over each
// dynamic body's weight impulse m |g| dt, summed per body in 44.20 fixed
// point (integer sums: the same in any order).  acc[nbodies] cleared first.
constexpr double kLoadScale = 1048576.0;
__global__ void k_body_load(ContactBuf c, BodyBuf b, uint32_t np, float gdt, unsigned long long* acc) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const uint4 bd = c.body[p];
    const float ln = fmax_(c.imp[p].x, 0.0f);
    if (!(ln > 0.0f)) return;
    const uint32_t xs[2] = {bd.x, bd.y};
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const uint32_t x = xs[k];
        if (meta_motion(b.meta[x]) != 2u) continue;
        const double w = double(ln) * double(b.x[x].w) / double(gdt);
        if (w > 0.0 && w < 1.0e12) atomicAdd(acc + x, (unsigned long long)(w * kLoadScale + 0.5));
    }
}
__global__ void k_fill_u64(unsigned long long* p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}
__device__ __forceinline__ float body_load(const unsigned long long* acc, uint32_t x) {
    return acc ? float(double(acc[x]) / kLoadScale) : 0.0f;
}
__device__ __forceinline__ float proj_inertia(const m3& inv_i, f3 r, f3 d) {
    const f3 x = cross(r, d);
    return dot(x, mmul(inv_i, x));
}
__device__ __forceinline__ float eff_mass(float ima, float imb, const m3& ia, const m3& ib, f3 ra, f3 rb, f3 d) {
    f3 xa = cross(ra, d), xb = cross(rb, d);
    float k = ima + imb;
    k = k + dot(xa, mmul(ia, xa));
    k = k + dot(xb, mmul(ib, xb));
    return k > kEps ? 1.0f / k : 0.0f;
}
__global__ void k_prepare(ContactBuf c, BodyBuf b, uint32_t np, const uint32_t* np_dev, PrepParams pp, VelRecBuf vr,
                          WsRecBuf wr, const uint32_t* vel_pos, const uint32_t* ws_pos) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (np_dev) np = *np_dev;
    if (p >= np) return;
    const uint4 bd = c.body[p];
    const float4 xa = b.x[bd.x], xb = b.x[bd.y];
    const f4q qa = ldq(b.q[bd.x]), qb = ldq(b.q[bd.y]);
    const m3 ia = ldm(b.iw + 3 * bd.x), ib = ldm(b.iw + 3 * bd.y);
    const f3 va = ld3(b.v[bd.x]), wa = ld3(b.w[bd.x]), vb = ld3(b.v[bd.y]), wb = ld3(b.w[bd.y]);
    const float4 pa4 = c.pa[p], pb4 = c.pb[p];
    const f3 pa = ld3(pa4), pb = ld3(pb4);
    const float sep = pa4.w;
    const float cf = c.la[p].w, cr = c.lb[p].w;
    float4 imp = c.imp[p];
    const f3 posa = ld3(xa), posb = ld3(xb);
    const f3 ra = sub(pa, posa), rb = sub(pb, posb);
    const f3 la = inv_rotate(qa, ra), lb = inv_rotate(qb, rb);
    const f3 n = normalized_or(ld3(c.n[p]), f3{0.0f, 1.0f, 0.0f});
    const f3 t1 = perpendicular(n);
    const f3 t2 = cross(n, t1);
    const float nm = eff_mass(xa.w, xb.w, ia, ib, ra, rb, n);
    const float tm1 = eff_mass(xa.w, xb.w, ia, ib, ra, rb, t1);
    const float tm2 = eff_mass(xa.w, xb.w, ia, ib, ra, rb, t2);
    const f3 pva = add(va, cross(wa, sub(pa, posa)));
    const f3 pvb = add(vb, cross(wb, sub(pb, posb)));
    const float nvel = dot(sub(pva, pvb), n);
    float rbias = 0.0f;
    if (nvel < -pp.restitution_threshold) rbias = -cr * nvel;
    float pbias = 0.0f;
    if (sep < -pp.slop) {
        pbias = pp.baumgarte / pp.dt * (-sep - pp.slop);
        pbias = fmin_(pbias, pp.max_bias);
    }
    if (pp.feat) {
        // Feature row of the learned warm start (layout in mlp.h).
        const f3 vrel = sub(pva, pvb);
        float4* row = pp.feat + 4 * size_t(p);
        row[0] = make_float4(nvel, dot(vrel, t1), dot(vrel, t2), n.x);
        row[1] = make_float4(n.y, n.z, -sep, xa.w);
        row[2] = make_float4(xb.w, proj_inertia(ia, ra, n), proj_inertia(ib, rb, n),
                             0.5f * (proj_inertia(ia, ra, t1) + proj_inertia(ia, ra, t2)));
        row[3] = make_float4(0.5f * (proj_inertia(ib, rb, t1) + proj_inertia(ib, rb, t2)), imp.w,
                             body_load(pp.load, bd.x), body_load(pp.load, bd.y));
    }
    const float mf = cf * imp.x;
    imp.y = fclamp(imp.y, -mf, mf);
    imp.z = fclamp(imp.z, -mf, mf);
    imp.x = fmax_(imp.x, 0.0f);
    imp.w = 0.0f;
    if (pp.feat) pp.base[p] = make_float4(imp.x, imp.y, imp.z, nm * pp.gdt);
    uint4 bdo = bd;
    const uint32_t ma = meta_motion(b.meta[bd.x]), mb = meta_motion(b.meta[bd.y]);
    bdo.z = (bd.z & 1u) | (ma == 2u ? 2u : 0u) | (mb == 2u ? 4u : 0u);
    c.body[p] = bdo;
    c.n[p] = st3(n, nm);
    c.t1[p] = st3(t1, tm1);
    c.t2[p] = st3(t2, tm2);
    c.ra[p] = st3(ra, rbias);
    c.rb[p] = st3(rb, pbias);
    c.la[p] = st3(la, cf);
    c.lb[p] = st3(lb, cr);
    c.imp[p] = imp;
    c.nvel[p] = nvel;
    // Sweep records.  Body rotations do not change during the solve, so the
    // anchors the reference rotates on every visit (ContactSolverSIMD.cpp:
    // 106-107, ContactSolver.cpp:268-269) are rotated once here: same values.
    const f3 qra = rotate(qa, la), qrb = rotate(qb, lb);
    const uint32_t e = vel_pos ? vel_pos[p] : 0xFFFFFFFFu;  // none: k_velocity_records writes them later
    if (e != 0xFFFFFFFFu) {
        // SIMD lanes add (positionBias or 0) to the restitution bias; the
        // scalar path adds positionBias only when enabled (ContactSolverSIMD.cpp:128,
        // ContactSolver.cpp:169-170).
        const float nbias = pp.use_simd ? rbias + (pp.use_velocity_bias ? pbias : 0.0f)
                                        : (pp.use_velocity_bias ? rbias + pbias : rbias);
        vr.bd[e] = bdo;
        vr.ra[e] = st3(qra, xa.w);
        vr.rb[e] = st3(qrb, xb.w);
        vr.n[e] = st3(n, nm);
        vr.t1[e] = st3(t1, tm1);
        vr.t2[e] = st3(t2, tm2);
        vr.fb[e] = make_float2(cf, nbias);
        vr.imp[e] = make_float4(imp.x, imp.y, imp.z, 0.0f);
    }
    const uint32_t w = ws_pos[p];
    if (w == 0xFFFFFFFFu) return;  // in no island: solved nowhere, keeps the clamped impulses
    wr.bd[w] = bdo;
    wr.ra[w] = st3(ra, xa.w);
    wr.rb[w] = st3(rb, xb.w);
    wr.qa[w] = st3(qra, pbias);
    wr.qb[w] = st3(qrb, 0.0f);
    wr.n[w] = st3(n, nm);
    wr.t1[w] = st3(t1, 0.0f);
    wr.t2[w] = st3(t2, 0.0f);
    wr.imp[w] = imp;
}

// The velocity records of a prepare launch that skipped them, from what it
// wrote: the constraint table (normal and tangents with their masses, the
// clamped impulses, friction, restitution and position biases) and the
// warm-start records (rotated anchors, inverse masses).  Every field is the
// same value k_prepare stores, so the records are identical.
__global__ void k_velocity_records(ContactBuf c, WsRecBuf wr, VelRecBuf vr, const uint32_t* vel_pos,
                                   const uint32_t* ws_pos, uint32_t np, int use_simd, int use_vbias) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const uint32_t e = vel_pos[p];
    if (e == 0xFFFFFFFFu) return;
    const uint32_t w = ws_pos[p];
    const float4 ra = wr.ra[w], rb = wr.rb[w], qa = wr.qa[w], qb = wr.qb[w], imp = c.imp[p];
    const float rbias = c.ra[p].w, pbias = c.rb[p].w;
    const float nbias = use_simd ? rbias + (use_vbias ? pbias : 0.0f) : (use_vbias ? rbias + pbias : rbias);
    vr.bd[e] = c.body[p];
    vr.ra[e] = make_float4(qa.x, qa.y, qa.z, ra.w);
    vr.rb[e] = make_float4(qb.x, qb.y, qb.z, rb.w);
    vr.n[e] = c.n[p];
    vr.t1[e] = c.t1[p];
    vr.t2[e] = c.t2[p];
    vr.fb[e] = make_float2(c.la[p].w, nbias);
    vr.imp[e] = make_float4(imp.x, imp.y, imp.z, 0.0f);
}

// ---------------------------------------------------------------------------
// Solver bodies.  During the sweeps the state a point reads of a body lives in
// one 64-byte line per body (instead of five separate arrays): velocities,
// or pseudo velocities, and the inverse world inertia columns c0, c1, c2:
//   (vx, vy, vz, wx) (wy, wz, c0x, c0y) (c0z, c1x, c1y, c1z) (c2x, c2y, c2z, 0)
// ---------------------------------------------------------------------------
struct PackedBody {
    f3 v, w;
    m3 ii;
};
__device__ __forceinline__ PackedBody load_pb(const float4* p, uint32_t i) {
    const float4 a = p[4 * i], b = p[4 * i + 1], c = p[4 * i + 2], d = p[4 * i + 3];
    PackedBody r;
    r.v = f3{a.x, a.y, a.z};
    r.w = f3{a.w, b.x, b.y};
    r.ii = m3{f3{b.z, b.w, c.x}, f3{c.y, c.z, c.w}, f3{d.x, d.y, d.z}};
    return r;
}
__device__ __forceinline__ void store_vw(float4* p, uint32_t i, f3 v, f3 w, const m3& ii) {
    p[4 * i] = make_float4(v.x, v.y, v.z, w.x);
    p[4 * i + 1] = make_float4(w.y, w.z, ii.c0.x, ii.c0.y);
}
__device__ __forceinline__ PackedBody unpack_pb(float4 a, float4 b, float4 c, float4 d) {
    PackedBody r;
    r.v = f3{a.x, a.y, a.z};
    r.w = f3{a.w, b.x, b.y};
    r.ii = m3{f3{b.z, b.w, c.x}, f3{c.y, c.z, c.w}, f3{d.x, d.y, d.z}};
    return r;
}

// How a sweep reaches the body lines.  Lines: plain loads, for one
// workgroup (its waves share the CU's L1) or workgroups that order levels
// with agent-scope fences.
struct Lines {
    float4* p;
    __device__ __forceinline__ PackedBody load(uint32_t i) const { return load_pb(p, i); }
    __device__ __forceinline__ void store(uint32_t i, f3 v, f3 w, const m3& ii) const { store_vw(p, i, v, w, ii); }
};
// LinesL2: loads at device scope (the sc1 cache policy: they miss the CU's
// L1 and are served by the L2), for workgroups of one XCD that order their
// levels through that XCD's L2 without cache maintenance: every store has
// reached the L2 before the barrier (xcd_sync), and no load can hit a line
// its CU's L1 kept from an earlier level.  Stores are plain: the L1 writes
// through.
constexpr int kBufferDword3 = 0x00020000;  // raw buffer, gfx9 family
constexpr int kLoadDeviceScope = 16;       // cache policy bit SC1
struct LinesL2 {
    float4* p;
    __device__ __forceinline__ PackedBody load(uint32_t i) const {
        const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(p, 0, 0x7FFFFFFF, kBufferDword3);
        const uint32_t o = 64u * i;
        const float4 a = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, o, 0, kLoadDeviceScope));
        const float4 b = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, o + 16u, 0, kLoadDeviceScope));
        const float4 c = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, o + 32u, 0, kLoadDeviceScope));
        const float4 d = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, o + 48u, 0, kLoadDeviceScope));
        return unpack_pb(a, b, c, d);
    }
    __device__ __forceinline__ void store(uint32_t i, f3 v, f3 w, const m3& ii) const { store_vw(p, i, v, w, ii); }
};

__global__ void k_pack_bodies(BodyBuf b, uint32_t n, float4* vel, float4* pse) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 c0 = b.iw[3 * i], c1 = b.iw[3 * i + 1], c2 = b.iw[3 * i + 2];
    const float4 v = b.v[i], w = b.w[i], pv = b.pv[i], pw = b.pw[i];
    vel[4 * i] = make_float4(v.x, v.y, v.z, w.x);
    vel[4 * i + 1] = make_float4(w.y, w.z, c0.x, c0.y);
    vel[4 * i + 2] = make_float4(c0.z, c1.x, c1.y, c1.z);
    vel[4 * i + 3] = make_float4(c2.x, c2.y, c2.z, 0.0f);
    pse[4 * i] = make_float4(pv.x, pv.y, pv.z, pw.x);
    pse[4 * i + 1] = make_float4(pw.y, pw.z, c0.x, c0.y);
    pse[4 * i + 2] = vel[4 * i + 2];
    pse[4 * i + 3] = vel[4 * i + 3];
}

__global__ void k_unpack_bodies(BodyBuf b, uint32_t n, const float4* vel, const float4* pse) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 a = vel[4 * i], c = vel[4 * i + 1], pa = pse[4 * i], pc = pse[4 * i + 1];
    b.v[i] = make_float4(a.x, a.y, a.z, 0.0f);
    b.w[i] = make_float4(a.w, c.x, c.y, 0.0f);
    b.pv[i] = make_float4(pa.x, pa.y, pa.z, 0.0f);
    b.pw[i] = make_float4(pa.w, pc.x, pc.y, 0.0f);
}

// ---------------------------------------------------------------------------
// Per-unit sweep bodies, reading sweep records.  A unit is n <= U records of
// one body pair that the sequence applies one after the other: its records
// are loaded first, its two bodies once, and the points are applied in order
// with the bodies in registers, which is the arithmetic of applying them one
// point at a time.
// ---------------------------------------------------------------------------
// WarmStart (ContactSolver.cpp:117-132): stored rA / rB.
template <uint32_t U, class SB>
__device__ __forceinline__ void warm_unit(const WsRecBuf& r, const SB& sb, uint32_t e0, uint32_t n, uint4 bd,
                                          float wsf) {
    const bool da = bd.z & 2u, db = bd.z & 4u;
    float4 imp[U], n4[U], t14[U], t24[U], ra4[U], rb4[U];
#pragma unroll
    for (uint32_t k = 0; k < U; ++k)
        if (k < n) {
            const uint32_t e = e0 + k;
            imp[k] = r.imp[e];
            n4[k] = r.n[e];
            t14[k] = r.t1[e];
            t24[k] = r.t2[e];
            if (da) ra4[k] = r.ra[e];
            if (db) rb4[k] = r.rb[e];
        }
    PackedBody A, B;
    if (da) A = sb.load(bd.x);
    if (db) B = sb.load(bd.y);
#pragma unroll
    for (uint32_t k = 0; k < U; ++k)
        if (k < n) {
            f3 j = add(add(mul(ld3(n4[k]), imp[k].x), mul(ld3(t14[k]), imp[k].y)), mul(ld3(t24[k]), imp[k].z));
            j = mul(j, wsf);
            if (da) {
                A.v = add(A.v, mul(j, ra4[k].w));
                A.w = add(A.w, mmul(A.ii, cross(ld3(ra4[k]), j)));
            }
            if (db) {
                B.v = sub(B.v, mul(j, rb4[k].w));
                B.w = sub(B.w, mmul(B.ii, cross(ld3(rb4[k]), j)));
            }
        }
    if (da) sb.store(bd.x, A.v, A.w, A.ii);
    if (db) sb.store(bd.y, B.v, B.w, B.ii);
}

// Lanes of SolveVelocityBatchSIMD (ContactSolverSIMD.cpp:93-256).
// _mm_max_ps(a, b) = a > b ? a : b and _mm_min_ps(a, b) = a < b ? a : b.
// dsq accumulates each point's sum of squared impulse increments.
template <uint32_t U, class SB>
__device__ __forceinline__ void velocity_unit(const VelRecBuf& r, const SB& sb, uint32_t e0, uint32_t n, uint4 bd,
                                              float& dsq) {
    float4 ra4[U], rb4[U], n4[U], t14[U], t24[U], imp[U];
    float2 fb[U];
#pragma unroll
    for (uint32_t k = 0; k < U; ++k)
        if (k < n) {
            const uint32_t e = e0 + k;
            ra4[k] = r.ra[e];
            rb4[k] = r.rb[e];
            n4[k] = r.n[e];
            t14[k] = r.t1[e];
            t24[k] = r.t2[e];
            fb[k] = r.fb[e];
            imp[k] = r.imp[e];
        }
    const PackedBody A = sb.load(bd.x), B = sb.load(bd.y);
    const m3 ia = A.ii, ib = B.ii;
    f3 va = A.v, wa = A.w, vb = B.v, wb = B.w;
    const float zero = 0.0f;
#pragma unroll
    for (uint32_t k = 0; k < U; ++k)
        if (k < n) {
            const f3 ra = ld3(ra4[k]), rb = ld3(rb4[k]);
            const float ima = ra4[k].w, imb = rb4[k].w, cf = fb[k].x, nbias = fb[k].y;
            float4 im = imp[k];
            float pd = 0.0f;
            auto rel = [&]() {
                f3 xa2 = cross(wa, ra), xb2 = cross(wb, rb);
                return f3{(va.x + xa2.x) - (vb.x + xb2.x), (va.y + xa2.y) - (vb.y + xb2.y),
                          (va.z + xa2.z) - (vb.z + xb2.z)};
            };
            auto apply = [&](f3 j) {
                va = f3{va.x + j.x * ima, va.y + j.y * ima, va.z + j.z * ima};
                wa = add(wa, mmul(ia, cross(ra, j)));
                vb = f3{vb.x - j.x * imb, vb.y - j.y * imb, vb.z - j.z * imb};
                wb = sub(wb, mmul(ib, cross(rb, j)));
            };
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const f3 tv = t == 0 ? ld3(t14[k]) : ld3(t24[k]);
                const float mass = t == 0 ? t14[k].w : t24[k].w;
                const float acc = t == 0 ? im.y : im.z;
                const float vt = dot(rel(), tv);
                const float lam = (zero - mass) * vt;
                const float mf = cf * im.x;
                float na = acc + lam;
                const float lo = zero - mf;
                na = na > lo ? na : lo;
                na = na < mf ? na : mf;
                const float dl = na - acc;
                if (t == 0) im.y = na; else im.z = na;
                pd += dl * dl;
                apply(mul(tv, dl));
            }
            {
                const f3 nv = ld3(n4[k]);
                const float vn = dot(rel(), nv);
                const float lam = (zero - n4[k].w) * (vn - nbias);
                float na = im.x + lam;
                na = na > zero ? na : zero;
                const float dl = na - im.x;
                im.x = na;
                pd += dl * dl;
                apply(mul(nv, dl));
            }
            r.imp[e0 + k] = im;
            dsq += pd;
        }
    if (bd.z & 2u) sb.store(bd.x, va, wa, ia);
    if (bd.z & 4u) sb.store(bd.y, vb, wb, ib);
}

// The same lanes on two adjacent threads (a lane pair): side 0 holds body
// A, side 1 body B.  Each side forms its own term of the relative velocity
// and swaps it with its partner (DPP), so both compute the same relative
// velocity, impulse and clamp with the same operations, and each applies
// the impulse to its own body.  Body B's updates are the original products
// subtracted, written as the negated product added (a - b is a + (-b) in
// IEEE arithmetic, signed zeros included), so the results are bit-identical
// to velocity_unit's with about 40% fewer instructions per thread.
__device__ __forceinline__ float pair_swap(float v) {
    // quad_perm [1, 0, 3, 2]: lanes 2k and 2k + 1 exchange
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, false));
}
template <uint32_t U, class SB>
__device__ __forceinline__ void velocity_pair(const VelRecBuf& r, const SB& sb, uint32_t e0, uint32_t n, uint4 bd,
                                              uint32_t side, float& dsq) {
    float4 rr4[U], n4[U], t14[U], t24[U], imp[U];
    float2 fb[U];
    const float4* anchor = side ? r.rb : r.ra;
#pragma unroll
    for (uint32_t k = 0; k < U; ++k)
        if (k < n) {
            const uint32_t e = e0 + k;
            rr4[k] = anchor[e];
            n4[k] = r.n[e];
            t14[k] = r.t1[e];
            t24[k] = r.t2[e];
            fb[k] = r.fb[e];
            imp[k] = r.imp[e];
        }
    const uint32_t body = side ? bd.y : bd.x;
    const PackedBody X = sb.load(body);
    const m3 ii = X.ii;
    f3 v = X.v, w = X.w;
    const float zero = 0.0f;
    auto sgn = [side](float x) { return side ? -x : x; };
#pragma unroll
    for (uint32_t k = 0; k < U; ++k)
        if (k < n) {
            const f3 rr = ld3(rr4[k]);
            const float im = rr4[k].w, cf = fb[k].x, nbias = fb[k].y;
            float4 m = imp[k];
            float pd = 0.0f;
            auto rel = [&]() {
                const f3 x = cross(w, rr);
                const f3 own = f3{v.x + x.x, v.y + x.y, v.z + x.z};
                const f3 oth = f3{pair_swap(own.x), pair_swap(own.y), pair_swap(own.z)};
                const f3 a = side ? oth : own, b = side ? own : oth;
                return f3{a.x - b.x, a.y - b.y, a.z - b.z};
            };
            auto apply = [&](f3 j) {
                const f3 y = mmul(ii, cross(rr, j));
                v = f3{v.x + sgn(j.x * im), v.y + sgn(j.y * im), v.z + sgn(j.z * im)};
                w = f3{w.x + sgn(y.x), w.y + sgn(y.y), w.z + sgn(y.z)};
            };
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const f3 tv = t == 0 ? ld3(t14[k]) : ld3(t24[k]);
                const float mass = t == 0 ? t14[k].w : t24[k].w;
                const float acc = t == 0 ? m.y : m.z;
                const float vt = dot(rel(), tv);
                const float lam = (zero - mass) * vt;
                const float mf = cf * m.x;
                float na = acc + lam;
                const float lo = zero - mf;
                na = na > lo ? na : lo;
                na = na < mf ? na : mf;
                const float dl = na - acc;
                if (t == 0) m.y = na; else m.z = na;
                pd += dl * dl;
                apply(mul(tv, dl));
            }
            {
                const f3 nv = ld3(n4[k]);
                const float vn = dot(rel(), nv);
                const float lam = (zero - n4[k].w) * (vn - nbias);
                float na = m.x + lam;
                na = na > zero ? na : zero;
                const float dl = na - m.x;
                m.x = na;
                pd += dl * dl;
                apply(mul(nv, dl));
            }
            if (!side) {
                r.imp[e0 + k] = m;
                dsq += pd;
            }
        }
    if (bd.z & (side ? 4u : 2u)) sb.store(body, v, w, ii);
}

// SolveContactVelocity, the scalar path (ContactSolver.cpp:137-181).
template <uint32_t U, class SB>
__device__ __forceinline__ void scalar_unit(const VelRecBuf& r, const SB& sb, uint32_t e0, uint32_t n, uint4 bd,
                                            float& dsq) {
    float4 ra4[U], rb4[U], n4[U], t14[U], t24[U], imp[U];
    float2 fb[U];
#pragma unroll
    for (uint32_t k = 0; k < U; ++k)
        if (k < n) {
            const uint32_t e = e0 + k;
            ra4[k] = r.ra[e];
            rb4[k] = r.rb[e];
            n4[k] = r.n[e];
            t14[k] = r.t1[e];
            t24[k] = r.t2[e];
            fb[k] = r.fb[e];
            imp[k] = r.imp[e];
        }
    const PackedBody A = sb.load(bd.x), B = sb.load(bd.y);
    const m3 ia = A.ii, ib = B.ii;
    f3 va = A.v, wa = A.w, vb = B.v, wb = B.w;
    const bool da = bd.z & 2u, db = bd.z & 4u;
#pragma unroll
    for (uint32_t k = 0; k < U; ++k)
        if (k < n) {
            const f3 ra = ld3(ra4[k]), rb = ld3(rb4[k]);
            float4 im = imp[k];
            float pd = 0.0f;
            auto apply = [&](f3 j) {
                if (da) {
                    va = add(va, mul(j, ra4[k].w));
                    wa = add(wa, mmul(ia, cross(ra, j)));
                }
                if (db) {
                    vb = sub(vb, mul(j, rb4[k].w));
                    wb = sub(wb, mmul(ib, cross(rb, j)));
                }
            };
            f3 vr = sub(add(va, cross(wa, ra)), add(vb, cross(wb, rb)));
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const f3 tv = t == 0 ? ld3(t14[k]) : ld3(t24[k]);
                const float mass = t == 0 ? t14[k].w : t24[k].w;
                float acc = t == 0 ? im.y : im.z;
                float lam = -mass * dot(vr, tv);
                const float mf = fb[k].x * im.x;
                const float ni = fclamp(acc + lam, -mf, mf);
                lam = ni - acc;
                if (t == 0) im.y = ni; else im.z = ni;
                pd += lam * lam;
                apply(mul(tv, lam));
            }
            vr = sub(add(va, cross(wa, ra)), add(vb, cross(wb, rb)));
            float lam = -n4[k].w * (dot(vr, ld3(n4[k])) - fb[k].y);
            const float ni = fmax_(im.x + lam, 0.0f);
            lam = ni - im.x;
            im.x = ni;
            pd += lam * lam;
            apply(mul(ld3(n4[k]), lam));
            r.imp[e0 + k] = im;
            dsq += pd;
        }
    if (da) sb.store(bd.x, va, wa, ia);
    if (db) sb.store(bd.y, vb, wb, ib);
}

// SolveBiasConstraint (ContactSolver.cpp:257-290).
template <uint32_t U, class SB>
__device__ __forceinline__ void bias_unit(const WsRecBuf& r, const SB& ps, uint32_t e0, uint32_t n, uint4 bd) {
    float4 qa4[U], qb4[U], n4[U], imp[U];
    float ima[U], imb[U];
#pragma unroll
    for (uint32_t k = 0; k < U; ++k)
        if (k < n) {
            const uint32_t e = e0 + k;
            qa4[k] = r.qa[e];
            qb4[k] = r.qb[e];
            n4[k] = r.n[e];
            imp[k] = r.imp[e];
            ima[k] = r.ra[e].w;
            imb[k] = r.rb[e].w;
        }
    const PackedBody A = ps.load(bd.x), B = ps.load(bd.y);
    f3 pva = A.v, pwa = A.w, pvb = B.v, pwb = B.w;
    bool moved = false;
#pragma unroll
    for (uint32_t k = 0; k < U; ++k) {
        if (k >= n) continue;
        const float pbias = qa4[k].w;
        if (!(ima[k] <= 0.0f && imb[k] <= 0.0f) && !(pbias <= 0.0f)) {
            const f3 ra = ld3(qa4[k]), rb = ld3(qb4[k]);
            const f3 vr = sub(add(pva, cross(pwa, ra)), add(pvb, cross(pwb, rb)));
            float4 im = imp[k];
            float lam = -n4[k].w * (dot(vr, ld3(n4[k])) - pbias);
            const float ni = fmax_(im.w + lam, 0.0f);
            lam = ni - im.w;
            im.w = ni;
            r.imp[e0 + k] = im;
            const f3 j = mul(ld3(n4[k]), lam);
            if (bd.z & 2u) {
                pva = add(pva, mul(j, ima[k]));
                pwa = add(pwa, mmul(A.ii, cross(ra, j)));
            }
            if (bd.z & 4u) {
                pvb = sub(pvb, mul(j, imb[k]));
                pwb = sub(pwb, mmul(B.ii, cross(rb, j)));
            }
            moved = true;
        }
    }
    if (moved && (bd.z & 2u)) ps.store(bd.x, pva, pwa, A.ii);
    if (moved && (bd.z & 4u)) ps.store(bd.y, pvb, pwb, B.ii);
}

// Level offsets of one group, staged in LDS when they fit (the sweep reads
// them after every barrier).
constexpr uint32_t kLevelsLds = 2048;
struct LevelOffsets {
    const uint32_t* glob;
    const uint32_t* lds;
    uint32_t base;
    bool in_lds;
    __device__ __forceinline__ uint32_t operator()(uint32_t l) const {
        if (in_lds) return lds[l - base];
        return glob[l];
    }
};
__device__ __forceinline__ LevelOffsets stage_offsets(const Schedule& s, uint2 range, uint32_t* lds) {
    const uint32_t n = range.y - range.x + 1;
    LevelOffsets o{s.lvl_off, lds, range.x, n <= kLevelsLds};
    if (o.in_lds)
        for (uint32_t k = threadIdx.x; k < n; k += blockDim.x) lds[k] = s.lvl_off[range.x + k];
    return o;
}

// Levels are swept one at a time.  The unit descriptors of the next level
// are fetched before the current level is processed (they are constant
// during the solve), so after each barrier a unit waits for one round of
// loads only.  With `parts` workgroups sharing the levels, workgroup `part`
// takes the part-th contiguous slice of every level (its threads stride
// through it) and the barrier is device-wide; one workgroup alone uses
// __syncthreads().
struct Slice {
    uint32_t lo, hi;
};
__device__ __forceinline__ Slice level_slice(const LevelOffsets& off, uint32_t l, uint32_t parts, uint32_t part) {
    const uint32_t e0 = off(l), e1 = off(l + 1);
    if (parts == 1) return {e0, e1};
    const uint32_t chunk = (e1 - e0 + parts - 1) / parts;
    const uint32_t lo = min(e0 + part * chunk, e1);
    return {lo, min(lo + chunk, e1)};
}
// A unit as the sweep fetches it: the raw descriptor (U > 1) or the body
// entry of record u (U == 1).  It is decoded only where it is used, so the
// fetch of the next level's units does not wait for the load.
struct UnitRef {
    uint4 d;
    uint32_t u;
};
template <uint32_t U>
__device__ __forceinline__ UnitRef unit_fetch(const uint4* unit, const uint4* bd, uint32_t u) {
    return UnitRef{U > 1 ? unit[u] : bd[u], u};
}
template <uint32_t U, class F>
__device__ __forceinline__ void unit_apply(const UnitRef& r, uint32_t side, F&& f) {
    if (U > 1)
        f(r.d.w, min(r.d.z >> 8, U), make_uint4(r.d.x, r.d.y, r.d.z & 0xFFu, 0u), side);
    else
        f(r.u, 1u, r.d, side);
}
// Device-wide barriers of the cooperating workgroups.  The sync block
// (kSyncWords words, cleared before every cooperative launch) holds the
// arrivals of the launch, the XCC chosen for the sweep and each XCC's
// arrivals (k_solve_xcd), the level counter and the give-up flag.  Every
// wait is bounded by `limit` polls: a workgroup that waits longer raises
// the flag, every waiting workgroup then stops waiting, and the sweep ends
// early with the flag set for the host to report.
__device__ __forceinline__ uint32_t load_dev(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Thread 0 only: wait until *w >= target, or the flag is up, or the limit is
// spent (then raise the flag); false if the flag is up.
__device__ __forceinline__ bool spin_until(uint32_t* sync, const uint32_t* w, uint32_t target, uint32_t limit) {
    uint32_t polls = 0;
    while (load_dev(w) < target) {
        if (load_dev(sync + kSyncGiveUp)) return false;
        if (++polls >= limit) {
            atomicOr(sync + kSyncGiveUp, 1u);
            return false;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    return true;
}
// Workgroups anywhere on the device: an agent-scope release fence before
// the arrival publishes the workgroup's body updates of the level (it
// writes the XCD's dirty L2 lines back), an acquire fence after the wait
// makes the others' visible.
__device__ __forceinline__ bool grid_sync(uint32_t* sync, uint32_t target, uint32_t limit) {
    __shared__ uint32_t ok;
    __syncthreads();
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        atomicAdd(sync + kSyncLevel, 1u);
        ok = spin_until(sync, sync + kSyncLevel, target, limit);
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
    return ok != 0u;
}
// Workgroups of one XCD, which share its L2: each wave waits until its
// stores are acknowledged (they are then in the L2), the workgroup meets,
// and thread 0 arrives and waits.  The body lines are read with
// device-scope loads (LinesL2), which the L2 serves, so no cache is
// written back or invalidated between levels.
__device__ __forceinline__ void wait_stores() { __builtin_amdgcn_s_waitcnt(0x0F70); }  // vmcnt(0)
__device__ __forceinline__ bool xcd_sync(uint32_t* sync, uint32_t target, uint32_t limit) {
    __shared__ uint32_t ok;
    wait_stores();
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicAdd(sync + kSyncLevel, 1u);
        ok = spin_until(sync, sync + kSyncLevel, target, limit);
    }
    __syncthreads();
    return ok != 0u;
}
// The XCC (accelerator complex die) this workgroup runs on.
__device__ __forceinline__ uint32_t xcc_id() {
    return __builtin_amdgcn_s_getreg(20 | (0 << 6) | ((16 - 1) << 11)) & 0xFu;  // HW_REG_XCC_ID
}

// ---------------------------------------------------------------------------
// A sweep through one group's levels, one level per step, with the next
// level's units prefetched; after the group's last level it starts over
// (the next iteration).  L threads per unit.
// ---------------------------------------------------------------------------
template <uint32_t U, uint32_t L>
struct LevelStream {
    LevelOffsets off;
    uint2 range;
    const uint4 *bd, *unit;
    uint32_t parts, part, t, nt, side, l;
    Slice s;
    UnitRef nxt;
    __device__ __forceinline__ LevelStream(const LevelOffsets& o, uint2 r, const uint4* b, const uint4* un,
                                           uint32_t parts_, uint32_t part_)
        : off(o), range(r), bd(b), unit(un), parts(parts_), part(part_), t(threadIdx.x / L), nt(blockDim.x / L),
          side(threadIdx.x % L), l(r.x), s{0, 0}, nxt{make_uint4(0, 0, 0, 0), 0} {
        if (r.x < r.y) fetch(r.x);
    }
    __device__ __forceinline__ void fetch(uint32_t lv) {
        s = level_slice(off, lv, parts, part);
        if (s.lo + t < s.hi) nxt = unit_fetch<U>(unit, bd, s.lo + t);
    }
    // Applies f to this thread's units of the current level and moves on;
    // true when that was the group's last level.
    template <class F>
    __device__ __forceinline__ bool step(F&& f) {
        const Slice sl = s;
        const UnitRef cur = nxt;
        const bool last = l + 1 == range.y;
        l = last ? range.x : l + 1;
        fetch(l);
        for (uint32_t u = sl.lo + t; u < sl.hi; u += nt)
            unit_apply<U>(u == sl.lo + t ? cur : unit_fetch<U>(unit, bd, u), side, f);
        return last;
    }
};

// The solve of one group of islands (WarmStart, then SolveVelocity x
// velocityIterations, then SolveBias x positionIterations; PhysicsWorld.cpp
// SolveIsland), level by level with `sync` between levels.  The bias
// iterations change only pseudo velocities and the bias impulses, which no
// velocity iteration reads or writes, so bias level k runs in the same
// interval as velocity step k: the same per-body order of updates with
// fewer barriers.  `resid` (optional) receives the sum of squared impulse
// increments of each velocity sweep.
template <uint32_t UW, uint32_t UV, uint32_t L, class SB, class VelF, class Sync>
__device__ __forceinline__ void velocity_and_bias(const WsRecBuf& wr, const SB& sbp, const LevelOffsets& wo, uint2 wrange,
                                                  const LevelOffsets& vo, uint2 vrange, const uint4* vbd,
                                                  const Schedule& ws, const Schedule& vel, const SolveParams& sp,
                                                  float* resid, uint32_t parts, uint32_t part, VelF&& vf,
                                                  Sync&& sync) {
    const uint32_t vsteps = (vrange.y - vrange.x) * sp.velocity_iterations;
    const uint32_t bsteps =
        sp.use_velocity_bias || (sp.phases & kPhaseBiasDone) ? 0u : (wrange.y - wrange.x) * sp.position_iterations;
    LevelStream<UV, L> v(vo, vrange, vbd, vel.unit, parts, part);
    LevelStream<UW, 1> b(wo, wrange, wr.bd, ws.unit, parts, part);
    float dsq = 0.0f;
    uint32_t it = 0;
    const uint32_t steps = vsteps > bsteps ? vsteps : bsteps;
    for (uint32_t k = 0; k < steps; ++k) {
        if (k < vsteps && v.step([&](uint32_t e, uint32_t n, uint4 bd, uint32_t side) { vf(e, n, bd, side, dsq); })) {
            if (resid && dsq != 0.0f) atomicAdd(resid + it, dsq);
            dsq = 0.0f;
            ++it;
        }
        if (k < bsteps) b.step([&](uint32_t e, uint32_t n, uint4 bd, uint32_t) { bias_unit<UW>(wr, sbp, e, n, bd); });
        if (!sync()) return;
    }
}
template <uint32_t UW, uint32_t UV, class SB, class Sync>
__device__ __forceinline__ void solve_group(const VelRecBuf& vr, const WsRecBuf& wr, const SB& sbv, const SB& sbp,
                                            const Schedule& ws, const Schedule& vel, const SolveParams& sp,
                                            float* resid, uint32_t group, uint32_t parts, uint32_t part,
                                            uint32_t* ws_lds, uint32_t* vel_lds, Sync&& sync) {
    // A warm-start-only launch may run before the velocity schedule exists:
    // it reads nothing of it.
    const bool velocity = (sp.phases & kPhaseVelocity) != 0;
    const uint2 wrange = ws.grp[group];
    const uint2 vrange = velocity ? vel.grp[group] : make_uint2(0u, 0u);
    const LevelOffsets wo = stage_offsets(ws, wrange, ws_lds);
    const LevelOffsets vo = velocity ? stage_offsets(vel, vrange, vel_lds) : wo;
    __syncthreads();
    const float wsf = sp.warm_start_factor;
    // The bias sweeps read and write only pseudo velocities and the bias
    // impulses, which neither the warm start nor a velocity sweep touches,
    // so they may also run here, before the velocity iterations: every body
    // sees its pseudo-velocity updates in the same order.  Bias level k then
    // shares warm-start level k's interval.  The two sweeps take the same
    // levels in the same slices, so a record's warm-start unit (which reads
    // its cached impulses) and its bias unit (which rewrites them with the
    // bias impulse changed) run on one thread, in that order.
    const uint32_t wsteps = (sp.phases & kPhaseWarm) ? wrange.y - wrange.x : 0u;
    const uint32_t bsteps = (sp.phases & kPhaseBiasAlone) && !sp.use_velocity_bias
                                ? (wrange.y - wrange.x) * sp.position_iterations
                                : 0u;
    if (wsteps || bsteps) {
        LevelStream<UW, 1> w(wo, wrange, wr.bd, ws.unit, parts, part);
        LevelStream<UW, 1> b(wo, wrange, wr.bd, ws.unit, parts, part);
        const uint32_t steps = wsteps > bsteps ? wsteps : bsteps;
        for (uint32_t k = 0; k < steps; ++k) {
            if (k < wsteps)
                w.step([&](uint32_t e, uint32_t n, uint4 bd, uint32_t) { warm_unit<UW>(wr, sbv, e, n, bd, wsf); });
            if (k < bsteps)
                b.step([&](uint32_t e, uint32_t n, uint4 bd, uint32_t) { bias_unit<UW>(wr, sbp, e, n, bd); });
            if (!sync()) return;
        }
    }
    if (!velocity) return;
    if (sp.use_simd)
        velocity_and_bias<UW, UV, 2>(
            wr, sbp, wo, wrange, vo, vrange, vr.bd, ws, vel, sp, resid, parts, part,
            [&](uint32_t e, uint32_t n, uint4 bd, uint32_t side, float& dsq) {
                velocity_pair<UV>(vr, sbv, e, n, bd, side, dsq);
            },
            sync);
    else
        velocity_and_bias<UW, UV, 1>(
            wr, sbp, wo, wrange, vo, vrange, vr.bd, ws, vel, sp, resid, parts, part,
            [&](uint32_t e, uint32_t n, uint4 bd, uint32_t, float& dsq) { scalar_unit<UV>(vr, sbv, e, n, bd, dsq); },
            sync);
}

// The persistent group sweep: one workgroup per group of islands.  The units
// of a level share no dynamic body, so the threads of the workgroup process
// a level in parallel; __syncthreads() orders levels (the waves of a
// workgroup share the CU's L1, so no cache maintenance is needed between
// levels).  UW / UV: the largest warm-start / bias and velocity unit (1 also
// serves schedules without unit descriptors).
template <uint32_t UW, uint32_t UV>
__global__ void __launch_bounds__(solve_threads(UW, UV)) k_solve_groups(VelRecBuf vr, WsRecBuf wr, float4* sbv, float4* sbp,
                                                                Schedule ws, Schedule vel, SolveParams sp,
                                                                float* resid, uint32_t group0) {
    __shared__ uint32_t ws_lds[kLevelsLds], vel_lds[kLevelsLds];
    solve_group<UW, UV>(vr, wr, Lines{sbv}, Lines{sbp}, ws, vel, sp, resid, group0 + blockIdx.x, 1u, 0u, ws_lds,
                        vel_lds, [] {
                            __syncthreads();
                            return true;
                        });
}

// One group solved by gridDim.x cooperating workgroups (the launch makes
// them co-resident): the same sweeps, each level split over all of them.
template <uint32_t UW, uint32_t UV>
__global__ void __launch_bounds__(solve_threads(UW, UV)) k_solve_coop(VelRecBuf vr, WsRecBuf wr, float4* sbv, float4* sbp,
                                                                 Schedule ws, Schedule vel, SolveParams sp,
                                                                 float* resid, uint32_t group, uint32_t* sync,
                                                                 uint32_t limit) {
    __shared__ uint32_t ws_lds[kLevelsLds], vel_lds[kLevelsLds];
    uint32_t phase = 0;
    solve_group<UW, UV>(vr, wr, Lines{sbv}, Lines{sbp}, ws, vel, sp, resid, group, gridDim.x, blockIdx.x, ws_lds,
                        vel_lds, [&] { return grid_sync(sync, gridDim.x * ++phase, limit); });
}

// The same on the workgroups of one XCD.  The launch is cooperative and
// holds several workgroups per XCD (they are dealt to the XCDs in turn).
// Each workgroup reads its XCC, the first to arrive names the XCC of the
// sweep, and once all have arrived the workgroups of that XCC sweep the
// group's levels with xcd_sync between them; the others return.
template <uint32_t UW, uint32_t UV>
__global__ void __launch_bounds__(solve_threads(UW, UV)) k_solve_xcd(VelRecBuf vr, WsRecBuf wr, float4* sbv, float4* sbp,
                                                                Schedule ws, Schedule vel, SolveParams sp,
                                                                float* resid, uint32_t group, uint32_t* sync,
                                                                uint32_t limit) {
    __shared__ uint32_t ws_lds[kLevelsLds], vel_lds[kLevelsLds];
    __shared__ uint32_t role[2];
    if (threadIdx.x == 0) {
        const uint32_t x = xcc_id();
        atomicCAS(sync + kSyncTarget, 0xFFFFFFFFu, x);
        const uint32_t idx = atomicAdd(sync + kSyncXcc + x, 1u);
        atomicAdd(sync + kSyncArrive, 1u);
        const bool ok = spin_until(sync, sync + kSyncArrive, gridDim.x, limit);
        const uint32_t target = load_dev(sync + kSyncTarget);
        role[0] = load_dev(sync + kSyncXcc + target);
        role[1] = ok && x == target ? idx : 0xFFFFFFFFu;
    }
    __syncthreads();
    const uint32_t parts = role[0], part = role[1];
    if (part == 0xFFFFFFFFu) return;
    uint32_t phase = 0;
    solve_group<UW, UV>(vr, wr, LinesL2{sbv}, LinesL2{sbp}, ws, vel, sp, resid, group, parts, part, ws_lds, vel_lds,
                        [&] { return xcd_sync(sync, parts * ++phase, limit); });
}

// Clears a cooperative launch's sync block, the give-up flag included only
// with `all` (a kernel, not a memset: the launches may be graph-captured).
__global__ void k_sync_reset(uint32_t* sync, uint32_t all) {
    const uint32_t i = threadIdx.x;
    if (i < kSyncWords && (all || i != kSyncGiveUp)) sync[i] = i == kSyncTarget ? 0xFFFFFFFFu : 0u;
}

// ---------------------------------------------------------------------------
// Natural-map residual of the current iterate (diagnostic; the reference
// computes none).  Per non-sensor point, with the bodies' velocities as the
// solve left them and the anchors as the velocity sweep forms them (rA =
// qA localAnchorA):
//   normal:   min(lambda_n, m_n (vRel . n - bias))                 (N s)
//   friction: lambda_t - clamp(lambda_t - m_t (vRel . t), +-mu lambda_n),
//             for t1 and t2 (what one more projected step would change)
//   cone:     max(0, |lambda_t| - mu lambda_n), for t1 and t2 (inside the
//             friction term, reported apart)
// bias = the restitution bias, plus the position bias with velocity bias.
// acc (FP64): squared sums of the three terms and the point count.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_natural_residual(ContactBuf c, BodyBuf b, uint32_t np, int use_vbias,
                                                          double* acc) {
    __shared__ double red[4][256];
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    double rn = 0.0, rt = 0.0, rc = 0.0, cnt = 0.0;
    if (p < np) {
        const uint4 bd = c.body[p];
        if (!(bd.z & 1u)) {
            const f4q qa = ldq(b.q[bd.x]), qb = ldq(b.q[bd.y]);
            const f3 ra = rotate(qa, ld3(c.la[p])), rb = rotate(qb, ld3(c.lb[p]));
            const f3 va = ld3(b.v[bd.x]), wa = ld3(b.w[bd.x]), vb = ld3(b.v[bd.y]), wb = ld3(b.w[bd.y]);
            const f3 vr = sub(add(va, cross(wa, ra)), add(vb, cross(wb, rb)));
            const float4 n4 = c.n[p], t14 = c.t1[p], t24 = c.t2[p], imp = c.imp[p];
            const float mu = c.la[p].w;
            const float bias = c.ra[p].w + (use_vbias ? c.rb[p].w : 0.0f);
            const double ln = imp.x, lim = double(mu) * ln;
            const double w = double(dot(vr, ld3(n4))) - bias;
            const double r0 = fmin(ln, double(n4.w) * w);
            rn = r0 * r0;
            const double lt[2] = {imp.y, imp.z};
            const float4 tt[2] = {t14, t24};
            for (int k = 0; k < 2; ++k) {
                const double step = lt[k] - double(tt[k].w) * double(dot(vr, ld3(tt[k])));
                const double proj = fmin(fmax(step, -lim), lim);
                const double r = lt[k] - proj;
                const double over = fmax(0.0, fabs(lt[k]) - lim);
                rt += r * r;
                rc += over * over;
            }
            cnt = 1.0;
        }
    }
    red[0][threadIdx.x] = rn;
    red[1][threadIdx.x] = rt;
    red[2][threadIdx.x] = rc;
    red[3][threadIdx.x] = cnt;
    __syncthreads();
    for (uint32_t off = blockDim.x / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off)
            for (int k = 0; k < 4; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x < 4) atomicAdd(acc + threadIdx.x, red[threadIdx.x][0]);
}
void launch_natural_residual(hipStream_t s, const ContactBuf& c, const BodyBuf& b, uint32_t np, int use_vbias,
                             double* acc) {
    if (np) k_natural_residual<<<(np + 255) / 256, 256, 0, s>>>(c, b, np, use_vbias, acc);
}

// Unit descriptors from the unit bounds and the records' body entries.
__global__ void k_unit_desc(const uint32_t* unit, const uint4* bd, uint32_t nu, uint4* desc) {
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nu) return;
    const uint32_t e = unit[u], n = unit[u + 1] - e;
    const uint4 b = bd[e];
    desc[u] = make_uint4(b.x, b.y, (b.z & 0xFFu) | (n << 8), e);
}

// Final impulses back to constraint-point order (for the cache store and
// read-back): velocity impulses from the velocity records, the bias impulse
// from the warm-start/bias records.
__global__ void k_writeback(VelRecBuf vr, WsRecBuf wr, ContactBuf c, const uint32_t* vel_pos,
                            const uint32_t* ws_pos, uint32_t np, const uint32_t* np_dev) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (np_dev) np = *np_dev;
    if (p >= np) return;
    float4 imp = c.imp[p];
    const uint32_t e = vel_pos[p];
    if (e != 0xFFFFFFFFu) {
        const float4 v = vr.imp[e];
        imp.x = v.x; imp.y = v.y; imp.z = v.z;
    }
    const uint32_t w = ws_pos[p];
    if (w != 0xFFFFFFFFu) imp.w = wr.imp[w].w;
    c.imp[p] = imp;
}

// ---------------------------------------------------------------------------
// Cache store (PhysicsWorld.cpp:568-581), one thread per constraint.  The
// host guarantees free capacity (load factor <= 1/2), so probing ends.
// ---------------------------------------------------------------------------
__global__ void k_cache_store(ContactBuf c, BodyBuf b, CacheBuf cache, uint32_t nc, const uint32_t* nc_dev) {
    uint32_t ci = blockIdx.x * blockDim.x + threadIdx.x;
    if (nc_dev) nc = *nc_dev;
    if (ci >= nc) return;
    const uint4 info = c.info[ci];
    const unsigned long long key = pair_key(b.packed[info.x], b.packed[info.y]);
    uint32_t mask = cache.cap - 1u, s = cache_hash(key) & mask;
    for (uint32_t probe = 0; probe < cache.cap; ++probe) {
        unsigned long long prev = atomicCAS(cache.key + s, 0ull, key);
        if (prev == 0ull) atomicAdd(cache.used, 1u);
        if (prev == 0ull || prev == key) break;
        s = (s + 1u) & mask;
    }
    const uint32_t np = info.w & 0xFFu;
    cache.count[s] = np;
    for (uint32_t k = 0; k < np; ++k) {
        const uint32_t p = info.z + k;
        const float4 imp = c.imp[p];
        const float4 pa = c.pa[p];
        cache.p[s * kPairPoints + k] = make_float4(pa.x, pa.y, pa.z, imp.x);
        cache.t[s * kPairPoints + k] = make_float2(imp.y, imp.z);
    }
}

__global__ void k_cache_rehash(CacheBuf src, CacheBuf dst) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= src.cap) return;
    const unsigned long long key = src.key[i];
    if (key == 0ull) return;
    uint32_t mask = dst.cap - 1u, s = cache_hash(key) & mask;
    for (uint32_t probe = 0; probe < dst.cap; ++probe) {
        if (atomicCAS(dst.key + s, 0ull, key) == 0ull) {
            atomicAdd(dst.used, 1u);
            break;
        }
        s = (s + 1u) & mask;
    }
    dst.count[s] = src.count[i];
    for (int k = 0; k < kPairPoints; ++k) {
        dst.p[s * kPairPoints + k] = src.p[i * kPairPoints + k];
        dst.t[s * kPairPoints + k] = src.t[i * kPairPoints + k];
    }
}

// ---------------------------------------------------------------------------
// IntegratePositions (PhysicsWorld.cpp:785-802), pseudo-velocity integration
// (:275-289) and UpdateBodyInertias (:854-861), one thread per body.
// ---------------------------------------------------------------------------
__global__ void k_finish(BodyBuf b, uint32_t n, float dt, int split_impulse) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t m = b.meta[i], mo = meta_motion(m);
    if (mo == 0u || meta_asleep(m)) return;  // the active list's non-static bodies
    float4 x = b.x[i];
    f3 t = ld3(x);
    f4q q = ldq(b.q[i]);
    const f3 v = ld3(b.v[i]), w = ld3(b.w[i]);
    t = add(t, mul(v, dt));
    if (len2(w) > 0.0f) q = integrate_rotation(q, w, dt);
    if (split_impulse) {
        const f3 pv = ld3(b.pv[i]), pw = ld3(b.pw[i]);
        if (len2(pv) > 0.0f) t = add(t, mul(pv, dt));
        if (len2(pw) > 0.0f) q = integrate_rotation(q, pw, dt);
        b.pv[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        b.pw[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    b.x[i] = make_float4(t.x, t.y, t.z, x.w);
    b.q[i] = make_float4(q.x, q.y, q.z, q.w);
    if (mo == 2u) {
        const m3 iw = world_inertia(q, ldm(b.il + 3 * i));
        b.iw[3 * i + 0] = st3(iw.c0, 0.0f);
        b.iw[3 * i + 1] = st3(iw.c1, 0.0f);
        b.iw[3 * i + 2] = st3(iw.c2, 0.0f);
    }
}

// World inertia from the local one, for uploads of new body state.
__global__ void k_world_inertia(BodyBuf b, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (meta_motion(b.meta[i]) != 2u) {
        for (int k = 0; k < 3; ++k) b.iw[3 * i + k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return;
    }
    const m3 iw = world_inertia(ldq(b.q[i]), ldm(b.il + 3 * i));
    b.iw[3 * i + 0] = st3(iw.c0, 0.0f);
    b.iw[3 * i + 1] = st3(iw.c1, 0.0f);
    b.iw[3 * i + 2] = st3(iw.c2, 0.0f);
}

// ---------------------------------------------------------------------------
// Host launchers.
// ---------------------------------------------------------------------------
static inline uint32_t blocks(uint32_t n, uint32_t t) { return (n + t - 1) / t; }

void launch_integrate_velocities(hipStream_t s, const BodyBuf& b, uint32_t n, float dt, const float* g) {
    if (n) k_integrate_velocities<<<blocks(n, 256), 256, 0, s>>>(b, n, dt, g[0], g[1], g[2]);
}
void launch_build(hipStream_t s, const ContactBuf& c, const ManifoldBuf& m, const BodyBuf& b, const CacheBuf& cache,
                  uint32_t nc, int warm) {
    if (nc) k_build<<<blocks(nc, 256), 256, 0, s>>>(c, m, b, cache, nc, nullptr, warm);
}
void launch_prepare(hipStream_t s, const ContactBuf& c, const BodyBuf& b, uint32_t np, float dt, float rthr,
                    float baum, float slop, float maxb, int use_simd, int use_vbias, const VelRecBuf& vr,
                    const WsRecBuf& wr, const uint32_t* vel_pos, const uint32_t* ws_pos, float gdt, float4* feat,
                    float4* base, unsigned long long* load, uint32_t nbodies) {
    if (feat && load && np && nbodies) {
        k_fill_u64<<<blocks(nbodies, 256), 256, 0, s>>>(load, nbodies);
        k_body_load<<<blocks(np, 256), 256, 0, s>>>(c, b, np, gdt, load);
    }
    PrepParams pp{dt, rthr, baum, slop, maxb, use_simd, use_vbias, gdt, feat, base, feat ? load : nullptr};
    if (np) k_prepare<<<blocks(np, 256), 256, 0, s>>>(c, b, np, nullptr, pp, vr, wr, vel_pos, ws_pos);
}
void launch_velocity_records(hipStream_t s, const ContactBuf& c, const WsRecBuf& wr, const VelRecBuf& vr,
                             const uint32_t* vel_pos, const uint32_t* ws_pos, uint32_t np, int use_simd,
                             int use_vbias) {
    if (np) k_velocity_records<<<blocks(np, 256), 256, 0, s>>>(c, wr, vr, vel_pos, ws_pos, np, use_simd, use_vbias);
}
void launch_writeback(hipStream_t s, const VelRecBuf& vr, const WsRecBuf& wr, const ContactBuf& c,
                      const uint32_t* vel_pos, const uint32_t* ws_pos, uint32_t np) {
    if (np) k_writeback<<<blocks(np, 256), 256, 0, s>>>(vr, wr, c, vel_pos, ws_pos, np, nullptr);
}
// Descriptors of `nu` units whose bounds are unit[0..nu] (records written).
void launch_unit_desc(hipStream_t s, const uint32_t* unit, const uint4* bd, uint32_t nu, uint4* desc) {
    if (nu) k_unit_desc<<<blocks(nu, 256), 256, 0, s>>>(unit, bd, nu, desc);
}

// Sweep kernels by unit instantiation (see unit_kernel).
template <class F>
static auto by_units(UnitKernel uk, F&& f) {
    auto vel = [&](auto w) {
        if (uk.vel <= 1) return f(w, std::integral_constant<uint32_t, 1>{});
        if (uk.vel == 2) return f(w, std::integral_constant<uint32_t, 2>{});
        return f(w, std::integral_constant<uint32_t, kMaxUnit>{});
    };
    if (uk.ws <= 1) return vel(std::integral_constant<uint32_t, 1>{});
    if (uk.ws == 2) return vel(std::integral_constant<uint32_t, 2>{});
    return vel(std::integral_constant<uint32_t, kMaxUnit>{});
}
static void launch_groups(hipStream_t s, UnitKernel uk, uint32_t n, uint32_t threads, const VelRecBuf& vr,
                          const WsRecBuf& wr, float4* sbv, float4* sbp, const Schedule& ws, const Schedule& vel,
                          const SolveParams& sp, float* resid, uint32_t group0) {
    // Schedules without unit descriptors (the batched worlds) take one-point
    // units; a warm-start-only launch has no velocity schedule and needs none.
    if (!ws.unit || (!vel.unit && (sp.phases & kPhaseVelocity))) uk = UnitKernel{};
    by_units(uk, [&](auto w, auto v) {
        k_solve_groups<decltype(w)::value, decltype(v)::value>
            <<<n, threads, 0, s>>>(vr, wr, sbv, sbp, ws, vel, sp, resid, group0);
        return 0;
    });
}
static const void* coop_kernel(UnitKernel uk, bool one_xcd) {
    return by_units(uk, [one_xcd](auto w, auto v) {
        return one_xcd ? (const void*)k_solve_xcd<decltype(w)::value, decltype(v)::value>
                       : (const void*)k_solve_coop<decltype(w)::value, decltype(v)::value>;
    });
}

// packed: 8 float4 per body (velocity lines, then pseudo-velocity lines).
// uk: the unit instantiation (unit_kernel of the schedule's largest units;
// ignored without unit descriptors).
void launch_solve_groups(hipStream_t s, const VelRecBuf& vr, const WsRecBuf& wr, const BodyBuf& b, uint32_t nbodies,
                         float4* packed, const Schedule& ws, const Schedule& vel, const SolveParams& sp,
                         uint32_t groups, uint32_t threads, float* resid, UnitKernel uk, uint32_t bodies) {
    if (!groups || !nbodies) return;
    float4* sbv = packed;
    float4* sbp = packed + 4 * size_t(nbodies);
    if (bodies & kPackBodies) k_pack_bodies<<<blocks(nbodies, 256), 256, 0, s>>>(b, nbodies, sbv, sbp);
    launch_groups(s, uk, groups, threads, vr, wr, sbv, sbp, ws, vel, sp, resid, 0u);
    if (bodies & kUnpackBodies) k_unpack_bodies<<<blocks(nbodies, 256), 256, 0, s>>>(b, nbodies, sbv, sbp);
}

void launch_sync_reset(hipStream_t s, uint32_t* sync) { k_sync_reset<<<1, 64, 0, s>>>(sync, 1u); }

// Group 0 by `coop_blocks` cooperating workgroups of `coop_threads`, the
// other groups one workgroup each.  one_xcd: the cooperating workgroups all
// run on one XCD (k_solve_xcd, launched with kXcds x coop_blocks workgroups
// of which one XCD's take part); else anywhere (k_solve_coop).  sync:
// kSyncWords words; limit: most polls of one wait.
hipError_t launch_solve_split(hipStream_t s, const VelRecBuf& vr, const WsRecBuf& wr, const BodyBuf& b,
                              uint32_t nbodies, float4* packed, const Schedule& ws, const Schedule& vel,
                              const SolveParams& sp, uint32_t groups, uint32_t threads, float* resid,
                              UnitKernel uk, uint32_t coop_blocks, uint32_t coop_threads, uint32_t* sync,
                              uint32_t limit, bool one_xcd, hipEvent_t coop_begin, hipEvent_t coop_end,
                              uint32_t bodies) {
    if (!groups || !nbodies) return hipSuccess;
    // The cooperative sweep reads unit arrays (the velocity one only when it
    // runs the velocity iterations).
    if (!ws.unit || (!vel.unit && (sp.phases & kPhaseVelocity))) return hipErrorInvalidValue;
    float4* sbv = packed;
    float4* sbp = packed + 4 * size_t(nbodies);
    if (bodies & kPackBodies) k_pack_bodies<<<blocks(nbodies, 256), 256, 0, s>>>(b, nbodies, sbv, sbp);
    k_sync_reset<<<1, 64, 0, s>>>(sync, 0u);
    VelRecBuf vr_ = vr;
    WsRecBuf wr_ = wr;
    Schedule ws_ = ws, vel_ = vel;
    SolveParams sp_ = sp;
    uint32_t group = 0;
    void* args[] = {&vr_, &wr_, &sbv, &sbp, &ws_, &vel_, &sp_, &resid, &group, &sync, &limit};
    if (coop_begin) hipEventRecord(coop_begin, s);
    const uint32_t grid = one_xcd ? kXcds * coop_blocks : coop_blocks;
    hipError_t e = hipLaunchCooperativeKernel(coop_kernel(uk, one_xcd), dim3(grid), dim3(coop_threads), args, 0, s);
    if (e != hipSuccess) return e;
    if (coop_end) hipEventRecord(coop_end, s);
    if (groups > 1) launch_groups(s, uk, groups - 1, threads, vr, wr, sbv, sbp, ws, vel, sp, resid, 1u);
    if (bodies & kUnpackBodies) k_unpack_bodies<<<blocks(nbodies, 256), 256, 0, s>>>(b, nbodies, sbv, sbp);
    return hipGetLastError();
}
void launch_cache_store(hipStream_t s, const ContactBuf& c, const BodyBuf& b, const CacheBuf& cache, uint32_t nc) {
    if (nc) k_cache_store<<<blocks(nc, 256), 256, 0, s>>>(c, b, cache, nc, nullptr);
}
void launch_cache_rehash(hipStream_t s, const CacheBuf& src, const CacheBuf& dst) {
    if (src.cap) k_cache_rehash<<<blocks(src.cap, 256), 256, 0, s>>>(src, dst);
}
void launch_finish(hipStream_t s, const BodyBuf& b, uint32_t n, float dt, int split) {
    if (n) k_finish<<<blocks(n, 256), 256, 0, s>>>(b, n, dt, split);
}
void launch_world_inertia(hipStream_t s, const BodyBuf& b, uint32_t n) {
    if (n) k_world_inertia<<<blocks(n, 256), 256, 0, s>>>(b, n);
}

// Capacity-sized launches whose counts are read on the device (graph-captured
// steps of the batched worlds).
void launch_build_counted(hipStream_t s, const ContactBuf& c, const ManifoldBuf& m, const BodyBuf& b,
                          const CacheBuf& cache, uint32_t cap, const uint32_t* nc_dev, int warm) {
    if (cap) k_build<<<blocks(cap, 256), 256, 0, s>>>(c, m, b, cache, cap, nc_dev, warm);
}
void launch_prepare_counted(hipStream_t s, const ContactBuf& c, const BodyBuf& b, uint32_t cap, const uint32_t* np_dev,
                            float dt, float rthr, float baum, float slop, float maxb, int use_simd, int use_vbias,
                            const VelRecBuf& vr, const WsRecBuf& wr, const uint32_t* vel_pos, const uint32_t* ws_pos) {
    PrepParams pp{dt, rthr, baum, slop, maxb, use_simd, use_vbias, 0.0f, nullptr, nullptr, nullptr};
    if (cap) k_prepare<<<blocks(cap, 256), 256, 0, s>>>(c, b, cap, np_dev, pp, vr, wr, vel_pos, ws_pos);
}
void launch_writeback_counted(hipStream_t s, const VelRecBuf& vr, const WsRecBuf& wr, const ContactBuf& c,
                              const uint32_t* vel_pos, const uint32_t* ws_pos, uint32_t cap, const uint32_t* np_dev) {
    if (cap) k_writeback<<<blocks(cap, 256), 256, 0, s>>>(vr, wr, c, vel_pos, ws_pos, cap, np_dev);
}
void launch_cache_store_counted(hipStream_t s, const ContactBuf& c, const BodyBuf& b, const CacheBuf& cache,
                                uint32_t cap, const uint32_t* nc_dev) {
    if (cap) k_cache_store<<<blocks(cap, 256), 256, 0, s>>>(c, b, cache, cap, nc_dev);
}

}  // namespace kz
