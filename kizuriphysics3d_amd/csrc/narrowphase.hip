// narrowphase.hip - convex narrowphase over broadphase pairs, hand-written
// for gfx950 (wave64).
//
// Two kernels:
//   k_collide_simple  one thread per pair.  Orders the pair (lower packed id
//                     first, src/World/PhysicsWorld.cpp:381-392), emits the
//                     closed-form paths (plane-convex, sphere-sphere) and
//                     appends every other pair to a work list.
//   k_collide_wave    one wavefront per work-list pair: GJK, EPA and
//                     polyhedral face clipping.  Hull vertices are held in
//                     registers, four per lane (256 max), and a support query
//                     is a wave arg-max; the EPA polytope lives in LDS and
//                     every per-face step (closest face, visibility, horizon,
//                     stable compaction, new faces) is a wave-parallel pass
//                     with ballot/prefix-sum compaction that preserves the
//                     reference's sequential order.
// Each function cites the reference routine whose arithmetic it follows;
// operation order is kept so that results are bit-identical.
#include "kz_device.h"
#include "kz_math.h"

namespace kz {

__device__ __forceinline__ f3 l3(const float4& v) { return f3{v.x, v.y, v.z}; }
__device__ __forceinline__ f4q lq(const float4& v) { return f4q{v.x, v.y, v.z, v.w}; }

struct ShapeView {
    uint32_t type;
    float4 prm;
    uint4 hull;  // vert offset, count, tri offset, count
};

struct Contact {
    f3 pa, pb, n;
    float sep;
    uint32_t feature;
};

struct PairOut {
    uint32_t count;
};

constexpr int kMaxFaceVerts = 16;

// Closed-form supports (include/Kizuri/Geometry/Shape.h:152-154,183-187,223-226).
__device__ __forceinline__ f3 support_closed(const ShapeView& s, f3 d) {
    if (s.type == kSphere) return mul(normalized_or(d, f3{0.0f, 1.0f, 0.0f}), s.prm.x);
    if (s.type == kBox)
        return f3{d.x >= 0.0f ? s.prm.x : -s.prm.x, d.y >= 0.0f ? s.prm.y : -s.prm.y, d.z >= 0.0f ? s.prm.z : -s.prm.z};
    // capsule
    float h = d.y >= 0.0f ? s.prm.x : -s.prm.x;
    return add(f3{0.0f, h, 0.0f}, mul(normalized_or(d, f3{0.0f, 1.0f, 0.0f}), s.prm.y));
}

// ===========================================================================
// Thread-per-pair kernel: ordering, plane and sphere-sphere paths.
// ===========================================================================
__device__ __forceinline__ void swap_contact(Contact& c) {
    f3 t = c.pa;
    c.pa = c.pb;
    c.pb = t;
    c.n = neg(c.n);
}

__device__ __forceinline__ void write_contact(const ManifoldBuf& m, uint32_t pair, int k, const Contact& c) {
    m.pa[pair * kPairPoints + k] = make_float4(c.pa.x, c.pa.y, c.pa.z, c.sep);
    m.pb[pair * kPairPoints + k] = make_float4(c.pb.x, c.pb.y, c.pb.z, __uint_as_float(c.feature));
    m.n[pair * kPairPoints + k] = make_float4(c.n.x, c.n.y, c.n.z, 0.0f);
}

// CollidePlaneConvex (src/Collision/CollideShape.cpp:370-430): A of the
// result is the convex shape, the normal points from the plane to it.
__device__ int plane_convex(const ShapeView& P, f4q qp, f3 tp, const ShapeView& B, f4q qb, f3 tb,
                            const float4* hverts, float max_sep, Contact* out) {
    const f3 pn = rotate(qp, l3(P.prm));
    const float pd = P.prm.w + dot(pn, tp);
    if (B.type == kBox || B.type == kHull) {
        Contact cand[8];
        int k = 0;
        const int nv = B.type == kBox ? 8 : (B.hull.y < uint32_t(kMaxFaceVerts) ? int(B.hull.y) : kMaxFaceVerts);
        for (int i = 0; i < nv && k < 8; ++i) {
            f3 v;
            if (B.type == kBox)
                v = xform(qb, tb, f3{(i & 1) ? B.prm.x : -B.prm.x, (i & 2) ? B.prm.y : -B.prm.y, (i & 4) ? B.prm.z : -B.prm.z});
            else
                v = xform(qb, tb, l3(hverts[B.hull.x + i]));
            const float d = dot(pn, v) - pd;
            if (d > max_sep) continue;
            cand[k].n = pn;
            cand[k].pa = v;
            cand[k].pb = sub(v, mul(pn, d));
            cand[k].sep = d;
            cand[k].feature = 0;
            ++k;
        }
        if (k > 4) {
            for (int i = 0; i < 4; ++i) out[i] = cand[i * (k - 1) / 3];
            return 4;
        }
        for (int i = 0; i < k; ++i) out[i] = cand[i];
        return k;
    }
    const f3 ld = inv_rotate(qb, neg(pn));
    const f3 deep = xform(qb, tb, support_closed(B, ld));
    const float d = dot(pn, deep) - pd;
    if (d > max_sep) return 0;
    out[0].n = pn;
    out[0].pa = deep;
    out[0].pb = sub(deep, mul(pn, d));
    out[0].sep = d;
    out[0].feature = 0;
    return 1;
}

// CollideSphereSphere (CollideShape.cpp:349-365).
__device__ int sphere_sphere(float ra, f3 ta, float rb, f3 tb, float max_sep, Contact* out) {
    const f3 d = sub(ta, tb);
    const float dist = len(d);
    const float sep = dist - (ra + rb);
    if (sep > max_sep) return 0;
    const f3 n = dist > kEps ? divs(d, dist) : f3{0.0f, 1.0f, 0.0f};
    out[0].n = n;
    out[0].pa = sub(ta, mul(n, ra));
    out[0].pb = add(tb, mul(n, rb));
    out[0].sep = sep;
    out[0].feature = 0;
    return 1;
}

__device__ __forceinline__ ShapeView load_shape(const ShapeBuf& s, uint32_t i) {
    ShapeView v;
    v.type = s.type[i];
    v.prm = s.params[i];
    v.hull = v.type == kHull ? s.hull[i] : make_uint4(0, 0, 0, 0);
    return v;
}

// Pair inputs: body indices.  For the world path body = slot; for the
// standalone batch every pair side is its own "body" carrying a transform.
// The pair i: the closed-form classes (PhysicsWorld::Collide's skip and
// order, then NarrowPhase::Collide's plane and sphere-sphere paths) write
// their manifold; an iterative pair returns its work item.
__device__ __forceinline__ void collide_simple_pair(const uint2* pairs, uint32_t i, const BodyBuf& b, const ShapeBuf& s,
                                                    const ManifoldBuf& m, float max_sep, int world_mode,
                                                    uint32_t& item) {
    const uint2 pr = pairs[i];
    uint32_t ia = pr.x, ib = pr.y;
    m.count[i] = 0;
    bool swapped = false;
    if (world_mode) {
        // PhysicsWorld::Collide: skip unless one body is dynamic; collide in
        // packed-id order and swap the manifold back afterwards.
        if (meta_motion(b.meta[ia]) != 2u && meta_motion(b.meta[ib]) != 2u) return;
        swapped = b.packed[ia] > b.packed[ib];
        if (swapped) { uint32_t t = ia; ia = ib; ib = t; }
    }
    const ShapeView A = load_shape(s, meta_shape(b.meta[ia]));
    const ShapeView B = load_shape(s, meta_shape(b.meta[ib]));
    const float4 xa = b.x[ia], xb = b.x[ib];
    const f4q qa = lq(b.q[ia]), qb = lq(b.q[ib]);
    Contact c[4];
    int k = -1;
    bool swap_result = swapped;
    // NarrowPhase::Collide dispatch order (CollideShape.cpp:509-591).
    if (A.type == kPlane) {
        k = plane_convex(A, qa, l3(xa), B, qb, l3(xb), s.hverts, max_sep, c);
        if (k > 0) swap_result = !swap_result;
    } else if (B.type == kPlane) {
        k = plane_convex(B, qb, l3(xb), A, qa, l3(xa), s.hverts, max_sep, c);
    } else if (A.type == kSphere && B.type == kSphere) {
        k = sphere_sphere(A.prm.x, l3(xa), B.prm.x, l3(xb), max_sep, c);
    } else {
        // Iterative class: hand the (ordered) pair to the GJK kernel.
        item = i | (work_class(A.type, B.type) << kWorkPairBits) | (swapped ? 0x80000000u : 0u);
        return;
    }
    for (int j = 0; j < k; ++j) {
        if (swap_result) swap_contact(c[j]);
        write_contact(m, i, j, c[j]);
    }
    m.count[i] = uint32_t(k);
}

__global__ void k_collide_simple(const uint2* pairs, uint32_t n, const uint32_t* n_dev, BodyBuf b, ShapeBuf s,
                                 ManifoldBuf m, uint32_t* work, uint32_t* work_count, float max_sep, int world_mode) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (n_dev) n = *n_dev;  // count kept on the device (graph-captured steps)
    if (i >= n) return;
    uint32_t item = 0xFFFFFFFFu;
    collide_simple_pair(pairs, i, b, s, m, max_sep, world_mode, item);
    if (item != 0xFFFFFFFFu) work[atomicAdd(work_count, 1u)] = item;
}

// The work list in class-major order (the order inside a class is
// arbitrary: the pairs are independent), in three passes over chunks of
// kWorkChunk items: class counts per chunk (LDS), an exclusive scan over
// (class, chunk), and the scatter.  No global atomics.
__device__ __forceinline__ uint32_t item_class(uint32_t item) {
    return (item >> kWorkPairBits) & (kWorkClasses - 1u);
}
__global__ void __launch_bounds__(kWorkChunk) k_bin_count(const uint32_t* work, const uint32_t* work_count,
                                                          uint32_t* hist) {
    __shared__ uint32_t cnt[kWorkClasses];
    const uint32_t total = *work_count, lo = blockIdx.x * kWorkChunk;
    if (lo >= total) return;
    if (threadIdx.x < kWorkClasses) cnt[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t i = lo + threadIdx.x;
    if (i < total) atomicAdd(cnt + item_class(work[i]), 1u);
    __syncthreads();
    if (threadIdx.x < kWorkClasses) hist[size_t(threadIdx.x) * gridDim.x + blockIdx.x] = cnt[threadIdx.x];
}
// One workgroup: hist (classes x chunks, class-major) -> exclusive offsets.
__global__ void __launch_bounds__(1024) k_bin_scan(uint32_t* hist, const uint32_t* work_count, uint32_t chunks_cap) {
    __shared__ uint32_t part[1024];
    const uint32_t chunks = (*work_count + kWorkChunk - 1) / kWorkChunk;
    const uint32_t n = kWorkClasses * chunks_cap;
    // Each thread scans a contiguous run; runs are then offset by a scan of their sums.
    const uint32_t per = (n + blockDim.x - 1) / blockDim.x, lo = threadIdx.x * per;
    uint32_t sum = 0;
    for (uint32_t k = lo; k < lo + per && k < n; ++k)
        if (k % chunks_cap < chunks) sum += hist[k];
    part[threadIdx.x] = sum;
    __syncthreads();
    for (uint32_t off = 1; off < blockDim.x; off <<= 1) {
        const uint32_t v = threadIdx.x >= off ? part[threadIdx.x - off] : 0u;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t acc = part[threadIdx.x] - sum;
    for (uint32_t k = lo; k < lo + per && k < n; ++k)
        if (k % chunks_cap < chunks) {
            const uint32_t v = hist[k];
            hist[k] = acc;
            acc += v;
        }
}
__global__ void __launch_bounds__(kWorkChunk) k_bin_scatter(const uint32_t* work, const uint32_t* work_count,
                                                            const uint32_t* hist, uint32_t* out) {
    __shared__ uint32_t cnt[kWorkClasses];
    const uint32_t total = *work_count, lo = blockIdx.x * kWorkChunk;
    if (lo >= total) return;
    if (threadIdx.x < kWorkClasses) cnt[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t i = lo + threadIdx.x;
    if (i < total) {
        const uint32_t item = work[i], c = item_class(item);
        const uint32_t rank = atomicAdd(cnt + c, 1u);
        out[hist[size_t(c) * gridDim.x + blockIdx.x] + rank] = item;
    }
}

// ===========================================================================
// Wave-per-pair kernel.
// ===========================================================================
constexpr int kEpaMaxVerts = 72;    // 4 seed vertices + one per iteration (64)
constexpr int kEpaMaxFaces = 512;   // EPA.cpp:206
constexpr int kEpaMaxHorizon = 256; // EPA.cpp:253
constexpr int kEdgeStride = 72;     // > 4 + 64 vertices
constexpr int kEdgeWords = (kEdgeStride * kEdgeStride + 31) / 32;

struct SP {
    f3 a, b, p;
};

// Up to four support points in named registers.  Indices are uniform
// across the wave; the accessors are select chains, so no private array
// (and no scratch memory) is involved.
struct Simplex {
    SP v0, v1, v2, v3;
};
// Component-wise value selects: no address is ever selected, so the
// simplex stays in registers.
__device__ __forceinline__ f3 sel3(bool c, f3 a, f3 b) {
    return f3{c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z};
}
__device__ __forceinline__ SP selsp(bool c, const SP& a, const SP& b) {
    return SP{sel3(c, a.a, b.a), sel3(c, a.b, b.b), sel3(c, a.p, b.p)};
}
__device__ __forceinline__ SP sget(const Simplex& s, int i) {
    return selsp(i == 0, s.v0, selsp(i == 1, s.v1, selsp(i == 2, s.v2, s.v3)));
}
__device__ __forceinline__ void sset(Simplex& s, int i, const SP& v) {
    s.v0 = selsp(i == 0, v, s.v0);
    s.v1 = selsp(i == 1, v, s.v1);
    s.v2 = selsp(i == 2, v, s.v2);
    s.v3 = selsp(i == 3, v, s.v3);
}
__device__ __forceinline__ f3 pick(bool c, f3 a, f3 b) { return sel3(c, a, b); }

struct WaveLds {
    // EPA polytope.
    float vx[kEpaMaxVerts * 9];           // SP records: a, b, p
    float fnx[kEpaMaxFaces], fny[kEpaMaxFaces], fnz[kEpaMaxFaces], fd[kEpaMaxFaces];
    uint32_t fabc[kEpaMaxFaces];          // a | b << 8 | c << 16
    uint16_t vis[kEpaMaxFaces];
    uint16_t hor[kEpaMaxHorizon];         // edge u | v << 8
    // Directed edges (u, v) of the visible faces, bit u * 72 + v (at most
    // 4 + 64 polytope vertices).
    uint32_t edges[kEdgeWords];
    // Exhaustive EPA seed candidates (EPA.cpp:116-146), 9 floats each.
    float cand16[16][9];
    // Clipping scratch (written by lane 0 only).
    float poly[2][2 * kMaxFaceVerts][3];
    float cand[8][11];
    // Face polygons of A and B (GetShapeFace, at most 4 vertices).
    float face[2][4][3];
    // Result contacts of the pair: pa, pb, normal, separation, feature id.
    float out[4][11];
};

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// ---- Wave arg-max / arg-min -------------------------------------------
// The reference scans candidates in index order keeping the first strict
// best ("v > best" from -inf, or "v < best" from +inf).  Here candidate
// index i sits in lane i % 64, slot i / 64; each lane keeps its first best
// slot, turns the value into an unsigned key whose order is the value order
// (-0 and +0 equal; NaN and the start value never win: key 0), and the wave
// takes the largest key with DPP row reductions and four lane reads.  The
// lowest index among the lanes holding that key wins, as in the reference.
// Every lane of the wave must be active.
__device__ __forceinline__ uint32_t orderable(float v) {
    const uint32_t u = v == 0.0f ? 0u : __float_as_uint(v);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ uint32_t max_key(float v) { return v > -kInf ? orderable(v) : 0u; }
__device__ __forceinline__ uint32_t min_key(float v) { return v < kInf ? ~orderable(v) : 0u; }

template <int Ctrl> __device__ __forceinline__ uint32_t dpp_max(uint32_t v) {
    const uint32_t o = uint32_t(__builtin_amdgcn_update_dpp(0, int(v), Ctrl, 0xF, 0xF, false));
    return o > v ? o : v;
}
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
    v = dpp_max<0xB1>(v);   // quad_perm [1,0,3,2]
    v = dpp_max<0x4E>(v);   // quad_perm [2,3,0,1]
    v = dpp_max<0x141>(v);  // row_half_mirror
    v = dpp_max<0x140>(v);  // row_mirror: every lane holds its 16-lane row's maximum
    const uint32_t r0 = __builtin_amdgcn_readlane(v, 0), r1 = __builtin_amdgcn_readlane(v, 16);
    const uint32_t r2 = __builtin_amdgcn_readlane(v, 32), r3 = __builtin_amdgcn_readlane(v, 48);
    const uint32_t a = r0 > r1 ? r0 : r1, b = r2 > r3 ? r2 : r3;
    return a > b ? a : b;
}
// Winning index (lane + 64 slot) of per-lane candidates (key, first best
// slot), or INT_MAX when no lane has a candidate.  Wave-uniform.
__device__ __forceinline__ int wave_arg_by_key(uint32_t key, int slot) {
    const uint32_t mk = wave_max_u32(key);
    if (mk == 0u) return 0x7FFFFFFF;
    for (int k = 0; k < 64; ++k) {
        const unsigned long long m = __ballot(key == mk && slot == k);
        if (m) return int(__builtin_ctzll(m)) + 64 * k;
    }
    return 0x7FFFFFFF;
}
__device__ __forceinline__ float read_lane(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// A shape with its transform, as seen by the wave; hull vertices are held
// four per lane (vertex lane + 64 j in slot j).
struct WShape {
    ShapeView s;
    f4q q;
    f3 t;
    f3 hv[4];
    int nverts;
};

__device__ __forceinline__ void load_wshape(WShape& w, const ShapeBuf& sb, uint32_t shape, float4 x, float4 q) {
    w.s = load_shape(sb, shape);
    w.q = lq(q);
    w.t = l3(x);
    w.nverts = 0;
    if (w.s.type == kHull) {
        w.nverts = int(w.s.hull.y);
        const int lane = lane_id();
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int vi = lane + 64 * j;
            w.hv[j] = vi < w.nverts ? l3(sb.hverts[w.s.hull.x + vi]) : f3{0.0f, 0.0f, 0.0f};
        }
    }
}

// ConvexHullShape::GetSupport (src/Geometry/Shapes.cpp:538-547) as a wave
// arg-max; other shapes are closed forms evaluated by every lane.
__device__ __forceinline__ f3 wsupport(const WShape& w, f3 d) {
    if (w.s.type != kHull) return support_closed(w.s, d);
    float best = -kInf;
    int bj = 0;
    const int lane = lane_id();
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int vi = lane + 64 * j;
        if (vi < w.nverts) {
            float v = dot(w.hv[j], d);
            if (v > best) { best = v; bj = j; }
        }
    }
    int bi = wave_arg_by_key(max_key(best), bj);
    if (bi == 0x7FFFFFFF) bi = 0;
    const int slot = bi >> 6, src = bi & 63;
    f3 mine = w.hv[3];
    if (slot == 0) mine = w.hv[0];
    else if (slot == 1) mine = w.hv[1];
    else if (slot == 2) mine = w.hv[2];
    return f3{read_lane(mine.x, src), read_lane(mine.y, src), read_lane(mine.z, src)};
}

// SupportMap::GetSupport (include/Kizuri/Collision/GJK.h:31-40).
__device__ __forceinline__ SP msupport(const WShape& A, const WShape& B, f3 d) {
    SP r;
    const f3 la = inv_rotate(A.q, d);
    const f3 lb = inv_rotate(B.q, neg(d));
    r.a = xform(A.q, A.t, wsupport(A, la));
    r.b = xform(B.q, B.t, wsupport(B, lb));
    r.p = sub(r.a, r.b);
    return r;
}

// --------------------------- GJK (src/Collision/GJK.cpp) -------------------
__device__ __forceinline__ f3 closest_on_triangle(f3 p, f3 a, f3 b, f3 c, float& u, float& v, float& w) {
    const f3 ab = sub(b, a), ac = sub(c, a), ap = sub(p, a);
    const float d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.0f && d2 <= 0.0f) { u = 1; v = 0; w = 0; return a; }
    const f3 bp = sub(p, b);
    const float d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0.0f && d4 <= d3) { u = 0; v = 1; w = 0; return b; }
    const float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
        const float t = d1 / (d1 - d3);
        u = 1 - t; v = t; w = 0;
        return add(a, mul(ab, t));
    }
    const f3 cp = sub(p, c);
    const float d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0.0f && d5 <= d6) { u = 0; v = 0; w = 1; return c; }
    const float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
        const float t = d2 / (d2 - d6);
        u = 1 - t; v = 0; w = t;
        return add(a, mul(ac, t));
    }
    const float va = d3 * d6 - d5 * d4;
    if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
        const float t = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        u = 0; v = 1 - t; w = t;
        return add(b, mul(sub(c, b), t));
    }
    const float den = 1.0f / (va + vb + vc);
    v = vb * den;
    w = vc * den;
    u = 1.0f - v - w;
    return add(add(a, mul(ab, v)), mul(ac, w));
}

__device__ __forceinline__ bool reduce_simplex(Simplex& s, int& n, f3& closest) {
    const f3 O{0.0f, 0.0f, 0.0f};
    if (n == 1) {
        closest = s.v0.p;
    } else if (n == 2) {
        const f3 a = s.v0.p, b = s.v1.p, ab = sub(b, a);
        const float t = -dot(a, ab);
        if (t <= 0.0f) { closest = a; n = 1; }
        else if (t >= len2(ab)) { s.v0 = s.v1; closest = b; n = 1; }
        else closest = add(a, mul(ab, t / len2(ab)));
    } else if (n == 3) {
        float u, v, w;
        closest = closest_on_triangle(O, s.v0.p, s.v1.p, s.v2.p, u, v, w);
        Simplex keep = s;
        int k = 0;
        if (u > 0.0f) sset(keep, k++, s.v0);
        if (v > 0.0f) sset(keep, k++, s.v1);
        if (w > 0.0f) sset(keep, k++, s.v2);
        if (k == 0) {
            // no positive weight: the reference keeps one default (zero) point
            keep.v0 = SP{f3{0.0f, 0.0f, 0.0f}, f3{0.0f, 0.0f, 0.0f}, f3{0.0f, 0.0f, 0.0f}};
            k = 1;
        }
        s.v0 = keep.v0;
        if (k > 1) s.v1 = keep.v1;
        if (k > 2) s.v2 = keep.v2;
        n = k;
    } else if (n == 4) {
        auto outside = [&](f3 p0, f3 p1, f3 p2, f3 opp) {
            f3 nn = cross(sub(p1, p0), sub(p2, p0));
            if (dot(nn, sub(opp, p0)) > 0.0f) nn = neg(nn);
            return dot(nn, sub(O, p0)) > 0.0f;
        };
        const bool oABC = outside(s.v0.p, s.v1.p, s.v2.p, s.v3.p);
        const bool oACD = outside(s.v0.p, s.v2.p, s.v3.p, s.v1.p);
        const bool oADB = outside(s.v0.p, s.v3.p, s.v1.p, s.v2.p);
        const bool oBDC = outside(s.v1.p, s.v3.p, s.v2.p, s.v0.p);
        if (!oABC && !oACD && !oADB && !oBDC) return true;
        float best = kInf;
        Simplex tri = s;
        int cnt = 0;
        // The faces ABC, ACD, ADB, BDC in this order; strict < keeps the first.
        auto consider = [&](bool out, const SP& x, const SP& y, const SP& z) {
            if (!out) return;
            float u, v, w;
            const f3 p = closest_on_triangle(O, x.p, y.p, z.p, u, v, w);
            const float d2 = len2(p);
            if (d2 < best) {
                best = d2;
                int k = 0;
                if (u > 0.0f) sset(tri, k++, x);
                if (v > 0.0f) sset(tri, k++, y);
                if (w > 0.0f) sset(tri, k++, z);
                cnt = k;
                closest = p;
            }
        };
        const Simplex o = s;
        consider(oABC, o.v0, o.v1, o.v2);
        consider(oACD, o.v0, o.v2, o.v3);
        consider(oADB, o.v0, o.v3, o.v1);
        consider(oBDC, o.v1, o.v3, o.v2);
        if (cnt > 0) s.v0 = tri.v0;
        if (cnt > 1) s.v1 = tri.v1;
        if (cnt > 2) s.v2 = tri.v2;
        n = cnt;
    }
    return false;
}

struct Gjk {
    bool intersect;
    float distance;
    f3 pa, pb, normal;
    int n;
    Simplex s;
};
// A GJK result in memory (16 x 16 B), handed from k_collide_lanes to the
// wave kernel.
struct GjkRec {
    float4 w[16];
};
static_assert(sizeof(GjkRec) == kGjkRecBytes, "GJK record size");
__device__ __forceinline__ Gjk load_gjk(const GjkRec& o) {
    Gjk g;
    const float4 w0 = o.w[0], w1 = o.w[1], w2 = o.w[2], w3 = o.w[3];
    g.intersect = w0.x != 0.0f;
    g.distance = w0.y;
    g.n = __float_as_int(w0.z);
    g.pa = f3{w1.x, w1.y, w1.z};
    g.pb = f3{w1.w, w2.x, w2.y};
    g.normal = f3{w2.z, w2.w, w3.x};
    SP* v[4] = {&g.s.v0, &g.s.v1, &g.s.v2, &g.s.v3};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const float4 a = o.w[4 + 2 * k], bq = o.w[5 + 2 * k], c = o.w[12 + k];
        v[k]->a = f3{a.x, a.y, a.z};
        v[k]->b = f3{a.w, bq.x, bq.y};
        v[k]->p = f3{bq.z, bq.w, c.x};
    }
    return g;
}

// GJK::GetClosestPoints (GJK.cpp:155-244) over a support map `ms` (d -> the
// support point of A - B).  In the wave kernel every lane runs the same
// scalar code and the support queries are wave-cooperative; in the
// thread-per-pair kernel each lane runs its own pair.
template <class MS> __device__ __forceinline__ void gjk_core(MS&& ms, float tol, Gjk& r) {
    r.intersect = false;
    r.distance = 0.0f;
    r.pa = r.pb = r.normal = f3{0.0f, 0.0f, 0.0f};
    Simplex& s = r.s;
    int n = 0;
    s.v0 = ms(f3{1.0f, 0.0f, 0.0f});
    s.v1 = s.v2 = s.v3 = s.v0;
    n = 1;
    f3 closest = s.v0.p;
    const float tol2 = tol * tol;
    for (int it = 0; it < 32; ++it) {
        if (len2(closest) <= tol2) { r.intersect = true; r.distance = 0.0f; break; }
        const f3 d = neg(closest);
        const SP w = ms(d);
        const float progress = len2(closest) - dot(closest, w.p);
        if (progress <= tol * len2(closest)) break;
        bool dup = false;
#pragma unroll
        for (int i = 0; i < 4; ++i)
            if (i < n && len2(sub(sget(s, i).p, w.p)) < tol2) dup = true;
        if (dup) break;
        sset(s, n++, w);
        if (reduce_simplex(s, n, closest)) { r.intersect = true; r.distance = 0.0f; break; }
    }
    r.n = n;
    if (!r.intersect) {
        const float dist = len(closest);
        r.distance = dist;
        if (dist < kEps) r.intersect = true;
        else r.normal = divs(neg(closest), dist);
        if (n == 2) {
            const f3 a = s.v0.p, b = s.v1.p, ab = sub(b, a);
            const float t = fclamp(-dot(a, ab) / fmax_(len2(ab), kEps), 0.0f, 1.0f);
            r.pa = add(s.v0.a, mul(sub(s.v1.a, s.v0.a), t));
            r.pb = add(s.v0.b, mul(sub(s.v1.b, s.v0.b), t));
        } else if (n == 3) {
            float u, v, w;
            closest_on_triangle(f3{0.0f, 0.0f, 0.0f}, s.v0.p, s.v1.p, s.v2.p, u, v, w);
            r.pa = add(add(mul(s.v0.a, u), mul(s.v1.a, v)), mul(s.v2.a, w));
            r.pb = add(add(mul(s.v0.b, u), mul(s.v1.b, v)), mul(s.v2.b, w));
        } else {
            r.pa = s.v0.a;
            r.pb = s.v0.b;
        }
    }
}

__device__ __forceinline__ void gjk(const WShape& A, const WShape& B, float tol, Gjk& r) {
    gjk_core([&](f3 d) { return msupport(A, B, d); }, tol, r);
}

// --------------------------- EPA (src/Collision/EPA.cpp) -------------------
__device__ __forceinline__ f3 perp(f3 v) {
    if (fabs_(v.x) <= fabs_(v.y) && fabs_(v.x) <= fabs_(v.z)) return f3{0.0f, -v.z, v.y};
    if (fabs_(v.y) <= fabs_(v.z)) return f3{-v.z, 0.0f, v.x};
    return f3{-v.y, v.x, 0.0f};
}
__device__ __forceinline__ bool origin_in_tetra(const Simplex& t) {
    const f3 O{0.0f, 0.0f, 0.0f};
    auto ok = [&](f3 a, f3 b, f3 c, f3 d) {
        f3 n = cross(sub(b, a), sub(c, a));
        if (dot(n, sub(d, a)) > 0.0f) n = neg(n);
        return dot(n, sub(O, a)) < -1.0e-9f;
    };
    return ok(t.v0.p, t.v1.p, t.v2.p, t.v3.p) && ok(t.v0.p, t.v2.p, t.v3.p, t.v1.p) &&
           ok(t.v0.p, t.v3.p, t.v1.p, t.v2.p) && ok(t.v1.p, t.v3.p, t.v2.p, t.v0.p);
}

// BuildInitialTetra (EPA.cpp:58-147).
struct WaveLds;
__device__ __forceinline__ bool initial_tetra(const WShape& A, const WShape& B, const Gjk& g, Simplex& out,
                                              WaveLds& L);
__device__ __forceinline__ bool initial_tetra_fallback(const WShape& A, const WShape& B, const Gjk& g, Simplex& out,
                                                       WaveLds& L);

__device__ __forceinline__ bool initial_tetra(const WShape& A, const WShape& B, const Gjk& g, Simplex& out,
                                              WaveLds& L) {
    const float tiny = 1.0e-12f;
    int n = 0;
    out = g.s;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        if (i >= g.n || n >= 4) continue;
        const SP gi = sget(g.s, i);
        bool dup = false;
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (j < n && len2(sub(sget(out, j).p, gi.p)) < tiny) dup = true;
        if (!dup) sset(out, n++, gi);
    }
    if (n == 0) sset(out, n++, msupport(A, B, f3{1.0f, 0.0f, 0.0f}));
    if (n == 4 && !origin_in_tetra(out)) n = 3;
    if (n == 1) {
        out.v1 = msupport(A, B, neg(out.v0.p));
        if (len2(sub(out.v1.p, out.v0.p)) > tiny) ++n;
    }
    if (n == 2) {
        const f3 ab = sub(out.v1.p, out.v0.p);
        if (len2(ab) > tiny) {
            const f3 d1 = normalized_or(perp(ab), f3{0.0f, 0.0f, 0.0f});
            const f3 d2 = normalized_or(cross(ab, d1), f3{0.0f, 0.0f, 0.0f});
            const SP k0 = out.v0, k1 = out.v1;
            out.v0 = msupport(A, B, d1);
            out.v1 = msupport(A, B, neg(d1));
            out.v2 = msupport(A, B, d2);
            out.v3 = msupport(A, B, neg(d2));
            if (origin_in_tetra(out)) return true;
            out.v0 = k0;
            out.v1 = k1;
        }
        const SP sp = msupport(A, B, perp(ab));
        if (len2(sub(sp.p, out.v0.p)) > tiny && len2(sub(sp.p, out.v1.p)) > tiny) sset(out, n++, sp);
    }
    if (n == 3) {
        f3 tn = cross(sub(out.v1.p, out.v0.p), sub(out.v2.p, out.v0.p));
        const float l = len(tn);
        if (l > tiny) {
            tn = divs(tn, l);
            const float sd = dot(tn, sub(f3{0.0f, 0.0f, 0.0f}, out.v0.p));
            const SP a = msupport(A, B, sd > 0.0f ? neg(tn) : tn);
            const SP b = msupport(A, B, sd > 0.0f ? tn : neg(tn));
            out.v3 = a;
            if (!origin_in_tetra(out)) out.v3 = b;
            if (origin_in_tetra(out)) return true;
        }
        n = 3;
    }
    if (n == 4 && origin_in_tetra(out)) return true;
    return initial_tetra_fallback(A, B, g, out, L);
}

struct Epa {
    bool valid;
    float depth;
    f3 normal, pa, pb;
};

__device__ __forceinline__ SP lds_vert(const WaveLds& L, int i) {
    const float* v = L.vx + 9 * i;
    return SP{f3{v[0], v[1], v[2]}, f3{v[3], v[4], v[5]}, f3{v[6], v[7], v[8]}};
}

// BuildFace (EPA.cpp:16-31) for vertex indices (a, b, c); may swap b and c.
__device__ __forceinline__ bool build_face(const WaveLds& L, uint32_t& a, uint32_t& b, uint32_t& c, f3& n, float& d) {
    const f3 A = lds_vert(L, int(a)).p, B = lds_vert(L, int(b)).p, C = lds_vert(L, int(c)).p;
    n = cross(sub(B, A), sub(C, A));
    const float l = len(n);
    if (l < 1.0e-12f) return false;
    n = divs(n, l);
    d = dot(n, A);
    if (d < 0.0f) {
        n = neg(n);
        d = -d;
        uint32_t t = b;
        b = c;
        c = t;
    }
    return true;
}

// FinishFace (EPA.cpp:149-172).
__device__ __forceinline__ void finish_face(const WaveLds& L, int f, Epa& r) {
    const f3 n{L.fnx[f], L.fny[f], L.fnz[f]};
    const float dist = L.fd[f];
    const uint32_t abc = L.fabc[f];
    const SP va = lds_vert(L, int(abc & 0xFF)), vb = lds_vert(L, int((abc >> 8) & 0xFF)), vc = lds_vert(L, int(abc >> 16));
    const f3 p = mul(n, dist);
    const f3 v0 = sub(vb.p, va.p), v1 = sub(vc.p, va.p), v2 = sub(p, va.p);
    const float d00 = dot(v0, v0), d01 = dot(v0, v1), d11 = dot(v1, v1);
    const float d20 = dot(v2, v0), d21 = dot(v2, v1);
    const float den = d00 * d11 - d01 * d01;
    float u, v, w;
    if (fabs_(den) < 1.0e-18f) { u = 1; v = 0; w = 0; }
    else {
        v = (d11 * d20 - d01 * d21) / den;
        w = (d00 * d21 - d01 * d20) / den;
        u = 1.0f - v - w;
    }
    r.valid = true;
    r.depth = dist;
    r.normal = neg(n);
    r.pa = add(add(mul(va.a, u), mul(vb.a, v)), mul(vc.a, w));
    r.pb = add(add(mul(va.b, u), mul(vb.b, v)), mul(vc.b, w));
}

// Wave-parallel closest alive face: strict <, first wins.
__device__ __forceinline__ int closest_face(const WaveLds& L, int nf) {
    float best = kInf;
    int bk = 0, k = 0;
    for (int f = lane_id(); f < nf; f += 64, ++k) {
        const float d = L.fd[f];
        if (d < best) { best = d; bk = k; }
    }
    return wave_arg_by_key(min_key(best), bk);
}

__device__ __forceinline__ void lds_write_vert(WaveLds& L, int i, const SP& s) {
    if (lane_id() == 0) {
        float* v = L.vx + 9 * i;
        v[0] = s.a.x; v[1] = s.a.y; v[2] = s.a.z;
        v[3] = s.b.x; v[4] = s.b.y; v[5] = s.b.z;
        v[6] = s.p.x; v[7] = s.p.y; v[8] = s.p.z;
    }
}


// Exhaustive fallback of BuildInitialTetra (EPA.cpp:116-146): candidates are
// the GJK simplex and the supports along 14 fixed directions (duplicates
// dropped), the first enclosing tetrahedron in i<j<k<l order wins.  The
// candidate list lives in LDS (written by lane 0, read by all lanes).
__device__ __forceinline__ f3 fallback_dir(int i) {
    const float s = 0.5773502692f;
    switch (i) {
        case 0: return f3{1, 0, 0};
        case 1: return f3{-1, 0, 0};
        case 2: return f3{0, 1, 0};
        case 3: return f3{0, -1, 0};
        case 4: return f3{0, 0, 1};
        case 5: return f3{0, 0, -1};
        case 6: return f3{s, s, s};
        case 7: return f3{s, s, -s};
        case 8: return f3{s, -s, s};
        case 9: return f3{-s, s, s};
        case 10: return f3{-s, -s, s};
        case 11: return f3{-s, s, -s};
        case 12: return f3{s, -s, -s};
        default: return f3{-s, -s, -s};
    }
}
__device__ __forceinline__ bool initial_tetra_fallback(const WShape& A, const WShape& B, const Gjk& g, Simplex& out,
                                                       WaveLds& L) {
    const float tiny = 1.0e-12f;
    int nc = 0;
    auto cand = [&](int i) {
        const float* v = L.cand16[i];
        return SP{f3{v[0], v[1], v[2]}, f3{v[3], v[4], v[5]}, f3{v[6], v[7], v[8]}};
    };
    auto addc = [&](const SP& sp) {
        for (int i = 0; i < nc; ++i)
            if (len2(sub(cand(i).p, sp.p)) < tiny) return;
        if (nc < 16) {
            if (lane_id() == 0) {
                float* v = L.cand16[nc];
                v[0] = sp.a.x; v[1] = sp.a.y; v[2] = sp.a.z; v[3] = sp.b.x; v[4] = sp.b.y; v[5] = sp.b.z;
                v[6] = sp.p.x; v[7] = sp.p.y; v[8] = sp.p.z;
            }
            __syncthreads();
            ++nc;
        }
    };
#pragma unroll
    for (int i = 0; i < 4; ++i)
        if (i < g.n) addc(sget(g.s, i));
    for (int i = 0; i < 14; ++i) addc(msupport(A, B, fallback_dir(i)));
    for (int i = 0; i + 3 < nc; ++i)
        for (int j = i + 1; j + 2 < nc; ++j)
            for (int k = j + 1; k + 1 < nc; ++k)
                for (int l = k + 1; l < nc; ++l) {
                    const Simplex t{cand(i), cand(j), cand(k), cand(l)};
                    const float vol = dot(cross(sub(t.v1.p, t.v0.p), sub(t.v2.p, t.v0.p)), sub(t.v3.p, t.v0.p));
                    if (fabs_(vol) < tiny) continue;
                    if (origin_in_tetra(t)) {
                        out = t;
                        return true;
                    }
                }
    return false;
}

__device__ __forceinline__ uint32_t prefix_count(unsigned long long mask) {
    // number of set bits below this lane
    return __builtin_amdgcn_mbcnt_hi(uint32_t(mask >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(mask), 0u));
}

// EPA::Compute (EPA.cpp:176-296).  All faces in [0, nf) are alive at the start
// of every iteration (dead faces are compacted away), so the alive tests of
// the reference reduce to index bounds.
__device__ __forceinline__ void epa(const WShape& A, const WShape& B, const Gjk& g, float tol, WaveLds& L, Epa& r) {
    r.valid = false;
    r.depth = 0.0f;
    r.normal = f3{0.0f, 1.0f, 0.0f};
    r.pa = r.pb = f3{0.0f, 0.0f, 0.0f};
    Simplex tet;
    if (!initial_tetra(A, B, g, tet, L)) return;
    lds_write_vert(L, 0, tet.v0);
    lds_write_vert(L, 1, tet.v1);
    lds_write_vert(L, 2, tet.v2);
    lds_write_vert(L, 3, tet.v3);
    __syncthreads();
    int nv = 4, nf = 0;
    const uint32_t kT[4][3] = {{0, 1, 2}, {0, 2, 3}, {0, 3, 1}, {1, 3, 2}};
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        uint32_t a = kT[t][0], b = kT[t][1], c = kT[t][2];
        f3 n;
        float d;
        if (build_face(L, a, b, c, n, d)) {
            if (lane_id() == 0) {
                L.fnx[nf] = n.x; L.fny[nf] = n.y; L.fnz[nf] = n.z; L.fd[nf] = d;
                L.fabc[nf] = a | (b << 8) | (c << 16);
            }
            ++nf;
        }
    }
    __syncthreads();
    if (nf < 4) return;
    const int lane = lane_id();
    for (int it = 0; it < 64; ++it) {
        const int bf = closest_face(L, nf);
        if (bf == 0x7FFFFFFF) break;
        const f3 fn{L.fnx[bf], L.fny[bf], L.fnz[bf]};
        const float fdist = L.fd[bf];
        const SP w = msupport(A, B, fn);
        const float d = dot(w.p, fn);
        if (d - fdist < tol || nv >= 256) {
            finish_face(L, bf, r);
            return;
        }
        const int newi = nv;
        lds_write_vert(L, nv, w);
        ++nv;
        // Visible faces, listed in index order.
        int nvis = 0;
        for (int base = 0; base < nf; base += 64) {
            const int f = base + lane;
            bool vis = false;
            if (f < nf) vis = dot(f3{L.fnx[f], L.fny[f], L.fnz[f]}, w.p) - L.fd[f] > tol;
            const unsigned long long m = __ballot(vis);
            if (vis) L.vis[nvis + int(prefix_count(m))] = uint16_t(f);
            nvis += __popcll(m);
        }
        __syncthreads();
        // Horizon: edges (ab, bc, ca) of each visible face, in list order,
        // that no other visible face shares reversed (EPA.cpp:253-270).  A
        // face never holds both directions of one edge, so "another visible
        // face has (v, u)" is "(v, u) is among the visible faces' edges".
        for (int wd = lane; wd < kEdgeWords; wd += 64) L.edges[wd] = 0u;
        __syncthreads();
        for (int e = lane; e < nvis; e += 64) {
            const uint32_t abc = L.fabc[L.vis[e]];
            const uint32_t fa = abc & 0xFF, fb = (abc >> 8) & 0xFF, fc = abc >> 16;
            const uint32_t k0 = fa * kEdgeStride + fb, k1 = fb * kEdgeStride + fc, k2 = fc * kEdgeStride + fa;
            atomicOr(&L.edges[k0 >> 5], 1u << (k0 & 31u));
            atomicOr(&L.edges[k1 >> 5], 1u << (k1 & 31u));
            atomicOr(&L.edges[k2 >> 5], 1u << (k2 & 31u));
        }
        __syncthreads();
        int nh = 0;
        for (int base = 0; base < nvis; base += 64) {
            const int e = base + lane;
            uint32_t eu[3] = {0, 0, 0}, ev[3] = {0, 0, 0};
            int keep = 0;
            if (e < nvis) {
                const int fi = L.vis[e];
                const uint32_t abc = L.fabc[fi];
                const uint32_t fa = abc & 0xFF, fb = (abc >> 8) & 0xFF, fc = abc >> 16;
                const uint32_t es[3][2] = {{fa, fb}, {fb, fc}, {fc, fa}};
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const uint32_t rk = es[k][1] * kEdgeStride + es[k][0];
                    const bool shared = (L.edges[rk >> 5] >> (rk & 31u)) & 1u;
                    if (!shared) {
                        // select-chain stores keep eu / ev in registers
                        if (keep == 0) { eu[0] = es[k][0]; ev[0] = es[k][1]; }
                        else if (keep == 1) { eu[1] = es[k][0]; ev[1] = es[k][1]; }
                        else { eu[2] = es[k][0]; ev[2] = es[k][1]; }
                        ++keep;
                    }
                }
            }
            // Exclusive prefix sum of `keep` (0..3) over the lanes, in order,
            // from the ballots of its two bits.
            const unsigned long long m0 = __ballot(keep & 1), m1 = __ballot(keep & 2);
            const int pos = nh + int(prefix_count(m0)) + 2 * int(prefix_count(m1));
#pragma unroll
            for (int k = 0; k < 3; ++k)
                if (k < keep && pos + k < kEpaMaxHorizon) L.hor[pos + k] = uint16_t(eu[k] | (ev[k] << 8));
            nh += __popcll(m0) + 2 * __popcll(m1);
        }
        if (nh > kEpaMaxHorizon) nh = kEpaMaxHorizon;
        __syncthreads();
        // Stable compaction of the surviving faces, chunk by chunk in order
        // (destinations never exceed sources, so in-place is safe).
        int nalive = 0;
        for (int base = 0; base < nf; base += 64) {
            const int f = base + lane;
            bool alive = false;
            float x = 0, y = 0, z = 0, dd = 0;
            uint32_t abc = 0;
            if (f < nf) {
                alive = !(dot(f3{L.fnx[f], L.fny[f], L.fnz[f]}, w.p) - L.fd[f] > tol);
                x = L.fnx[f]; y = L.fny[f]; z = L.fnz[f]; dd = L.fd[f]; abc = L.fabc[f];
            }
            const unsigned long long m = __ballot(alive);
            __syncthreads();
            if (alive) {
                const int dst = nalive + int(prefix_count(m));
                L.fnx[dst] = x; L.fny[dst] = y; L.fnz[dst] = z; L.fd[dst] = dd; L.fabc[dst] = abc;
            }
            nalive += __popcll(m);
            __syncthreads();
        }
        nf = nalive;
        // New faces (v, u, new) per horizon edge, in order, until 512 faces.
        for (int base = 0; base < nh; base += 64) {
            const int e = base + lane;
            bool ok = false;
            f3 n{0.0f, 0.0f, 0.0f};
            float dd = 0.0f;
            uint32_t a = 0, b = 0, c = 0;
            if (e < nh) {
                const uint32_t h = L.hor[e];
                a = h >> 8;       // edge v
                b = h & 0xFF;     // edge u
                c = uint32_t(newi);
                ok = build_face(L, a, b, c, n, dd);
            }
            const unsigned long long m = __ballot(ok);
            const int dst = nf + int(prefix_count(m));
            if (ok && dst < kEpaMaxFaces) {
                L.fnx[dst] = n.x; L.fny[dst] = n.y; L.fnz[dst] = n.z; L.fd[dst] = dd;
                L.fabc[dst] = a | (b << 8) | (c << 16);
            }
            nf += __popcll(m);
            if (nf > kEpaMaxFaces) nf = kEpaMaxFaces;
        }
        __syncthreads();
        if (nf >= kEpaMaxFaces && it < 63) {
            // (the cap only stops appending; iteration continues as in the reference)
        }
    }
    const int bf = closest_face(L, nf);
    if (bf != 0x7FFFFFFF) finish_face(L, bf, r);
}

// ---------------------- Polyhedral clipping (CollideShape.cpp) -------------
// GetShapeFace (CollideShape.cpp:30-98); the hull triangle scan is a wave
// arg-max over precomputed local face normals.
__device__ __forceinline__ void put3(float* dst, f3 v) {
    if (lane_id() == 0) {
        dst[0] = v.x;
        dst[1] = v.y;
        dst[2] = v.z;
    }
}
__device__ __forceinline__ bool shape_face(const WShape& w, const ShapeBuf& sb, f3 dir, float (*out)[3], int& cnt,
                                           f3& nw) {
    if (w.s.type == kBox) {
        const f3 ld = inv_rotate(w.q, dir);
        const int ax = max_index(vabs(ld));
        const float sg = comp(ld, ax) >= 0.0f ? 1.0f : -1.0f;
        const f3 he = l3(w.s.prm);
        f3 n{0.0f, 0.0f, 0.0f};
        set_comp(n, ax, sg);
        const f3 c{he.x * n.x, he.y * n.y, he.z * n.z};
        const int a0 = (ax + 1) % 3, a1 = (ax + 2) % 3;
        const float s0 = comp(he, a0), s1 = comp(he, a1);
        f3 k0 = c, k1 = c, k2 = c, k3 = c;
        set_comp(k0, a0, s0); set_comp(k0, a1, s1);
        set_comp(k1, a0, -s0); set_comp(k1, a1, s1);
        set_comp(k2, a0, -s0); set_comp(k2, a1, -s1);
        set_comp(k3, a0, s0); set_comp(k3, a1, -s1);
        put3(out[0], xform(w.q, w.t, k0));
        put3(out[1], xform(w.q, w.t, k1));
        put3(out[2], xform(w.q, w.t, k2));
        put3(out[3], xform(w.q, w.t, k3));
        cnt = 4;
        nw = rotate(w.q, n);
        return true;
    }
    if (w.s.type == kHull) {
        const f3 ld = inv_rotate(w.q, dir);
        const int nt = int(w.s.hull.w);
        float best = -kInf;
        int bk = 0, k = 0;
        for (int f = lane_id(); f < nt; f += 64, ++k) {
            const float d = dot(l3(sb.hfn[w.s.hull.z + f]), ld);
            if (d > best) { best = d; bk = k; }
        }
        const int bi = wave_arg_by_key(max_key(best), bk);
        if (bi == 0x7FFFFFFF) return false;
        const uint32_t abc = sb.htris[w.s.hull.z + bi];
        put3(out[0], xform(w.q, w.t, l3(sb.hverts[w.s.hull.x + (abc & 0xFF)])));
        put3(out[1], xform(w.q, w.t, l3(sb.hverts[w.s.hull.x + ((abc >> 8) & 0xFF)])));
        put3(out[2], xform(w.q, w.t, l3(sb.hverts[w.s.hull.x + (abc >> 16)])));
        cnt = 3;
        nw = rotate(w.q, l3(sb.hfn[w.s.hull.z + bi]));
        return true;
    }
    return false;
}

// Sutherland-Hodgman against one plane, keeping d <= 0 (CollideShape.cpp:111-126),
// across the wave: lane i takes edge (i, i + 1) of the input polygon and
// emits its start vertex when inside and the crossing point when the edge
// crosses the plane.  Ballot prefix counts give each emitted point its place
// in the serial loop's output order; that loop stops appending at
// 2 * kMaxFaceVerts points, so points placed beyond are dropped.  Returns the
// output count (wave-uniform).  The caller orders the LDS writes before the
// next reads.
__device__ __forceinline__ int clip_plane(const float (*in)[3], int nin, f3 pn, float pd, float (*out)[3]) {
    if (nin == 0) return 0;
    const int i = lane_id();
    const bool live = i < nin;
    bool keep = false, cross_ = false;
    f3 cur{0.0f, 0.0f, 0.0f}, p{0.0f, 0.0f, 0.0f};
    if (live) {
        const int j = (i + 1) % nin;
        cur = f3{in[i][0], in[i][1], in[i][2]};
        const f3 nxt{in[j][0], in[j][1], in[j][2]};
        const float dc = dot(pn, cur) - pd, dn = dot(pn, nxt) - pd;
        keep = dc <= 0.0f;
        cross_ = (dc <= 0.0f) != (dn <= 0.0f);
        if (cross_) {
            const float t = dc / (dc - dn);
            p = add(cur, mul(sub(nxt, cur), t));
        }
    }
    const unsigned long long mk = __ballot(keep), mc = __ballot(cross_);
    const int at = int(prefix_count(mk) + prefix_count(mc));
    constexpr int cap = 2 * kMaxFaceVerts;
    if (keep && at < cap) { out[at][0] = cur.x; out[at][1] = cur.y; out[at][2] = cur.z; }
    const int at1 = at + (keep ? 1 : 0);
    if (cross_ && at1 < cap) { out[at1][0] = p.x; out[at1][1] = p.y; out[at1][2] = p.z; }
    const int n = __popcll(mk) + __popcll(mc);
    return n < cap ? n : cap;
}

// One result contact into the wave's LDS output slot k (lane 0 writes).
__device__ __forceinline__ void put_contact(WaveLds& L, int k, f3 pa, f3 pb, f3 n, float sep, uint32_t feature) {
    if (lane_id() == 0) {
        float* c = L.out[k];
        c[0] = pa.x; c[1] = pa.y; c[2] = pa.z; c[3] = pb.x; c[4] = pb.y; c[5] = pb.z;
        c[6] = n.x; c[7] = n.y; c[8] = n.z; c[9] = sep; c[10] = __uint_as_float(feature);
    }
}

// CollidePolyhedral (CollideShape.cpp:131-228).  Returns the point count;
// the points are left in L.out.
__device__ __forceinline__ int collide_polyhedral(const WShape& A, const WShape& B, const ShapeBuf& sb, f3 normal,
                                                  float max_sep, WaveLds& L) {
    f3 na, nb;
    int ca = 0, cb = 0;
    const bool ha = shape_face(A, sb, neg(normal), L.face[0], ca, na);
    const bool hb = shape_face(B, sb, normal, L.face[1], cb, nb);
    __syncthreads();
    if (!ha || !hb) return 0;
    const float align_a = dot(na, neg(normal)), align_b = dot(nb, normal);
    const bool ref_a = align_a >= align_b;
    const float (*rv)[3] = L.face[ref_a ? 0 : 1];
    const float (*iv)[3] = L.face[ref_a ? 1 : 0];
    const int rc = ref_a ? ca : cb;
    const int ic = ref_a ? cb : ca;
    const f3 rn = pick(ref_a, na, nb);
    const float rd = dot(rn, f3{rv[0][0], rv[0][1], rv[0][2]});
    // The incident face, clipped by the side planes of the reference face
    // (every lane computes the planes), then the points within max_sep
    // projected onto the reference plane, at most 8 in order.
    const int lane = lane_id();
    int cur = 0, np = ic;
    if (lane < ic) { L.poly[0][lane][0] = iv[lane][0]; L.poly[0][lane][1] = iv[lane][1]; L.poly[0][lane][2] = iv[lane][2]; }
    __syncthreads();
    f3 cen{0.0f, 0.0f, 0.0f};
    for (int i = 0; i < rc; ++i) cen = add(cen, f3{rv[i][0], rv[i][1], rv[i][2]});
    cen = divs(cen, float(rc));
    for (int i = 0; i < rc; ++i) {
        const int i1 = (i + 1) % rc;
        const f3 v0{rv[i][0], rv[i][1], rv[i][2]}, v1{rv[i1][0], rv[i1][1], rv[i1][2]};
        f3 sn = cross(sub(v1, v0), rn);
        const float l = len(sn);
        if (l < 1.0e-9f) continue;
        sn = divs(sn, l);
        if (dot(sn, sub(v0, cen)) < 0.0f) sn = neg(sn);
        np = clip_plane(L.poly[cur], np, sn, dot(sn, v0), L.poly[cur ^ 1]);
        cur ^= 1;
        __syncthreads();
        if (np == 0) break;
    }
    bool keep = false;
    float sep = 0.0f;
    f3 pt{0.0f, 0.0f, 0.0f};
    if (lane < np) {
        pt = f3{L.poly[cur][lane][0], L.poly[cur][lane][1], L.poly[cur][lane][2]};
        sep = dot(rn, pt) - rd;
        keep = !(sep > max_sep);
    }
    const unsigned long long mk = __ballot(keep);
    const int k = int(prefix_count(mk));
    if (keep && k < 8) {
        const f3 proj = sub(pt, mul(rn, sep));
        const f3 pa = pick(ref_a, proj, pt), pb = pick(ref_a, pt, proj);
        float* c = L.cand[k];
        c[0] = pa.x; c[1] = pa.y; c[2] = pa.z; c[3] = pb.x; c[4] = pb.y; c[5] = pb.z;
        c[6] = normal.x; c[7] = normal.y; c[8] = normal.z; c[9] = sep; c[10] = __uint_as_float(uint32_t(lane));
    }
    const int kept = __popcll(mk);
    const int count = kept < 8 ? kept : 8;
    __syncthreads();
    const int outn = count > 4 ? 4 : count;
    // Reduction to four points: cand[j * (count - 1) / 3] (CollideShape.cpp:216-224).
    const int j = lane_id();
    if (j < outn) {
        const int src = count > 4 ? j * (count - 1) / 3 : j;
#pragma unroll
        for (int q = 0; q < 11; ++q) L.out[j][q] = L.cand[src][q];
    }
    __syncthreads();
    return outn;
}

__device__ __forceinline__ bool polyhedral(const WShape& w) { return w.s.type == kBox || w.s.type == kHull; }

// The pairs k_collide_lanes hands on, each with its GJK result:
//   * CollideSphereConvex with an intersecting proxy sphere: EPA between the
//     proxy and the other shape; no point when EPA fails (:317-330);
//   * CollideConvexGeneric, separated within maxSeparation, both shapes
//     polyhedral: clipping with the GJK normal, else the GJK point
//     (:240-249);
//   * CollideConvexGeneric, intersecting: EPA, then clipping with the EPA
//     normal for polyhedral pairs (else, or without clipped points, the
//     EPA point); when EPA fails, the GJK witness with separation 0
//     (:261-298).  The reference runs EPA a second time for a polyhedral
//     pair whose first EPA failed; EPA is a function of the shapes and the
//     GJK simplex, so that run fails again and is not repeated here.
// One EPA and one clipping call site keep the kernel's code small.
// The next item of a wave after its first (its block index).  With `next`,
// the items come from one counter shared by the grid, so the waves that drew
// cheap pairs take more of them: on mixed hulls (8-64 vertices) and piles,
// where a few EPA pairs cost many times the rest, the slowest wave otherwise
// decides the kernel's time.  Without it (the batched worlds, boxes mostly,
// of even cost, where one atomic per item on one address costs more than the
// balance gains) the items go by grid stride.  Called with the whole wave
// active (the item is wave-uniform).
__device__ __forceinline__ uint32_t next_wave_item(uint32_t item, uint32_t* next) {
    if (!next) return item + gridDim.x;
    uint32_t k = 0;
    if (threadIdx.x == 0) k = atomicAdd(next, 1u);
    return gridDim.x + __builtin_amdgcn_readfirstlane(k);
}
__global__ void __launch_bounds__(64) k_collide_wave(const uint2* pairs, const uint32_t* work,
                                                     const uint32_t* work_count, uint32_t* next, BodyBuf b,
                                                     ShapeBuf sb, ManifoldBuf m, float max_sep, const GjkRec* pre) {
    __shared__ WaveLds L;
    const uint32_t total = *work_count;
    const float tol = 1.0e-4f;
    for (uint32_t item = blockIdx.x; item < total; item = next_wave_item(item, next)) {
        const Gjk g = load_gjk(pre[item]);
        const uint32_t wv = work[item];
        const uint32_t pi = wv & kWorkPairMask;
        const bool swapped = (wv >> 31) != 0u;
        uint32_t ia = pairs[pi].x, ib = pairs[pi].y;
        if (swapped) { uint32_t t = ia; ia = ib; ib = t; }
        const uint32_t sa = meta_shape(b.meta[ia]), sbi = meta_shape(b.meta[ib]);
        const uint32_t ta = sb.type[sa], tb = sb.type[sbi];
        const bool sph_a = ta == kSphere, sph_b = !sph_a && tb == kSphere, sphere = sph_a || sph_b;
        const bool both_poly = (ta == kBox || ta == kHull) && (tb == kBox || tb == kHull);
        // GJK operands: the sphere side (as the 1e-4 proxy) first, else A, B.
        const uint32_t ix = sph_b ? ib : ia, iy = sph_b ? ia : ib;
        WShape X, Y;
        load_wshape(X, sb, sph_b ? sbi : sa, b.x[ix], b.q[ix]);
        load_wshape(Y, sb, sph_b ? sa : sbi, b.x[iy], b.q[iy]);
        const float r = X.s.prm.x;
        if (sphere) {
            X.s.type = kSphere;
            X.s.prm.x = 1.0e-4f;
        }
        int k = 0;
        bool swap_result = swapped;
        Epa e;
        e.valid = false;
        const bool run_epa = g.intersect;
        if (run_epa) epa(X, Y, g, tol, L, e);
        const bool clip = !sphere && both_poly && ((!g.intersect && len2(g.normal) > 0.0f) || (g.intersect && e.valid));
        if (clip) k = collide_polyhedral(X, Y, sb, g.intersect ? e.normal : g.normal, max_sep, L);
        if (k == 0) {
            if (sphere) {
                if (e.valid) {
                    const f3 n = neg(e.normal);
                    put_contact(L, 0, add(X.t, mul(n, r)), e.pb, n, -(e.depth + r), 0);
                    k = 1;
                    if (sph_b) swap_result = !swap_result;
                }
            } else if (!g.intersect) {
                put_contact(L, 0, g.pa, g.pb, g.normal, g.distance, 0);
                k = 1;
            } else if (e.valid) {
                put_contact(L, 0, e.pa, e.pb, e.normal, -e.depth, 0);
                k = 1;
            } else {
                put_contact(L, 0, g.pa, g.pb, len2(g.normal) > 0.0f ? g.normal : f3{0.0f, 1.0f, 0.0f}, 0.0f, 0);
                k = 1;
            }
        }
        __syncthreads();
        // Lane j writes result point j.
        const int lane = lane_id();
        if (lane < k) {
            const float* o = L.out[lane];
            Contact c;
            c.pa = f3{o[0], o[1], o[2]};
            c.pb = f3{o[3], o[4], o[5]};
            c.n = f3{o[6], o[7], o[8]};
            c.sep = o[9];
            c.feature = __float_as_uint(o[10]);
            if (swap_result) swap_contact(c);
            write_contact(m, pi, lane, c);
        }
        if (lane == 0) m.count[pi] = uint32_t(k);
        __syncthreads();
    }
}

// ===========================================================================
// Thread-per-pair GJK: one lane per pair runs the reference GJK (hull
// supports as sequential scans: the same first strict maximum as the wave
// arg-max, so the same arithmetic and results as the wave kernel).  Pairs
// whose result needs nothing more are finished here: separated beyond
// maxSeparation (no point), or one GJK point (CollideSphereConvex
// separated; CollideConvexGeneric separated with a non-polyhedral shape).
// Intersecting pairs (EPA) and separated polyhedral pairs within
// maxSeparation (clipping) go to the wave kernel's list with their GJK
// result, from which the wave kernel continues.
// ===========================================================================
struct TShape {
    ShapeView s;
    f4q q;
    f3 t;
};
__device__ __forceinline__ f3 tsupport(const TShape& w, const float4* hverts, f3 d) {
    if (w.s.type != kHull) return support_closed(w.s, d);
    float best = -kInf;
    uint32_t bi = 0;
    const float4* v = hverts + w.s.hull.x;
    for (uint32_t i = 0; i < w.s.hull.y; ++i) {
        const float x = dot(l3(v[i]), d);
        if (x > best) { best = x; bi = i; }
    }
    return l3(v[bi]);
}

__device__ __forceinline__ void store_gjk(GjkRec& o, const Gjk& g) {
    o.w[0] = make_float4(g.intersect ? 1.0f : 0.0f, g.distance, __int_as_float(g.n), 0.0f);
    o.w[1] = make_float4(g.pa.x, g.pa.y, g.pa.z, g.pb.x);
    o.w[2] = make_float4(g.pb.y, g.pb.z, g.normal.x, g.normal.y);
    o.w[3] = make_float4(g.normal.z, 0.0f, 0.0f, 0.0f);
    const SP* v[4] = {&g.s.v0, &g.s.v1, &g.s.v2, &g.s.v3};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        o.w[4 + 2 * k] = make_float4(v[k]->a.x, v[k]->a.y, v[k]->a.z, v[k]->b.x);
        o.w[5 + 2 * k] = make_float4(v[k]->b.y, v[k]->b.z, v[k]->p.x, v[k]->p.y);
        o.w[12 + k] = make_float4(v[k]->p.z, 0.0f, 0.0f, 0.0f);
    }
}

// The list is sorted by class in chunks, so neighbouring lanes mostly run
// the same supports for a similar number of GJK iterations.
__global__ void __launch_bounds__(256) k_collide_lanes(const uint2* pairs, const uint32_t* work,
                                                       const uint32_t* work_count, uint32_t* fwd,
                                                       uint32_t* fwd_count, GjkRec* fwd_gjk, BodyBuf b, ShapeBuf sb,
                                                       ManifoldBuf m, float max_sep) {
    const uint32_t item = blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= *work_count) return;
    const uint32_t wv = work[item];
    const uint32_t pi = wv & kWorkPairMask;
    const bool swapped = (wv >> 31) != 0u;
    uint32_t ia = pairs[pi].x, ib = pairs[pi].y;
    if (swapped) { uint32_t t = ia; ia = ib; ib = t; }
    // The dispatch of k_collide_wave: a sphere side goes through
    // CollideSphereConvex (GJK of a 1e-4 proxy sphere at its centre against
    // the other shape), everything else through CollideConvexGeneric.
    const ShapeView sa = load_shape(sb, meta_shape(b.meta[ia])), sbv = load_shape(sb, meta_shape(b.meta[ib]));
    const bool sph_a = sa.type == kSphere, sph_b = !sph_a && sbv.type == kSphere;
    const bool both_poly = (sa.type == kBox || sa.type == kHull) && (sbv.type == kBox || sbv.type == kHull);
    const uint32_t ix = sph_b ? ib : ia, iy = sph_b ? ia : ib;
    TShape X{sph_b ? sbv : sa, lq(b.q[ix]), l3(b.x[ix])};
    const TShape Y{sph_b ? sa : sbv, lq(b.q[iy]), l3(b.x[iy])};
    const float r = X.s.prm.x;
    if (sph_a || sph_b) {
        X.s.type = kSphere;
        X.s.prm.x = 1.0e-4f;
    }
    Gjk g;
    gjk_core(
        [&](f3 d) {
            SP sp;
            sp.a = xform(X.q, X.t, tsupport(X, sb.hverts, inv_rotate(X.q, d)));
            sp.b = xform(Y.q, Y.t, tsupport(Y, sb.hverts, inv_rotate(Y.q, neg(d))));
            sp.p = sub(sp.a, sp.b);
            return sp;
        },
        1.0e-4f, g);
    int k = -1;  // -1: to the wave kernel
    bool swap_result = swapped;
    Contact c;
    if (sph_a || sph_b) {
        if (!g.intersect) {
            const float dist = g.distance;
            const float sep = dist - r;
            if (sep > max_sep) {
                k = 0;
            } else {
                const f3 n = dist > kEps ? divs(sub(X.t, g.pb), dist) : f3{0.0f, 1.0f, 0.0f};
                c = Contact{sub(X.t, mul(n, r)), g.pb, n, sep, 0u};
                k = 1;
                if (sph_b) swap_result = !swap_result;
            }
        }
    } else if (!g.intersect) {
        if (g.distance > max_sep) {
            k = 0;
        } else if (!(both_poly && len2(g.normal) > 0.0f)) {
            c = Contact{g.pa, g.pb, g.normal, g.distance, 0u};
            k = 1;
        }
    }
    if (k < 0) {
        const uint32_t slot = atomicAdd(fwd_count, 1u);
        fwd[slot] = wv;
        store_gjk(fwd_gjk[slot], g);
        return;
    }
    if (k == 1) {
        if (swap_result) swap_contact(c);
        write_contact(m, pi, 0, c);
    }
    m.count[pi] = uint32_t(k);
}

static inline uint32_t nblocks(uint32_t n, uint32_t t) { return (n + t - 1) / t; }

// work[0, cap) -> class-major work[cap, 2 cap); chunk counts after 3 cap.
static void launch_bins(hipStream_t s, uint32_t* work, const uint32_t* work_count, uint32_t cap) {
    const uint32_t chunks = nblocks(cap, kWorkChunk);
    uint32_t* hist = work + 3 * size_t(cap);
    k_bin_count<<<chunks, kWorkChunk, 0, s>>>(work, work_count, hist);
    k_bin_scan<<<1, 1024, 0, s>>>(hist, work_count, chunks);
    k_bin_scatter<<<chunks, kWorkChunk, 0, s>>>(work, work_count, hist, work + cap);
}

// `work` holds 3 n + 16 ceil(n / 1024) entries (the list, the class-major
// list, the list handed to the wave kernel, the chunk counts), `work_count`
// three counters (the two lists, the wave kernel's next item), `gjk` n
// records; n < 2^27 (the pair index bits of an item).
void launch_narrowphase(hipStream_t s, const uint2* pairs, uint32_t n, const BodyBuf& b, const ShapeBuf& sb,
                        const ManifoldBuf& m, uint32_t* work, uint32_t* work_count, void* gjk, float max_sep,
                        int world_mode, uint32_t wave_blocks) {
    if (!n) return;
    hipMemsetAsync(work_count, 0, 3 * sizeof(uint32_t), s);
    k_collide_simple<<<nblocks(n, 256), 256, 0, s>>>(pairs, n, nullptr, b, sb, m, work, work_count, max_sep,
                                                      world_mode);
    launch_bins(s, work, work_count, n);
    GjkRec* rec = static_cast<GjkRec*>(gjk);
    k_collide_lanes<<<nblocks(n, 256), 256, 0, s>>>(pairs, work + n, work_count, work + 2 * size_t(n), work_count + 1,
                                                     rec, b, sb, m, max_sep);
    uint32_t grid = wave_blocks ? wave_blocks : (n < 65536u ? n : 65536u);
    k_collide_wave<<<grid, 64, 0, s>>>(pairs, work + 2 * size_t(n), work_count + 1, work_count + 2, b, sb, m,
                                       max_sep, rec);
}

// Capacity-sized launch with the pair count read on the device; the two
// list counters must already be zero (no memset: the launch may be
// graph-captured).  The wave kernel takes its items by grid stride.  work: 3 cap + 16 ceil(cap / 1024) entries; gjk: cap
// records; cap < 2^27.
void launch_narrowphase_counted(hipStream_t s, const uint2* pairs, uint32_t cap, const uint32_t* n_dev,
                                const BodyBuf& b, const ShapeBuf& sb, const ManifoldBuf& m, uint32_t* work,
                                uint32_t* work_count, void* gjk, float max_sep, uint32_t wave_blocks) {
    if (!cap) return;
    k_collide_simple<<<nblocks(cap, 256), 256, 0, s>>>(pairs, cap, n_dev, b, sb, m, work, work_count, max_sep, 1);
    launch_bins(s, work, work_count, cap);
    GjkRec* rec = static_cast<GjkRec*>(gjk);
    k_collide_lanes<<<nblocks(cap, 256), 256, 0, s>>>(pairs, work + cap, work_count, work + 2 * size_t(cap),
                                                       work_count + 1, rec, b, sb, m, max_sep);
    k_collide_wave<<<wave_blocks ? wave_blocks : 1024u, 64, 0, s>>>(pairs, work + 2 * size_t(cap), work_count + 1,
                                                                    nullptr, b, sb, m, max_sep, rec);
}

}  // namespace kz
