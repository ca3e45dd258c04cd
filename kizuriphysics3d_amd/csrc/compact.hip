// compact.hip - the constraint table of one world built on the device from
// the narrowphase output (BuildContactConstraints, PhysicsWorld.cpp:454-471):
// the non-empty manifolds in pair order, each constraint's bodies ordered by
// packed id, the sensor flag, and the points numbered consecutively.
//
// Three launches over tiles of 1024 pairs: per-tile counts, a scan of the
// tile counts (one workgroup), and the emission with a scan inside the tile.
//
// Also the islands of that table (BuildIslands, :587-664), and, for worlds
// with sleeping, the active list's device side: the wake-ups of the build
// (:473-490), the island indices over the list, and UpdateSleeping's timers,
// islands and decisions (:866-951).
#include <hip/hip_runtime.h>

#include "kz_device.h"

namespace kz {

namespace {

constexpr uint32_t kTile = 1024;

// Exclusive scan of (a, b) over a workgroup of 1024 threads; tot receives the
// workgroup totals.  sh holds 2 x 16 words.
__device__ __forceinline__ uint2 tile_scan(uint32_t a, uint32_t b, uint2& tot, uint32_t* sh) {
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint32_t ia = a, ib = b;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t oa = __shfl_up(ia, off, 64), ob = __shfl_up(ib, off, 64);
        if (lane >= uint32_t(off)) {
            ia += oa;
            ib += ob;
        }
    }
    if (lane == 63) {
        sh[wave] = ia;
        sh[16 + wave] = ib;
    }
    __syncthreads();
    uint32_t ba = 0, bb = 0, ta = 0, tb = 0;
    for (uint32_t w = 0; w < blockDim.x / 64; ++w) {
        if (w < wave) {
            ba += sh[w];
            bb += sh[16 + w];
        }
        ta += sh[w];
        tb += sh[16 + w];
    }
    __syncthreads();
    tot = make_uint2(ta, tb);
    return make_uint2(ba + ia - a, bb + ib - b);
}

__global__ void __launch_bounds__(1024) k_compact_count(const uint32_t* m_count, uint32_t n, uint2* tile) {
    __shared__ uint32_t sh[32];
    const uint32_t i = blockIdx.x * kTile + threadIdx.x;
    const uint32_t k = i < n ? m_count[i] : 0u;
    uint2 tot;
    tile_scan(k ? 1u : 0u, k, tot, sh);
    if (threadIdx.x == 0) tile[blockIdx.x] = tot;
}

// Exclusive scan of the tile totals in place; totals = (constraints, points).
__global__ void __launch_bounds__(1024) k_compact_scan(uint2* tile, uint32_t tiles, uint32_t* totals) {
    __shared__ uint32_t sh[32];
    uint32_t ca = 0, cb = 0;
    for (uint32_t base = 0; base < tiles; base += kTile) {
        const uint32_t i = base + threadIdx.x;
        const uint2 v = i < tiles ? tile[i] : make_uint2(0u, 0u);
        uint2 tot;
        const uint2 e = tile_scan(v.x, v.y, tot, sh);
        if (i < tiles) tile[i] = make_uint2(ca + e.x, cb + e.y);
        ca += tot.x;
        cb += tot.y;
    }
    if (threadIdx.x == 0) {
        totals[0] = ca;
        totals[1] = cb;
    }
}

__global__ void __launch_bounds__(1024) k_compact_emit(const uint32_t* m_count, uint32_t n, const uint2* tile,
                                                       const uint2* pairs, BodyBuf b, ContactBuf c) {
    __shared__ uint32_t sh[32];
    const uint32_t i = blockIdx.x * kTile + threadIdx.x;
    const uint32_t k = i < n ? m_count[i] : 0u;
    uint2 tot;
    const uint2 e = tile_scan(k ? 1u : 0u, k, tot, sh);
    if (!k) return;
    const uint2 base = tile[blockIdx.x];
    uint32_t a = pairs[i].x, bb = pairs[i].y;
    if (b.packed[a] > b.packed[bb]) {
        const uint32_t t = a;
        a = bb;
        bb = t;
    }
    const uint32_t sensor = (meta_sensor(b.meta[a]) || meta_sensor(b.meta[bb])) ? 1u : 0u;
    const uint32_t ci = base.x + e.x;
    c.info[ci] = make_uint4(a, bb, base.y + e.y, k | (sensor << 8));
    c.pair[ci] = i;
}

// ---- Islands (BuildIslands, PhysicsWorld.cpp:587-664) -----------------------
// Lock-free union-find over the dynamic-dynamic constraints: a root is only
// ever linked under a smaller root (parent[x] <= x), so every find ends and
// each component's root is its smallest body slot.  The reference numbers
// islands by the first appearance of their root in slot order, i.e. by that
// smallest slot; island ids are therefore the exclusive count of roots below.
__device__ __forceinline__ bool dyn_body(const BodyBuf& b, uint32_t x) { return meta_motion(b.meta[x]) == 2u; }
__device__ __forceinline__ uint32_t uf_load(const uint32_t* parent, uint32_t x) {
    return __hip_atomic_load(parent + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t uf_find(const uint32_t* parent, uint32_t x) {
    uint32_t p = uf_load(parent, x);
    while (p != x) {
        x = p;
        p = uf_load(parent, x);
    }
    return x;
}
// Find with path halving: a non-root is pointed at its grandparent (an
// ancestor, so parent[x] <= x still holds; roots are never written here, so
// the linking CAS is unaffected).
__device__ __forceinline__ uint32_t uf_find_halve(uint32_t* parent, uint32_t x) {
    for (;;) {
        const uint32_t p = uf_load(parent, x);
        if (p == x) return x;
        const uint32_t gp = uf_load(parent, p);
        if (gp == p) return p;
        __hip_atomic_store(parent + x, gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        x = gp;
    }
}

__global__ void k_uf_init(uint32_t* parent, uint32_t nb) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nb) parent[i] = i;
}

__global__ void k_uf_union(ContactBuf c, const uint32_t* totals, BodyBuf b, uint32_t* parent) {
    const uint32_t ci = blockIdx.x * blockDim.x + threadIdx.x;
    if (ci >= totals[0]) return;
    const uint4 in = c.info[ci];
    if (!dyn_body(b, in.x) || !dyn_body(b, in.y)) return;
    uint32_t x = in.x, y = in.y;
    for (;;) {
        x = uf_find_halve(parent, x);
        y = uf_find_halve(parent, y);
        if (x == y) return;
        const uint32_t hi = x > y ? x : y, lo = x > y ? y : x;
        if (atomicCAS(parent + hi, hi, lo) == hi) return;
    }
}

// The same union-find for a world whose parent array fits in one
// workgroup's LDS: 4 bytes per body plus a bit per body for the dynamic
// flags.  LDS atomics and finds cost tens of cycles instead of a device-scope
// round trip, and the result is the same (each component's root is its
// smallest slot).  Writes parent[i] = root of i.
constexpr uint32_t kUfLdsBodies = 36864;
constexpr uint32_t uf_lds_bytes(uint32_t nb) { return 4u * nb + 4u * ((nb + 31u) / 32u); }
__device__ __forceinline__ uint32_t lds_find(uint32_t* par, uint32_t x) {
    for (;;) {
        const uint32_t p = __hip_atomic_load(par + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (p == x) return x;
        const uint32_t gp = __hip_atomic_load(par + p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (gp == p) return p;
        __hip_atomic_store(par + x, gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        x = gp;
    }
}
__global__ void __launch_bounds__(1024) k_uf_lds(ContactBuf c, const uint32_t* totals, BodyBuf b, uint32_t nb,
                                                 uint32_t* parent) {
    extern __shared__ uint32_t lds[];
    uint32_t* par = lds;
    uint32_t* dynbits = lds + nb;
    const uint32_t nw = (nb + 31u) / 32u;
    for (uint32_t i = threadIdx.x; i < nb; i += blockDim.x) par[i] = i;
    // Dynamic flags: one coalesced load per body, packed by a wave ballot
    // (the loop bound is a multiple of 64, so whole waves take part).
    const uint32_t nb64 = (nb + 63u) & ~63u;
    for (uint32_t i = threadIdx.x; i < nb64; i += blockDim.x) {
        const uint64_t m = __ballot(i < nb && dyn_body(b, i));
        if ((threadIdx.x & 63u) == 0) {
            dynbits[i >> 5] = uint32_t(m);
            if ((i >> 5) + 1 < nw) dynbits[(i >> 5) + 1] = uint32_t(m >> 32);
        }
    }
    __syncthreads();
    auto dyn = [&](uint32_t x) { return (dynbits[x >> 5] >> (x & 31u)) & 1u; };
    const uint32_t nc = totals[0];
    // Eight constraints per thread and round; the next round's are loaded
    // while this round's are linked.
    constexpr uint32_t R = 8;
    const uint32_t step = R * blockDim.x;
    auto load = [&](uint32_t base, uint4* in) {
#pragma unroll
        for (uint32_t k = 0; k < R; ++k) {
            const uint32_t ci = base + k * blockDim.x + threadIdx.x;
            in[k] = ci < nc ? c.info[ci] : make_uint4(0, 0, 0, 0);
        }
    };
    uint4 cur[R], nxt[R];
    load(0, cur);
    for (uint32_t base = 0; base < nc; base += step) {
        if (base + step < nc) load(base + step, nxt);
#pragma unroll
        for (uint32_t k = 0; k < R; ++k) {
            const uint32_t ci = base + k * blockDim.x + threadIdx.x;
            if (ci >= nc || !dyn(cur[k].x) || !dyn(cur[k].y)) continue;
            uint32_t x = cur[k].x, y = cur[k].y;
            for (;;) {
                x = lds_find(par, x);
                y = lds_find(par, y);
                if (x == y) break;
                const uint32_t hi = x > y ? x : y, lo = x > y ? y : x;
                if (atomicCAS(par + hi, hi, lo) == hi) break;
            }
        }
#pragma unroll
        for (uint32_t k = 0; k < R; ++k) cur[k] = nxt[k];
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nb; i += blockDim.x) parent[i] = lds_find(par, i);
}

// parent[i] = root of i; flag[i] = i is a dynamic root.
__global__ void __launch_bounds__(1024) k_uf_roots(uint32_t* parent, uint32_t nb, BodyBuf b, uint2* tile) {
    __shared__ uint32_t sh[32];
    const uint32_t i = blockIdx.x * kTile + threadIdx.x;
    uint32_t root = 0;
    if (i < nb) {
        root = uf_find(parent, i);
        parent[i] = root;
    }
    const uint32_t f = (i < nb && root == i && dyn_body(b, i)) ? 1u : 0u;
    uint2 tot;
    tile_scan(f, 0u, tot, sh);
    if (threadIdx.x == 0) tile[blockIdx.x] = tot;
}

// island[i] = island id of root i (other entries unused).
__global__ void __launch_bounds__(1024) k_island_ids(const uint32_t* parent, uint32_t nb, BodyBuf b,
                                                     const uint2* tile, uint32_t* island) {
    __shared__ uint32_t sh[32];
    const uint32_t i = blockIdx.x * kTile + threadIdx.x;
    const uint32_t f = (i < nb && parent[i] == i && dyn_body(b, i)) ? 1u : 0u;
    uint2 tot;
    const uint2 e = tile_scan(f, 0u, tot, sh);
    if (f) island[i] = tile[blockIdx.x].x + e.x;
}

// Island of each constraint: that of body A when dynamic, else of body B.
__global__ void k_con_island(ContactBuf c, const uint32_t* totals, BodyBuf b, const uint32_t* parent,
                             const uint32_t* island, uint32_t* con_island) {
    const uint32_t ci = blockIdx.x * blockDim.x + threadIdx.x;
    if (ci >= totals[0]) return;
    const uint4 in = c.info[ci];
    const uint32_t x = dyn_body(b, in.x) ? in.x : in.y;
    con_island[ci] = island[parent[x]];
}

// ---- Islands over the active list ------------------------------------------
// BuildIslands as the reference runs it over its active list: body x's
// union-find element is isl[x] (the island index the reference keeps on the
// body: compact index of an active dynamic body, the stale value of a
// sleeping one), valid below nd = sleep[kSlDynamic].  Unions join dynamic
// pairs whose two indices are valid; a constraint goes to the island of A
// when A is dynamic with a valid index, else to that of B, else to none
// (PhysicsWorld.cpp:612-654).  Roots are the smallest elements, so islands
// are numbered by their smallest element as the first-appearance scan of
// the reference numbers them.
enum : uint32_t { kSlActive = 0, kSlDynamic = 1, kSlWoken = 2, kSlGiveUp = 3 };
constexpr uint32_t kNoIndex = 0xFFFFFFFFu;

__global__ void k_uf_union_indexed(ContactBuf c, const uint32_t* totals, BodyBuf b, const uint32_t* isl,
                                   const uint32_t* sleep, uint32_t* parent) {
    const uint32_t ci = blockIdx.x * blockDim.x + threadIdx.x;
    if (ci >= totals[0]) return;
    const uint4 in = c.info[ci];
    if (!dyn_body(b, in.x) || !dyn_body(b, in.y)) return;
    const uint32_t nd = sleep[kSlDynamic];
    uint32_t x = isl[in.x], y = isl[in.y];
    if (x >= nd || y >= nd) return;
    for (;;) {
        x = uf_find_halve(parent, x);
        y = uf_find_halve(parent, y);
        if (x == y) return;
        const uint32_t hi = x > y ? x : y, lo = x > y ? y : x;
        if (atomicCAS(parent + hi, hi, lo) == hi) return;
    }
}

__global__ void __launch_bounds__(1024) k_uf_roots_indexed(uint32_t* parent, const uint32_t* sleep, uint2* tile) {
    __shared__ uint32_t sh[32];
    const uint32_t i = blockIdx.x * kTile + threadIdx.x, nd = sleep[kSlDynamic];
    uint32_t root = kNoIndex;
    if (i < nd) {
        root = uf_find(parent, i);
        parent[i] = root;
    }
    uint2 tot;
    tile_scan(root == i ? 1u : 0u, 0u, tot, sh);
    if (threadIdx.x == 0) tile[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(1024) k_island_ids_indexed(const uint32_t* parent, const uint32_t* sleep,
                                                             const uint2* tile, uint32_t* island) {
    __shared__ uint32_t sh[32];
    const uint32_t i = blockIdx.x * kTile + threadIdx.x;
    const uint32_t f = (i < sleep[kSlDynamic] && parent[i] == i) ? 1u : 0u;
    uint2 tot;
    const uint2 e = tile_scan(f, 0u, tot, sh);
    if (f) island[i] = tile[blockIdx.x].x + e.x;
}

__global__ void k_con_island_indexed(ContactBuf c, const uint32_t* totals, BodyBuf b, const uint32_t* isl,
                                     const uint32_t* sleep, const uint32_t* parent, const uint32_t* island,
                                     uint32_t* con_island) {
    const uint32_t ci = blockIdx.x * blockDim.x + threadIdx.x;
    if (ci >= totals[0]) return;
    const uint4 in = c.info[ci];
    const uint32_t nd = sleep[kSlDynamic];
    auto island_of = [&](uint32_t x) {
        if (!dyn_body(b, x)) return kNoIndex;
        const uint32_t e = isl[x];
        return e < nd ? island[parent[e]] : kNoIndex;
    };
    uint32_t is = island_of(in.x);
    if (is == kNoIndex) is = island_of(in.y);
    con_island[ci] = is;
}

// ---- Activation and sleeping ------------------------------------------------
// Exclusive count of `f` over one workgroup of 1024 threads (16 waves);
// total receives the workgroup's count.  sh holds 16 words.
__device__ __forceinline__ uint32_t block_rank(bool f, uint32_t& total, uint32_t* sh) {
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint64_t m = __ballot(f);
    const uint32_t below = uint32_t(__popcll(m & ((1ull << lane) - 1ull)));
    if (lane == 0) sh[wave] = uint32_t(__popcll(m));
    __syncthreads();
    uint32_t base = 0, tot = 0;
    for (uint32_t w = 0; w < blockDim.x / 64; ++w) {
        const uint32_t v = sh[w];
        base += w < wave ? v : 0u;
        tot += v;
    }
    __syncthreads();
    total = tot;
    return base + below;
}
__device__ __forceinline__ bool sleeper(uint32_t m) { return meta_motion(m) == 2u && meta_asleep(m); }

// BuildIslands' compact indices (PhysicsWorld.cpp:591-604): walking the
// active list, a dynamic body gets the next index, any other body none.
// One workgroup; sleep[kSlDynamic] receives the count.
__global__ void __launch_bounds__(1024) k_island_index(const uint32_t* act, uint32_t* sleep, BodyBuf b,
                                                       uint32_t* isl) {
    __shared__ uint32_t sh[16];
    const uint32_t n = sleep[kSlActive];
    uint32_t base = 0;
    for (uint32_t i0 = 0; i0 < n; i0 += blockDim.x) {
        const uint32_t i = i0 + threadIdx.x;
        const uint32_t x = i < n ? act[i] : 0u;
        const bool f = i < n && dyn_body(b, x);
        uint32_t tot;
        const uint32_t r = block_rank(f, tot, sh);
        if (i < n) isl[x] = f ? base + r : kNoIndex;
        base += tot;
    }
    if (threadIdx.x == 0) sleep[kSlDynamic] = base;
}

// The wake-ups of BuildContactConstraints (PhysicsWorld.cpp:473-490): walking
// the constraints in order, a sleeping dynamic body whose partner is moving
// (an active dynamic body, or a kinematic one with velocity) is activated,
// and counts as moving for the constraints after it.  As a fixed point: the
// step t(x) at which x moves is 0 for bodies moving from the start, and for
// a sleeping body the first constraint k (as k + 1) whose partner y has
// t(y) < k + 1.  Relaxing t(x) = min(t(x), k + 1) over all constraints until
// nothing changes never goes below the walk's value (a lowered value has a
// partner moving earlier, which the walk would have seen) and reaches it, so
// the order of the relaxation does not matter.
__global__ void k_wake_init(BodyBuf b, uint32_t nb, uint32_t* t, uint32_t* sleep) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) sleep[kSlGiveUp] = 0u;
    if (i >= nb) return;
    const uint32_t m = b.meta[i], mo = meta_motion(m);
    uint32_t v = kNoIndex;
    if (mo == 2u) {
        v = meta_asleep(m) ? kNoIndex : 0u;
    } else if (mo == 1u) {
        const float4 lv = b.v[i], av = b.w[i];
        if (lv.x * lv.x + lv.y * lv.y + lv.z * lv.z > 0.0f || av.x * av.x + av.y * av.y + av.z * av.z > 0.0f) v = 0u;
    }
    t[i] = v;
}

__device__ __forceinline__ uint32_t load_agent(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One workgroup relaxes until a round changes nothing; after `limit` rounds
// that still change something it sets the give-up flag and stops.
__global__ void __launch_bounds__(1024) k_wake_relax(ContactBuf c, const uint32_t* totals, BodyBuf b, uint32_t* t,
                                                     uint32_t* sleep, uint32_t limit) {
    __shared__ uint32_t changed;
    const uint32_t nc = totals[0];
    for (uint32_t round = 0;; ++round) {
        if (threadIdx.x == 0) changed = 0u;
        __syncthreads();
        for (uint32_t ci = threadIdx.x; ci < nc; ci += blockDim.x) {
            const uint4 in = c.info[ci];
            const bool sa = sleeper(b.meta[in.x]), sb = sleeper(b.meta[in.y]);
            if (!sa && !sb) continue;
            const uint32_t k1 = ci + 1u;
            if (sa && load_agent(t + in.y) < k1 && atomicMin(t + in.x, k1) > k1) atomicOr(&changed, 1u);
            if (sb && load_agent(t + in.x) < k1 && atomicMin(t + in.y, k1) > k1) atomicOr(&changed, 1u);
        }
        __syncthreads();
        const bool more = changed != 0u;
        __syncthreads();
        if (!more) return;
        if (round + 1u >= limit) {
            if (threadIdx.x == 0) sleep[kSlGiveUp] = 1u;
            return;
        }
    }
}

// The woken bodies onto the active list in constraint order (one constraint
// wakes at most one body: a body woken by its partner leaves the partner
// moving, so not asleep), then activated: asleep flag cleared, sleep timer 0
// (BodyManager::ActivateBody, src/Body/BodyManager.cpp:138-146).
__global__ void __launch_bounds__(1024) k_wake_append(ContactBuf c, const uint32_t* totals, BodyBuf b,
                                                      const uint32_t* t, uint32_t* act, uint32_t* sleep,
                                                      float* timer) {
    __shared__ uint32_t sh[16];
    const uint32_t nc = totals[0], n0 = sleep[kSlActive];
    uint32_t base = 0;
    for (uint32_t c0 = 0; c0 < nc; c0 += blockDim.x) {
        const uint32_t ci = c0 + threadIdx.x;
        uint32_t w = kNoIndex;
        if (ci < nc) {
            const uint4 in = c.info[ci];
            if (sleeper(b.meta[in.x]) && t[in.x] == ci + 1u) w = in.x;
            else if (sleeper(b.meta[in.y]) && t[in.y] == ci + 1u) w = in.y;
        }
        uint32_t tot;
        const uint32_t r = block_rank(w != kNoIndex, tot, sh);
        if (w != kNoIndex) act[n0 + base + r] = w;
        base += tot;
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < base; k += blockDim.x) {
        const uint32_t w = act[n0 + k];
        b.meta[w] &= ~kMetaAsleep;
        timer[w] = 0.0f;
    }
    if (threadIdx.x == 0) {
        sleep[kSlActive] = n0 + base;
        sleep[kSlWoken] = base;
    }
}

// UpdateSleeping (PhysicsWorld.cpp:866-951) over the n bodies of the active
// list.  Timers (:874-889) and the island indices set to list positions
// (:895-898), union-find elements and island minima reset.
__global__ void k_sleep_timers(const uint32_t* act, uint32_t n, BodyBuf b, float* timer, uint32_t* isl,
                               uint32_t* parent, uint32_t* tmin, float dt, float lin2, float ang2) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t x = act[i], m = b.meta[x];
    float tm = 0.0f;
    if (meta_motion(m) == 2u && meta_may_sleep(m)) {
        const float4 lv = b.v[x], av = b.w[x];
        const float ls = lv.x * lv.x + lv.y * lv.y + lv.z * lv.z, as = av.x * av.x + av.y * av.y + av.z * av.z;
        if (ls < lin2 && as < ang2) tm = timer[x] + dt;
    }
    timer[x] = tm;
    isl[x] = i;
    parent[i] = i;
    tmin[i] = __float_as_uint(1.0e18f);  // math::kBigNumber
}
// Unions over the step's contact constraints with both bodies dynamic and
// both indices below n (:900-908); sleeping bodies keep stale indices.
__global__ void k_sleep_union(ContactBuf c, const uint32_t* totals, BodyBuf b, const uint32_t* isl, uint32_t n,
                              uint32_t* parent) {
    const uint32_t ci = blockIdx.x * blockDim.x + threadIdx.x;
    if (ci >= totals[0]) return;
    const uint4 in = c.info[ci];
    if (!dyn_body(b, in.x) || !dyn_body(b, in.y)) return;
    uint32_t x = isl[in.x], y = isl[in.y];
    if (x >= n || y >= n) return;
    for (;;) {
        x = uf_find_halve(parent, x);
        y = uf_find_halve(parent, y);
        if (x == y) return;
        const uint32_t hi = x > y ? x : y, lo = x > y ? y : x;
        if (atomicCAS(parent + hi, hi, lo) == hi) return;
    }
}
// Smallest timer of each island (:921-930); timers are non-negative, so
// their bit patterns order as the values do.
__global__ void k_sleep_min(const uint32_t* act, uint32_t n, BodyBuf b, const float* timer, const uint32_t* parent,
                            uint32_t* tmin) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t x = act[i], m = b.meta[x];
    if (meta_motion(m) != 2u || !meta_may_sleep(m)) return;
    atomicMin(tmin + uf_find(parent, i), __float_as_uint(timer[x]));
}
// The decisions (:932-941): position i of the list as the loop starts holds
// the body the reference's loop reads there (its swap-removes only write
// positions it has passed), so q[i] = that body falls asleep.  Its
// velocities are zeroed and it is flagged asleep here; the host replays the
// swap-removes on its copy of the list.
__global__ void k_sleep_decide(const uint32_t* act, uint32_t n, BodyBuf b, const uint32_t* parent,
                               const uint32_t* tmin, float thr, uint8_t* q) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t x = act[i], m = b.meta[x];
    const bool d = meta_motion(m) == 2u && meta_may_sleep(m) && __uint_as_float(tmin[uf_find(parent, i)]) >= thr;
    q[i] = d ? 1u : 0u;
    if (d) {
        b.meta[x] = m | kMetaAsleep;
        b.v[x] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        b.w[x] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
}
// A new active list (kzg_set_active_bodies): flags from `active` (1 per
// body), sleeping bodies' velocities zeroed, timers 0, island indices unset.
__global__ void k_set_activity(BodyBuf b, uint32_t nb, const uint8_t* active, float* timer, uint32_t* isl) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb) return;
    const uint32_t m = b.meta[i];
    if (active[i]) {
        b.meta[i] = m & ~kMetaAsleep;
    } else {
        b.meta[i] = m | kMetaAsleep;
        b.v[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        b.w[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    timer[i] = 0.0f;
    isl[i] = kNoIndex;
}

}  // namespace

// Compact island indices of the active list d_act (length sleep[0]).
void launch_island_index(hipStream_t s, const uint32_t* act, uint32_t* sleep, const BodyBuf& b, uint32_t* isl) {
    k_island_index<<<1, 1024, 0, s>>>(act, sleep, b, isl);
}

// Islands over the active list's indices (see k_uf_union_indexed); the
// outputs as launch_islands gives them.  parent and island hold nb entries.
void launch_islands_indexed(hipStream_t s, const ContactBuf& c, uint32_t cap, const BodyBuf& b, uint32_t nb,
                            const uint32_t* isl, const uint32_t* sleep, uint32_t* parent, uint32_t* island,
                            uint2* tile, uint32_t* totals, uint32_t* con_island) {
    const uint32_t tiles = (nb + kTile - 1) / kTile;
    const uint32_t cb = (cap + 255) / 256;
    if (nb) k_uf_init<<<(nb + 255) / 256, 256, 0, s>>>(parent, nb);
    if (cb) k_uf_union_indexed<<<cb, 256, 0, s>>>(c, totals, b, isl, sleep, parent);
    if (tiles) k_uf_roots_indexed<<<tiles, kTile, 0, s>>>(parent, sleep, tile);
    k_compact_scan<<<1, kTile, 0, s>>>(tile, tiles, totals + 2);
    if (tiles) k_island_ids_indexed<<<tiles, kTile, 0, s>>>(parent, sleep, tile, island);
    if (cb) k_con_island_indexed<<<cb, 256, 0, s>>>(c, totals, b, isl, sleep, parent, island, con_island);
}

// The wake pass of a build over the constraint table (totals[0]
// constraints): woken bodies appended to act (capacity nb), sleep[0..2]
// updated; `limit` bounds the relaxation rounds.
void launch_wake(hipStream_t s, const ContactBuf& c, const uint32_t* totals, const BodyBuf& b, uint32_t nb,
                 uint32_t* t, uint32_t* act, uint32_t* sleep, float* timer, uint32_t limit) {
    if (!nb) return;
    k_wake_init<<<(nb + 255) / 256, 256, 0, s>>>(b, nb, t, sleep);
    k_wake_relax<<<1, 1024, 0, s>>>(c, totals, b, t, sleep, limit);
    k_wake_append<<<1, 1024, 0, s>>>(c, totals, b, t, act, sleep, timer);
}

// UpdateSleeping's device part over the first n entries of act; q[i] = 1
// where the body at position i falls asleep.  parent and tmin hold n entries.
void launch_sleep_decide(hipStream_t s, const ContactBuf& c, uint32_t cap, const uint32_t* totals, const BodyBuf& b,
                         const uint32_t* act, uint32_t n, float* timer, uint32_t* isl, uint32_t* parent,
                         uint32_t* tmin, uint8_t* q, float dt, float lin2, float ang2, float thr) {
    if (!n) return;
    const uint32_t nbk = (n + 255) / 256, cb = (cap + 255) / 256;
    k_sleep_timers<<<nbk, 256, 0, s>>>(act, n, b, timer, isl, parent, tmin, dt, lin2, ang2);
    if (cb) k_sleep_union<<<cb, 256, 0, s>>>(c, totals, b, isl, n, parent);
    k_sleep_min<<<nbk, 256, 0, s>>>(act, n, b, timer, parent, tmin);
    k_sleep_decide<<<nbk, 256, 0, s>>>(act, n, b, parent, tmin, thr, q);
}

void launch_set_activity(hipStream_t s, const BodyBuf& b, uint32_t nb, const uint8_t* active, float* timer,
                         uint32_t* isl) {
    if (nb) k_set_activity<<<(nb + 255) / 256, 256, 0, s>>>(b, nb, active, timer, isl);
}

// Islands of the constraint table built by launch_compact: con_island[ci]
// for ci < totals[0], and totals[2] = number of islands (every dynamic body
// belongs to one).  parent and island hold nb entries, tile ceil(nb / 1024)
// (and at least 1).  Constraints are bounded by cap.
void launch_islands(hipStream_t s, const ContactBuf& c, uint32_t cap, const BodyBuf& b, uint32_t nb,
                    uint32_t* parent, uint32_t* island, uint2* tile, uint32_t* totals, uint32_t* con_island,
                    uint32_t lds_limit) {
    const uint32_t tiles = (nb + kTile - 1) / kTile;
    const uint32_t cb = (cap + 255) / 256;
    // The LDS variant where the device lets one workgroup hold the array
    // (worlds of at most min(lds_limit, kUfLdsBodies) bodies; 0 = no limit
    // of the caller's).
    const uint32_t limit = lds_limit ? std::min(lds_limit, kUfLdsBodies) : kUfLdsBodies;
    int dev = 0, optin = 0;
    bool lds = nb && nb <= limit && hipGetDevice(&dev) == hipSuccess &&
               hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, dev) == hipSuccess &&
               uint32_t(optin) >= uf_lds_bytes(nb);
    if (lds)
        lds = hipFuncSetAttribute((const void*)k_uf_lds, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  int(uf_lds_bytes(nb))) == hipSuccess;
    if (lds) {
        k_uf_lds<<<1, 1024, uf_lds_bytes(nb), s>>>(c, totals, b, nb, parent);
    } else {
        if (nb) k_uf_init<<<(nb + 255) / 256, 256, 0, s>>>(parent, nb);
        if (cb) k_uf_union<<<cb, 256, 0, s>>>(c, totals, b, parent);
    }
    if (tiles) k_uf_roots<<<tiles, kTile, 0, s>>>(parent, nb, b, tile);
    k_compact_scan<<<1, kTile, 0, s>>>(tile, tiles, totals + 2);
    if (tiles) k_island_ids<<<tiles, kTile, 0, s>>>(parent, nb, b, tile, island);
    if (cb) k_con_island<<<cb, 256, 0, s>>>(c, totals, b, parent, island, con_island);
}

// Constraint table of n pairs into c.info / c.pair; totals[0..1] = constraints,
// points.  tile holds ceil(n / 1024) entries.
void launch_compact(hipStream_t s, const uint32_t* m_count, uint32_t n, const uint2* pairs, const BodyBuf& b,
                    const ContactBuf& c, uint2* tile, uint32_t* totals) {
    const uint32_t tiles = (n + kTile - 1) / kTile;
    if (tiles) k_compact_count<<<tiles, kTile, 0, s>>>(m_count, n, tile);
    k_compact_scan<<<1, kTile, 0, s>>>(tile, tiles, totals);
    if (tiles) k_compact_emit<<<tiles, kTile, 0, s>>>(m_count, n, tile, pairs, b, c);
}

}  // namespace kz
