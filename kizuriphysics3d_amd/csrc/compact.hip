// compact.hip - the constraint table of one world built on the device from
// the narrowphase output (BuildContactConstraints, PhysicsWorld.cpp:454-471):
// the non-empty manifolds in pair order, each constraint's bodies ordered by
// packed id, the sensor flag, and the points numbered consecutively.
//
// Three launches over tiles of 1024 pairs: per-tile counts, a scan of the
// tile counts (one workgroup), and the emission with a scan inside the tile.
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

}  // namespace

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
