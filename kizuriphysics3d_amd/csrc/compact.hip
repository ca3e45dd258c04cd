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

}  // namespace

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
