// bucket.hip - the velocity sequence's bucketing by level, on the device.
//
// The host's level walk leaves, per point of the first-fit sequence, its
// point index, its unit (unit << 2 | index in the unit; units numbered in the
// order of their first points) and its level within its group.  The sweeps
// need the units grouped by level (level offsets), each unit's points in
// consecutive records (unit bounds) and every point's record (pos).  This is
// the host's bucketing (ScheduleBuilder::flatten_points) done in parallel:
//   count   units per global level (group base + level - 1)
//   scan    level offsets (tile sums, their scan, tile scans)
//   slot    each unit's place among its level's units
//   size    points per unit
//   scan    unit bounds
//   place   record of every point
// The units of group 0 (the giant island's group, ~100 levels of ~1,000 units
// each on a settled pile) come with their counts per level and their places
// in their levels from the host's walk, which counts them anyway; the other
// groups' units take theirs from atomic counters, so their order within a
// level differs from the host's.  The units of a level share no dynamic
// body, so the sweeps' results do not depend on that order.  A unit's points
// keep their sequence order.
#include "ctx_internal.h"

namespace kz {

namespace {

constexpr uint32_t kNoRecord = 0xFFFFFFFFu;
constexpr uint32_t kScanThreads = 1024;

inline uint32_t blocks_of(uint32_t n, uint32_t t) { return (n + t - 1) / t; }

__global__ void k_vb_fill(uint32_t* p, uint32_t n, uint32_t v) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// The group of point i: the first run whose end lies beyond i.
__device__ __forceinline__ uint32_t run_group(const uint2* runs, uint32_t nr, uint32_t i) {
    uint32_t lo = 0, hi = nr;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (runs[mid].x <= i) lo = mid + 1;
        else hi = mid;
    }
    return runs[lo < nr ? lo : nr - 1].y;
}

// Neighbouring points of the sequence mostly share a level, and a level's
// counter would take one atomic per unit (~1,200 on each of the ~100 levels
// of a 32k pile's giant island).  So the lanes of a wave that add to the same
// counter are served by one atomic: the first pending lane's value is
// broadcast, every lane with that value joins, and the joined lanes take
// consecutive values in lane order from the atomic's result.  Called by
// whole waves; `active` lanes take part.  Returns the lane's value.
__device__ __forceinline__ uint32_t wave_claim(uint32_t* counter, uint32_t key, bool active) {
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t got = 0;
    unsigned long long pending = __ballot(active);
    while (pending) {
        const uint32_t leader = uint32_t(__builtin_ctzll(pending));
        const uint32_t k = __builtin_amdgcn_readlane(key, leader);
        const unsigned long long group = __ballot(active && key == k) & pending;
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(counter + k, uint32_t(__popcll(group)));
        base = __builtin_amdgcn_readlane(base, leader);
        if ((group >> lane) & 1ull)
            got = base + uint32_t(__popcll(group & ((1ull << lane) - 1ull)));
        pending &= ~group;
    }
    return got;
}

// Group 0's level counts come from the host (copied into count[0 ..
// g0_levels) before this kernel); the other groups' units are counted here.
__global__ void k_vb_count(VelBucketArgs a) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = i < a.n;
    uint32_t l = 0;
    bool start = false;
    if (live) {
        l = a.group_base[run_group(a.runs, a.num_runs, i)] + a.lvl[i] - 1u;
        a.gl[i] = l;
        start = (a.uid[i] & 3u) == 0u && a.rank[i] == kNoRecord;
    }
    wave_claim(a.count, l, start);
}

// Exclusive scans: in[0..n) -> out[0..n], out[n] = the total.  A workgroup of
// 256 threads owns a tile of 1024 entries, 4 consecutive per thread (loads
// coalesced across the wave): per-tile sums, a scan of the sums in one
// workgroup, then each tile's scan plus its offset.
constexpr uint32_t kTile = 1024, kTileThreads = 256;
__device__ __forceinline__ uint32_t tile_load(const uint32_t* in, uint32_t n, uint32_t i) { return i < n ? in[i] : 0u; }
// Exclusive scan of v over the workgroup; total of all threads in `total`.
__device__ __forceinline__ uint32_t tile_scan(uint32_t v, uint32_t& total) {
    __shared__ uint32_t wsum[kTileThreads / 64];
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (uint32_t o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    uint32_t before = 0;
    total = 0;
#pragma unroll
    for (uint32_t k = 0; k < kTileThreads / 64; ++k) {
        const uint32_t ws = wsum[k];
        before += k < w ? ws : 0u;
        total += ws;
    }
    __syncthreads();
    return before + x - v;
}
__global__ void __launch_bounds__(kTileThreads) k_vb_tile_sums(const uint32_t* in, uint32_t n, uint32_t* sums) {
    const uint32_t i0 = blockIdx.x * kTile + 4u * threadIdx.x;
    const uint32_t v = tile_load(in, n, i0) + tile_load(in, n, i0 + 1) + tile_load(in, n, i0 + 2) + tile_load(in, n, i0 + 3);
    uint32_t total;
    tile_scan(v, total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}
__global__ void __launch_bounds__(kTileThreads) k_vb_tile_apply(const uint32_t* in, uint32_t n, const uint32_t* base,
                                                               uint32_t* out) {
    const uint32_t i0 = blockIdx.x * kTile + 4u * threadIdx.x;
    const uint32_t a = tile_load(in, n, i0), b = tile_load(in, n, i0 + 1), c = tile_load(in, n, i0 + 2),
                   d = tile_load(in, n, i0 + 3);
    uint32_t total;
    uint32_t run = base[blockIdx.x] + tile_scan(a + b + c + d, total);
    const uint32_t v[4] = {a, b, c, d};
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) {
        if (i0 + k < n) out[i0 + k] = run;
        run += v[k];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = base[gridDim.x];
}

// Exclusive scan of in[0..n) in one workgroup: out[0..n], out[n] = the total
// (the tile sums: a few hundred entries).
__global__ void __launch_bounds__(kScanThreads) k_vb_scan(const uint32_t* in, uint32_t n, uint32_t* out) {
    __shared__ uint32_t part[kScanThreads];
    const uint32_t t = threadIdx.x, per = (n + kScanThreads - 1) / kScanThreads;
    const uint32_t b0 = min(t * per, n), b1 = min(b0 + per, n);
    uint32_t s = 0;
    for (uint32_t i = b0; i < b1; ++i) s += in[i];
    part[t] = s;
    __syncthreads();
    for (uint32_t d = 1; d < kScanThreads; d <<= 1) {
        const uint32_t v = t >= d ? part[t - d] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint32_t run = part[t] - s;
    for (uint32_t i = b0; i < b1; ++i) {
        const uint32_t c = in[i];
        out[i] = run;
        run += c;
    }
    if (t == kScanThreads - 1) out[n] = part[t];
}

__global__ void k_vb_slot(VelBucketArgs a) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = i < a.n;
    const uint32_t u = live ? a.uid[i] : 1u;
    const bool start = live && (u & 3u) == 0u;
    const uint32_t rk = start ? a.rank[i] : kNoRecord;
    const uint32_t l = start ? a.gl[i] : 0u;
    const uint32_t k = wave_claim(a.fill, l, start && rk == kNoRecord);
    if (start) a.uslot[u >> 2] = a.off[l] + (rk == kNoRecord ? k : rk);
}

__global__ void k_vb_size(VelBucketArgs a) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const uint32_t u = a.uid[i];
    atomicMax(a.usize + a.uslot[u >> 2], (u & 3u) + 1u);
}

__global__ void k_vb_place(VelBucketArgs a) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const uint32_t u = a.uid[i];
    a.pos[a.pts[i]] = a.unit[a.uslot[u >> 2]] + (u & 3u);
}

// in[0..n) -> out[0..n] through the tiles; sums holds 2 * (tiles + 1) words.
void scan(hipStream_t s, const uint32_t* in, uint32_t n, uint32_t* out, uint32_t* sums) {
    const uint32_t tiles = blocks_of(n ? n : 1u, kTile);
    k_vb_tile_sums<<<tiles, kTileThreads, 0, s>>>(in, n, sums);
    k_vb_scan<<<1, kScanThreads, 0, s>>>(sums, tiles, sums + tiles + 1);
    k_vb_tile_apply<<<tiles, kTileThreads, 0, s>>>(in, n, sums + tiles + 1, out);
}

}  // namespace

uint32_t vel_bucket_sums_words(uint32_t levels, uint32_t units) {
    const uint32_t m = levels > units ? levels : units;
    return 2u * (blocks_of(m ? m : 1u, kTile) + 1u);
}

void launch_vel_bucket(hipStream_t s, const VelBucketArgs& a) {
    if (a.levels > a.g0_levels)
        k_vb_fill<<<blocks_of(a.levels - a.g0_levels, 256), 256, 0, s>>>(a.count + a.g0_levels,
                                                                       a.levels - a.g0_levels, 0u);
    if (a.levels) k_vb_fill<<<blocks_of(a.levels, 256), 256, 0, s>>>(a.fill, a.levels, 0u);
    if (a.units) k_vb_fill<<<blocks_of(a.units, 256), 256, 0, s>>>(a.usize, a.units, 0u);
    if (a.num_points) k_vb_fill<<<blocks_of(a.num_points, 256), 256, 0, s>>>(a.pos, a.num_points, kNoRecord);
    if (!a.n) {
        k_vb_fill<<<blocks_of(a.levels + 1, 256), 256, 0, s>>>(a.off, a.levels + 1, 0u);
        k_vb_fill<<<1, 64, 0, s>>>(a.unit, 1, 0u);
        return;
    }
    k_vb_count<<<blocks_of(a.n, 256), 256, 0, s>>>(a);
    scan(s, a.count, a.levels, a.off, a.sums);
    k_vb_slot<<<blocks_of(a.n, 256), 256, 0, s>>>(a);
    k_vb_size<<<blocks_of(a.n, 256), 256, 0, s>>>(a);
    scan(s, a.usize, a.units, a.unit, a.sums);
    k_vb_place<<<blocks_of(a.n, 256), 256, 0, s>>>(a);
}

}  // namespace kz
