// batch.hip - batched worlds (include/kzb.h): many independent worlds stepped
// as one contact table on one GPU, with a whole step captured in a HIP graph.
//
// Every world keeps its own host broadphase (the dynamic AABB tree of
// src/BroadPhase/BroadPhase.cpp, restated in host_geometry.cpp).  The
// device sees one body table (worlds concatenated, packed id = global slot,
// so each world keeps its own packed-id order) and one pair list (worlds
// concatenated, each in its broadphase order).  Pairs never cross worlds,
// so islands never do, and every world computes exactly what it computes
// alone.  What the single-world path does on the host - constraint
// compaction, islands, the first-fit SIMD batches and the level schedule
// (PhysicsWorld.cpp:454-535, 587-717) - runs here on the device, one wave
// per world, so that no host round trip splits the step.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <thread>

#include "ctx_internal.h"
#include "kzb.h"
#include "kzw.h"

namespace kz {
void launch_narrowphase_counted(hipStream_t s, const uint2* pairs, uint32_t cap, const uint32_t* n_dev,
                                const BodyBuf& b, const ShapeBuf& sb, const ManifoldBuf& m, uint32_t* work,
                                uint32_t* work_count, void* gjk, float max_sep, uint32_t wave_blocks);
void launch_build_counted(hipStream_t s, const ContactBuf& c, const ManifoldBuf& m, const BodyBuf& b,
                          const CacheBuf& cache, uint32_t cap, const uint32_t* nc_dev, int warm);
void launch_prepare_counted(hipStream_t s, const ContactBuf& c, const BodyBuf& b, uint32_t cap, const uint32_t* np_dev,
                            float dt, float rthr, float baum, float slop, float maxb, int use_simd, int use_vbias,
                            const VelRecBuf& vr, const WsRecBuf& wr, const uint32_t* vel_pos, const uint32_t* ws_pos);
void launch_writeback_counted(hipStream_t s, const VelRecBuf& vr, const WsRecBuf& wr, const ContactBuf& c,
                              const uint32_t* vel_pos, const uint32_t* ws_pos, uint32_t cap, const uint32_t* np_dev);
void launch_cache_store_counted(hipStream_t s, const ContactBuf& c, const BodyBuf& b, const CacheBuf& cache,
                                uint32_t cap, const uint32_t* nc_dev);
}  // namespace kz

namespace {

constexpr uint32_t kWorldBodies = 1024;   // LDS tables of the schedule kernel
static_assert((kWorldBodies & (kWorldBodies - 1u)) == 0 && kWorldBodies <= 0x8000u,
              "body indices are masked into the tables and carry a flag in bit 15");
constexpr uint32_t kWorldBatches = 2048;
constexpr uint32_t kWorldLevels = 4096;
constexpr uint32_t kNone = 0xFFFFFFFFu;

// Device counters, zeroed at the start of every step.
enum { kCntCons = 0, kCntPts = 1, kCntError = 2, kCntMaxVel = 3, kCntMaxWs = 4, kCntN = 8 };
enum { kErrBodies = 1u, kErrBatches = 2u, kErrLevels = 4u };

struct BatchDev {
    uint32_t W;
    const uint32_t* body_off;  // [W + 1]
    const uint32_t* pair_off;  // [W + 1], this step
    uint32_t* wc;              // [W] contact constraints per world
    uint32_t* wp;              // [W] contact points per world
    uint32_t* con_base;        // [W] first constraint of the world
    uint32_t* pt_base;         // [W] first point of the world
    uint32_t* cnt;             // [kCntN]
    uint32_t* ws_level;        // [points] warm-start / bias level, from 1
    uint32_t* vel_level;       // [points] velocity level, 0 = not swept
    uint32_t* depth_ws;        // [W]
    uint32_t* depth_vel;       // [W]
    uint32_t* vel_slots;       // [W] points of the world in the velocity sweep
    uint32_t* werr;            // [W] 1 when the world exceeded a schedule table
    uint32_t* lvb_ws;          // [W] first level of the world (each world also owns a sentinel)
    uint32_t* lvb_vel;         // [W]
    const uint32_t* order;     // [W] worlds by body count, largest first (schedule kernel)
    uint16_t* lane_pt;         // [4 * points] world-local point of each first-fit batch lane
};

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Exclusive scan over the block (blockDim a multiple of 64, at most 1024).
__device__ uint32_t block_scan(uint32_t v, uint32_t& total, uint32_t* sh) {
    const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (uint32_t o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63u) sh[wid] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t s = 0;
        for (uint32_t k = 0; k < nw; ++k) {
            const uint32_t t = sh[k];
            sh[k] = s;
            s += t;
        }
        sh[16] = s;
    }
    __syncthreads();
    const uint32_t r = sh[wid] + x - v;
    total = sh[16];
    __syncthreads();
    return r;
}

__global__ void k_batch_begin(BatchDev d, uint32_t* work_count, float* resid, uint32_t nres) {
    const uint32_t i = threadIdx.x;
    if (i < kCntN) d.cnt[i] = 0;
    if (i < 2) work_count[i] = 0;
    for (uint32_t k = i; k < nres; k += blockDim.x) resid[k] = 0.0f;
}

// Non-empty manifolds and points per world.
__global__ void __launch_bounds__(256) k_world_count(BatchDev d, const uint32_t* m_count) {
    __shared__ uint32_t sh[17];
    const uint32_t w = blockIdx.x;
    const uint32_t p0 = d.pair_off[w], p1 = d.pair_off[w + 1];
    uint32_t nc = 0, np = 0;
    for (uint32_t i = p0 + threadIdx.x; i < p1; i += blockDim.x) {
        const uint32_t k = m_count[i];
        nc += k ? 1u : 0u;
        np += k;
    }
    uint32_t tc, tp;
    block_scan(nc, tc, sh);
    block_scan(np, tp, sh);
    if (threadIdx.x == 0) {
        d.wc[w] = tc;
        d.wp[w] = tp;
    }
}

// Exclusive scan of the per-world counts (one block).
__global__ void __launch_bounds__(1024) k_world_scan(BatchDev d) {
    __shared__ uint32_t sh[17];
    uint32_t cc = 0, cp = 0;
    for (uint32_t base = 0; base < d.W; base += blockDim.x) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t vc = i < d.W ? d.wc[i] : 0u, vp = i < d.W ? d.wp[i] : 0u;
        uint32_t tc, tp;
        const uint32_t ec = block_scan(vc, tc, sh), ep = block_scan(vp, tp, sh);
        if (i < d.W) {
            d.con_base[i] = cc + ec;
            d.pt_base[i] = cp + ep;
        }
        cc += tc;
        cp += tp;
    }
    if (threadIdx.x == 0) {
        d.cnt[kCntCons] = cc;
        d.cnt[kCntPts] = cp;
    }
}

// Constraint table in pair order (BuildContactConstraints, PhysicsWorld.cpp:
// 454-471: empty manifolds skipped, bodies ordered by packed id).
// info.w = points | sensor << 8 | dynamic A << 9 | dynamic B << 10.
__global__ void __launch_bounds__(256) k_world_emit(BatchDev d, ContactBuf c, BodyBuf b, const uint2* pairs,
                                                    const uint32_t* m_count) {
    __shared__ uint32_t sh[17];
    const uint32_t w = blockIdx.x;
    const uint32_t p0 = d.pair_off[w], p1 = d.pair_off[w + 1];
    const uint32_t cb = d.con_base[w], pb = d.pt_base[w];
    uint32_t carry_c = 0, carry_p = 0;
    for (uint32_t base = p0; base < p1; base += blockDim.x) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t k = i < p1 ? m_count[i] : 0u;
        uint32_t tc, tp;
        const uint32_t ec = block_scan(k ? 1u : 0u, tc, sh), ep = block_scan(k, tp, sh);
        if (k) {
            uint32_t a = pairs[i].x, bb = pairs[i].y;
            if (b.packed[a] > b.packed[bb]) {
                const uint32_t t = a;
                a = bb;
                bb = t;
            }
            const uint32_t ma = b.meta[a], mb = b.meta[bb];
            const uint32_t flags = ((meta_sensor(ma) || meta_sensor(mb)) ? 1u : 0u) |
                                   (meta_motion(ma) == 2u ? 2u : 0u) | (meta_motion(mb) == 2u ? 4u : 0u);
            const uint32_t ci = cb + carry_c + ec;
            c.info[ci] = make_uint4(a, bb, pb + carry_p + ep, k | (flags << 8));
            c.pair[ci] = i;
        }
        carry_c += tc;
        carry_p += tp;
    }
}

// ---------------------------------------------------------------------------
// Schedule, one wave per world.
//   islands   union-find over dynamic-dynamic contacts (BuildIslands,
//             PhysicsWorld.cpp:587-664); only the partition matters here,
//             because islands share no body and levels follow per-body order
//   ws        warm-start / bias levels in constraint order (SolveIsland
//             :672-674, 736-742; sensors included)
//   batches   the reference's first fit (:677-717): a point goes to the
//             first batch of its island with fewer than four lanes and no
//             lane sharing body A or B (static bodies included), else to a
//             new batch.  The wave scans 64 batches per step from the
//             island's first non-full batch.
//   vel       velocity levels in batch order
// A level is 1 + the largest last level of the point's dynamic bodies.
// ---------------------------------------------------------------------------
struct alignas(16) SchedLds {
    uint32_t parent[kWorldBodies];
    uint16_t last_ws[kWorldBodies], last_vel[kWorldBodies], hint[kWorldBodies];
    uint32_t dyn[kWorldBodies / 32];
    uint4 stage[64];
    uint16_t ba[kWorldBatches * 4], bb[kWorldBatches * 4];  // 8-byte aligned: one quad per batch
    uint16_t bm[kWorldBatches];  // lanes used | island root << 3
};
static_assert(offsetof(SchedLds, ba) % 8 == 0 && offsetof(SchedLds, bb) % 8 == 0, "lane quads are read as 8 bytes");

__device__ __forceinline__ bool is_dyn(const SchedLds& L, uint32_t x) { return (L.dyn[x >> 5] >> (x & 31u)) & 1u; }
__device__ __forceinline__ uint32_t lds_find(const volatile uint32_t* parent, uint32_t x) {
    uint32_t p = parent[x];
    while (p != x) {
        x = p;
        p = parent[x];
    }
    return x;
}

// One wave per workgroup; workgroups take the worlds largest first so that
// the small worlds fill the tail.
__global__ void __launch_bounds__(64) k_world_schedule(BatchDev d, ContactBuf c, BodyBuf b, int use_simd) {
    __shared__ SchedLds L;
    const uint32_t lane = threadIdx.x;
    const uint32_t w = d.order[blockIdx.x];
    const uint32_t b0 = d.body_off[w], nb = d.body_off[w + 1] - b0;
    const uint32_t c0 = d.con_base[w], nc = d.wc[w], p0 = d.pt_base[w], np = d.wp[w];
    uint32_t err = 0;
    if (nb > kWorldBodies || np > 65535u) err = kErrBodies;
    if (!err) {
        for (uint32_t x = lane; x < nb; x += 64) {
            L.parent[x] = x;
            L.last_ws[x] = 0;
            L.last_vel[x] = 0;
            L.hint[x] = 0xFFFFu;
        }
        for (uint32_t k = lane; k < kWorldBodies / 32; k += 64) L.dyn[k] = 0;
        wave_sync();
        for (uint32_t x = lane; x < nb; x += 64)
            if (meta_motion(b.meta[b0 + x]) == 2u) atomicOr(&L.dyn[x >> 5], 1u << (x & 31u));
        wave_sync();
        for (uint32_t ci = lane; ci < nc; ci += 64) {
            const uint4 in = c.info[c0 + ci];
            uint32_t x = in.x - b0, y = in.y - b0;
            if (!is_dyn(L, x) || !is_dyn(L, y)) continue;
            for (;;) {
                x = lds_find(L.parent, x);
                y = lds_find(L.parent, y);
                if (x == y) break;
                const uint32_t hi = x > y ? x : y, lo = x > y ? y : x;
                if (atomicCAS(&L.parent[hi], hi, lo) == hi) break;
            }
        }
        wave_sync();
        for (uint32_t x = lane; x < nb; x += 64) L.parent[x] = lds_find(L.parent, x);
        wave_sync();
    }
    // Warm-start / bias levels, constraint order.
    uint32_t depth_ws = 0, slots = 0;
    for (uint32_t base = 0; !err && base < nc; base += 64) {
        const uint32_t cnt = nc - base < 64u ? nc - base : 64u;
        if (lane < cnt) L.stage[lane] = c.info[c0 + base + lane];
        wave_sync();
        if (lane == 0) {
            for (uint32_t j = 0; j < cnt; ++j) {
                const uint4 in = L.stage[j];
                const uint32_t x = in.x - b0, y = in.y - b0, k = in.w & 0xFFu, first = in.z - p0;
                const bool sensor = (in.w >> 8) & 1u;
                const bool dx = (in.w >> 9) & 1u, dy = (in.w >> 10) & 1u;
                // The points of a constraint follow each other: the first
                // takes 1 + the bodies' last level, each next one the level
                // after (its bodies were last updated by the point before),
                // capped at kWorldLevels; with no dynamic body every point
                // takes level 1.  Both entries are read (a static body's is
                // never used) so that the reads issue together.
                const uint32_t rx = L.last_ws[x], ry = L.last_ws[y];
                const uint32_t lx = dx ? rx : 0u, ly = dy ? ry : 0u;
                const uint32_t l0 = 1u + (lx > ly ? lx : ly);
                uint32_t l = l0;
                for (uint32_t q = 0; q < k; ++q) {
                    l = (dx || dy) ? l0 + q : 1u;
                    if (l > kWorldLevels) {
                        err = kErrLevels;
                        l = kWorldLevels;
                    }
                    d.ws_level[p0 + first + q] = l;
                    if (use_simd && sensor) d.vel_level[p0 + first + q] = 0;
                }
                if (k) {
                    L.last_ws[x] = uint16_t(l);  // a static body's entry is never read as a level
                    L.last_ws[y] = uint16_t(l);
                    depth_ws = l > depth_ws ? l : depth_ws;
                }
                slots += (use_simd && sensor) ? 0u : k;
            }
        }
        err = __shfl(err, 0);
        wave_sync();
    }
    // First-fit batches (SIMD solver).  A batch's lanes take points of the
    // world only, so its lane table fits in 4 entries per point.
    uint32_t nbat = 0;
    uint16_t* lane_pt = d.lane_pt + 4 * size_t(p0);
    // The hint (first non-full batch) of the island of the last point placed,
    // kept in a register; L.hint holds every island's.
    uint32_t hint_isl = kNone, hint_val = 0xFFFFu;
    if (use_simd) {
        for (uint32_t base = 0; !err && base < nc; base += 64) {
            const uint32_t cnt = nc - base < 64u ? nc - base : 64u;
            if (lane < cnt) L.stage[lane] = c.info[c0 + base + lane];
            wave_sync();
            for (uint32_t j = 0; !err && j < cnt; ++j) {
                const uint4 in = L.stage[j];
                if ((in.w >> 8) & 1u) continue;  // sensors are not swept (PhysicsWorld.cpp:683)
                const uint32_t x = in.x - b0, y = in.y - b0, k = in.w & 0xFFu, first = in.z - p0;
                const bool dx = (in.w >> 9) & 1u;
                const uint32_t isl = dx ? L.parent[x] : L.parent[y];
                // Lane entries: body | dynamic << 15.
                const uint32_t ex = x | (dx ? 0x8000u : 0u), ey = y | (((in.w >> 10) & 1u) ? 0x8000u : 0u);
                for (uint32_t q = 0; q < k; ++q) {
                    const uint32_t h = isl == hint_isl ? hint_val : uint32_t(L.hint[isl]);
                    uint32_t found = kNone, open = kNone, mfound = 0;
                    for (uint32_t s = (h == 0xFFFFu ? nbat : h); s < nbat; s += 64) {
                        const uint32_t jb = s + lane;
                        bool mine = false, fit = false;
                        uint32_t m = 0;
                        if (jb < nbat) {
                            // The batch's lane count and both lane quads at
                            // once (8 bytes each); lanes past the count hold
                            // stale entries and are masked.
                            m = L.bm[jb];
                            const uint2 qa = reinterpret_cast<const uint2*>(L.ba)[jb];
                            const uint2 qb = reinterpret_cast<const uint2*>(L.bb)[jb];
                            const uint32_t n = m & 7u;
                            mine = (m >> 3) == isl && n < 4u;
                            const uint32_t ea[4] = {qa.x & 0x7FFFu, (qa.x >> 16) & 0x7FFFu, qa.y & 0x7FFFu,
                                                    (qa.y >> 16) & 0x7FFFu};
                            const uint32_t eb[4] = {qb.x & 0x7FFFu, (qb.x >> 16) & 0x7FFFu, qb.y & 0x7FFFu,
                                                    (qb.y >> 16) & 0x7FFFu};
                            bool clash = false;
#pragma unroll
                            for (uint32_t t = 0; t < 4; ++t)
                                clash |= t < n && (ea[t] == x || ea[t] == y || eb[t] == x || eb[t] == y);
                            fit = mine && !clash;
                        }
                        const unsigned long long om = __ballot(mine), fm = __ballot(fit);
                        if (open == kNone && om) open = s + uint32_t(__ffsll(om)) - 1u;
                        if (fm) {
                            const uint32_t src = uint32_t(__ffsll(fm)) - 1u;
                            found = s + src;
                            mfound = __builtin_amdgcn_readlane(m, src);
                            break;
                        }
                    }
                    if (found == kNone) {
                        if (nbat >= kWorldBatches) {
                            err = kErrBatches;
                            break;
                        }
                        found = nbat++;
                        mfound = isl << 3;
                        if (open == kNone) open = found;
                    }
                    wave_sync();
                    if (lane == 0) {
                        const uint32_t n = mfound & 7u;
                        L.ba[found * 4 + n] = uint16_t(ex);
                        L.bb[found * 4 + n] = uint16_t(ey);
                        lane_pt[found * 4 + n] = uint16_t(first + q);
                        L.bm[found] = uint16_t(mfound + 1u);
                        L.hint[isl] = uint16_t(open);
                    }
                    hint_isl = isl;
                    hint_val = open;
                    wave_sync();
                }
            }
            wave_sync();
        }
    }
    // Velocity levels in batch order; the lane points come through LDS,
    // 128 batches at a time.
    uint32_t depth_vel = 0;
    uint16_t* stage_pt = reinterpret_cast<uint16_t*>(L.stage);
    for (uint32_t jb0 = 0; use_simd && !err && jb0 < nbat; jb0 += 128) {
        const uint32_t nj = nbat - jb0 < 128u ? nbat - jb0 : 128u;
        wave_sync();
        for (uint32_t k = lane; k < 4 * nj; k += 64)
            if ((L.bm[jb0 + k / 4] & 7u) > k % 4) stage_pt[k] = lane_pt[4 * jb0 + k];
        wave_sync();
        if (lane != 0) continue;
        // The lanes of a batch share no body, so its points read and write
        // disjoint level entries: all four are taken at once.
        for (uint32_t jb = jb0; jb < jb0 + nj; ++jb) {
            const uint32_t n = L.bm[jb] & 7u;
            const uint2 qa = reinterpret_cast<const uint2*>(L.ba)[jb];
            const uint2 qb = reinterpret_cast<const uint2*>(L.bb)[jb];
            const uint2 qp = reinterpret_cast<const uint2*>(stage_pt)[jb - jb0];
            const uint32_t xs[4] = {qa.x & 0xFFFFu, qa.x >> 16, qa.y & 0xFFFFu, qa.y >> 16};
            const uint32_t ys[4] = {qb.x & 0xFFFFu, qb.x >> 16, qb.y & 0xFFFFu, qb.y >> 16};
            const uint32_t ps[4] = {qp.x & 0xFFFFu, qp.x >> 16, qp.y & 0xFFFFu, qp.y >> 16};
            // Every entry is read, a static body's too (its level is written
            // and never used); entries past the lane count hold stale values,
            // masked into the table and never used: no read waits on a flag.
            uint32_t lv[4];
#pragma unroll
            for (uint32_t t = 0; t < 4; ++t) {
                const uint32_t rx = L.last_vel[xs[t] & (kWorldBodies - 1u)];
                const uint32_t ry = L.last_vel[ys[t] & (kWorldBodies - 1u)];
                const uint32_t lx = (xs[t] >> 15) ? rx : 0u, ly = (ys[t] >> 15) ? ry : 0u;
                lv[t] = 1u + (lx > ly ? lx : ly);
            }
#pragma unroll
            for (uint32_t t = 0; t < 4; ++t) {
                if (t >= n) continue;
                uint32_t l = lv[t];
                if (l > kWorldLevels) {
                    err = kErrLevels;
                    l = kWorldLevels;
                }
                L.last_vel[xs[t] & (kWorldBodies - 1u)] = uint16_t(l);
                L.last_vel[ys[t] & (kWorldBodies - 1u)] = uint16_t(l);
                d.vel_level[p0 + ps[t]] = l;
                depth_vel = l > depth_vel ? l : depth_vel;
            }
        }
    }
    err = __shfl(err, 0);
    if (lane == 0) {
        if (err) {
            depth_ws = depth_vel = 0;
            atomicOr(d.cnt + kCntError, err);
        }
        if (!use_simd) {
            depth_vel = depth_ws;
            slots = np;
        }
        d.werr[w] = err ? 1u : 0u;
        d.depth_ws[w] = depth_ws;
        d.depth_vel[w] = depth_vel;
        d.vel_slots[w] = err ? 0u : slots;
        atomicMax(d.cnt + kCntMaxWs, depth_ws);
        atomicMax(d.cnt + kCntMaxVel, depth_vel);
    }
}

// Level bases: each world owns depth + 1 entries of the level-offset arrays
// (its levels and its own end sentinel).
__global__ void __launch_bounds__(1024) k_level_scan(BatchDev d, uint2* ws_grp, uint2* vel_grp) {
    __shared__ uint32_t sh[17];
    uint32_t cw = 0, cv = 0;
    for (uint32_t base = 0; base < d.W; base += blockDim.x) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t dw = i < d.W ? d.depth_ws[i] : 0u, dv = i < d.W ? d.depth_vel[i] : 0u;
        uint32_t tw, tv;
        const uint32_t ew = block_scan(i < d.W ? dw + 1u : 0u, tw, sh);
        const uint32_t ev = block_scan(i < d.W ? dv + 1u : 0u, tv, sh);
        if (i < d.W) {
            d.lvb_ws[i] = cw + ew;
            d.lvb_vel[i] = cv + ev;
            ws_grp[i] = make_uint2(cw + ew, cw + ew + dw);
            vel_grp[i] = make_uint2(cv + ev, cv + ev + dv);
        }
        cw += tw;
        cv += tv;
    }
}

// Points bucketed by level: record positions and level offsets.  Points of
// one level share no dynamic body, so their order within the level is free.
__global__ void __launch_bounds__(256) k_world_bucket(BatchDev d, uint32_t* ws_off, uint32_t* vel_off,
                                                      uint32_t* ws_pos, uint32_t* vel_pos, int use_simd) {
    __shared__ uint32_t cws[kWorldLevels], cvel[kWorldLevels];
    __shared__ uint32_t sh[17];
    const uint32_t w = blockIdx.x, t = threadIdx.x, nt = blockDim.x;
    const uint32_t p0 = d.pt_base[w], np = d.wp[w];
    const uint32_t dw = d.depth_ws[w], dv = d.depth_vel[w];
    const uint32_t* vlev = use_simd ? d.vel_level : d.ws_level;
    if (d.werr[w]) {  // not swept; every point still gets a record slot
        for (uint32_t p = t; p < np; p += nt) {
            ws_pos[p0 + p] = p0 + p;
            vel_pos[p0 + p] = kNone;
        }
        if (t == 0) {
            ws_off[d.lvb_ws[w]] = p0 + np;
            vel_off[d.lvb_vel[w]] = p0;
        }
        return;
    }
    for (uint32_t l = t; l < dw; l += nt) cws[l] = 0;
    for (uint32_t l = t; l < dv; l += nt) cvel[l] = 0;
    __syncthreads();
    for (uint32_t p = t; p < np; p += nt) {
        atomicAdd(&cws[d.ws_level[p0 + p] - 1u], 1u);
        const uint32_t lv = vlev[p0 + p];
        if (lv) atomicAdd(&cvel[lv - 1u], 1u);
    }
    __syncthreads();
    uint32_t carry = 0;
    for (uint32_t base = 0; base < dw; base += nt) {
        const uint32_t l = base + t;
        const uint32_t v = l < dw ? cws[l] : 0u;
        uint32_t tot;
        const uint32_t e = block_scan(v, tot, sh);
        if (l < dw) {
            cws[l] = carry + e;
            ws_off[d.lvb_ws[w] + l] = p0 + carry + e;
        }
        carry += tot;
    }
    if (t == 0) ws_off[d.lvb_ws[w] + dw] = p0 + carry;
    carry = 0;
    for (uint32_t base = 0; base < dv; base += nt) {
        const uint32_t l = base + t;
        const uint32_t v = l < dv ? cvel[l] : 0u;
        uint32_t tot;
        const uint32_t e = block_scan(v, tot, sh);
        if (l < dv) {
            cvel[l] = carry + e;
            vel_off[d.lvb_vel[w] + l] = p0 + carry + e;
        }
        carry += tot;
    }
    if (t == 0) vel_off[d.lvb_vel[w] + dv] = p0 + carry;
    __syncthreads();
    for (uint32_t p = t; p < np; p += nt) {
        ws_pos[p0 + p] = p0 + atomicAdd(&cws[d.ws_level[p0 + p] - 1u], 1u);
        const uint32_t lv = vlev[p0 + p];
        vel_pos[p0 + p] = lv ? p0 + atomicAdd(&cvel[lv - 1u], 1u) : kNone;
    }
}

inline uint32_t cdiv(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

}  // namespace

// ===========================================================================
// Host side.
// ===========================================================================
struct kzb_batch {
    kz_world_desc settings{};
    std::string err;
    kzg_ctx* ctx = nullptr;
    struct World {
        kzw_world* w = nullptr;
        std::vector<kz_shape_desc> shapes;
        std::vector<kz_body_desc> bodies;
        std::vector<uint32_t> pairs;  // local slots, this step
    };
    std::vector<World> worlds;
    std::vector<uint32_t> body_off{0};
    uint32_t nbodies = 0;
    bool tables_dirty = true;
    std::vector<uint32_t> pair_off;
    // device-side batch arrays
    DevVec<uint32_t> world_order;
    DevVec<uint16_t> lane_pt;
    DevVec<uint32_t> d_body_off, d_pair_off, wc, wp, con_base, pt_base, cnt, ws_level, vel_level, depth_ws,
        depth_vel, vel_slots, werr, lvb_ws, lvb_vel;
    Pinned h_pairs, h_out;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool graph_valid = false;
    uint32_t cap_pairs = 0;
    uint32_t cache_used = 0;
    kzb_config cfg{};
    kzb_stats stats{};
    hipEvent_t e0 = nullptr, e1 = nullptr, emid = nullptr;
    hipEvent_t phase_ev[7] = {};  // direct launches: phase boundaries

    BatchDev dev() {
        return BatchDev{uint32_t(worlds.size()), d_body_off.p, d_pair_off.p, wc.p, wp.p, con_base.p, pt_base.p, cnt.p,
                        ws_level.p, vel_level.p, depth_ws.p, depth_vel.p, vel_slots.p, werr.p, lvb_ws.p, lvb_vel.p,
                        world_order.p, lane_pt.p};
    }
};

namespace {

int bfail(kzb_batch* b, int code, const std::string& m) {
    if (b) b->err = m;
    return code;
}

#define KZB_HIP(b, call)                                                                              \
    do {                                                                                              \
        hipError_t e_ = (call);                                                                       \
        if (e_ != hipSuccess)                                                                         \
            return bfail(b, KZG_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_));      \
    } while (0)

void drop_graph(kzb_batch* B) {
    if (B->exec) hipGraphExecDestroy(B->exec);
    if (B->graph) hipGraphDestroy(B->graph);
    B->exec = nullptr;
    B->graph = nullptr;
    B->graph_valid = false;
}

// Body and shape tables of every world, concatenated.
int upload_tables(kzb_batch* B) {
    kzg_ctx* c = B->ctx;
    std::vector<kz_shape_desc> sd;
    std::vector<kzg_hull_range> hr;
    std::vector<float> hv;
    std::vector<uint32_t> ht;
    std::vector<kzg_body> bodies;
    B->body_off.assign(1, 0);
    for (auto& W : B->worlds) {
        const uint32_t shape0 = uint32_t(sd.size());
        for (uint32_t s = 0; s < W.shapes.size(); ++s) {
            kz_shape_desc d = W.shapes[s];
            kzg_hull_range r{0, 0, 0, 0};
            if (d.type == KZ_SHAPE_CONVEX_HULL) {
                uint32_t nv = 0, nt = 0;
                kzw_get_hull(W.w, s, nullptr, nullptr, &nv, &nt);
                std::vector<float> v(3 * size_t(nv));
                std::vector<uint32_t> t(3 * size_t(nt));
                kzw_get_hull(W.w, s, v.data(), t.data(), nullptr, nullptr);
                r = kzg_hull_range{uint32_t(hv.size() / 3), nv, uint32_t(ht.size() / 3), nt};
                hv.insert(hv.end(), v.begin(), v.end());
                ht.insert(ht.end(), t.begin(), t.end());
            }
            sd.push_back(d);
            hr.push_back(r);
        }
        const uint32_t n = uint32_t(W.bodies.size());
        std::vector<kz_body_state> st(n);
        std::vector<float> mass(10 * size_t(n));
        kzw_get_states(W.w, st.data(), n);
        kzw_get_mass(W.w, mass.data(), n);
        for (uint32_t i = 0; i < n; ++i) {
            const kz_body_desc& d = W.bodies[i];
            kzg_body o{};
            o.packed_id = uint32_t(bodies.size());
            o.shape = shape0 + d.shape;
            o.motion_type = d.motion_type;
            o.is_sensor = d.is_sensor;
            std::memcpy(o.position, st[i].position, sizeof(o.position));
            std::memcpy(o.rotation, st[i].rotation, sizeof(o.rotation));
            std::memcpy(o.linear_velocity, st[i].linear_velocity, sizeof(o.linear_velocity));
            std::memcpy(o.angular_velocity, st[i].angular_velocity, sizeof(o.angular_velocity));
            o.inverse_mass = mass[10 * i];
            std::memcpy(o.inverse_inertia_local, &mass[10 * i + 1], sizeof(float) * 9);
            o.friction = d.friction;
            o.restitution = d.restitution;
            o.linear_damping = d.linear_damping;
            o.angular_damping = d.angular_damping;
            o.gravity_factor = d.gravity_factor;
            o.max_linear_velocity = d.max_linear_velocity;
            o.max_angular_velocity = d.max_angular_velocity;
            bodies.push_back(o);
        }
        B->body_off.push_back(uint32_t(bodies.size()));
    }
    int rc = kzg_upload_shapes(c, sd.data(), uint32_t(sd.size()), hr.data(), hv.data(), uint32_t(hv.size() / 3),
                               ht.data(), uint32_t(ht.size() / 3));
    if (rc) return bfail(B, rc, kzg_last_error(c));
    rc = kzg_upload_bodies(c, bodies.data(), uint32_t(bodies.size()));
    if (rc) return bfail(B, rc, kzg_last_error(c));
    B->nbodies = uint32_t(bodies.size());
    const size_t W = B->worlds.size();
    if ((rc = upload(c, B->d_body_off, B->body_off.data(), W + 1))) return bfail(B, rc, kzg_last_error(c));
    std::vector<uint32_t> order(W);
    for (uint32_t i = 0; i < W; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        return B->body_off[a + 1] - B->body_off[a] > B->body_off[b + 1] - B->body_off[b];
    });
    KZB_HIP(B, B->world_order.ensure(std::max<size_t>(W, 1)));
    if (W && (rc = upload(c, B->world_order, order.data(), W))) return bfail(B, rc, kzg_last_error(c));
    for (auto* v : {&B->wc, &B->wp, &B->con_base, &B->pt_base, &B->depth_ws, &B->depth_vel, &B->vel_slots, &B->werr,
                    &B->lvb_ws, &B->lvb_vel})
        KZB_HIP(B, v->ensure(std::max<size_t>(W, 1)));
    KZB_HIP(B, B->d_pair_off.ensure(W + 1));
    KZB_HIP(B, B->cnt.ensure(kCntN));
    KZB_HIP(B, c->vel_grp.ensure(std::max<size_t>(W, 1)));
    KZB_HIP(B, c->ws_grp.ensure(std::max<size_t>(W, 1)));
    KZB_HIP(B, B->h_out.ensure(4 * sizeof(float4) * size_t(B->nbodies) + 64));
    KZB_HIP(B, c->b_packed_solver.ensure(8 * size_t(std::max<uint32_t>(B->nbodies, 1))));
    B->tables_dirty = false;
    drop_graph(B);
    return KZG_OK;
}

// Per-pair and per-point buffers for up to `pairs` pairs (points <= 4 per pair).
int ensure_capacity(kzb_batch* B, uint32_t pairs) {
    if (pairs <= B->cap_pairs && B->cap_pairs) return KZG_OK;
    kzg_ctx* c = B->ctx;
    const uint32_t cap = std::max<uint32_t>(1024, pairs + pairs / 4);
    if (cap > kWorkPairMask) return bfail(B, KZG_ERR_UNSUPPORTED, "more than 2^27 - 1 broadphase pairs");
    const size_t pts = 4 * size_t(cap), W = B->worlds.size();
    KZB_HIP(B, c->d_pairs.ensure(cap));
    KZB_HIP(B, c->m_count.ensure(cap));
    KZB_HIP(B, c->m_pa.ensure(pts));
    KZB_HIP(B, c->m_pb.ensure(pts));
    KZB_HIP(B, c->m_n.ensure(pts));
    KZB_HIP(B, c->work.ensure(work_entries(cap)));
    KZB_HIP(B, c->gjk.ensure(size_t(cap) * kGjkRecBytes));
    KZB_HIP(B, c->c_info.ensure(cap));
    KZB_HIP(B, c->c_pair.ensure(cap));
    KZB_HIP(B, c->p_body.ensure(pts));
    for (auto* d : {&c->p_pa, &c->p_pb, &c->p_n, &c->p_t1, &c->p_t2, &c->p_ra, &c->p_rb, &c->p_la, &c->p_lb, &c->p_imp,
                    &c->vr_ra, &c->vr_rb, &c->vr_n, &c->vr_t1, &c->vr_t2, &c->vr_imp, &c->wr_ra, &c->wr_rb, &c->wr_qa,
                    &c->wr_qb, &c->wr_n, &c->wr_t1, &c->wr_t2, &c->wr_imp})
        KZB_HIP(B, d->ensure(pts));
    KZB_HIP(B, c->vr_bd.ensure(pts));
    KZB_HIP(B, c->wr_bd.ensure(pts));
    KZB_HIP(B, c->vr_fb.ensure(pts));
    KZB_HIP(B, c->p_nvel.ensure(pts));
    KZB_HIP(B, c->vel_pos.ensure(pts));
    KZB_HIP(B, c->ws_pos.ensure(pts));
    KZB_HIP(B, c->ws_off.ensure(pts + W + 1));
    KZB_HIP(B, c->vel_off.ensure(pts + W + 1));
    KZB_HIP(B, B->ws_level.ensure(pts));
    KZB_HIP(B, B->vel_level.ensure(pts));
    KZB_HIP(B, B->lane_pt.ensure(4 * pts));
    KZB_HIP(B, B->h_pairs.ensure(2 * sizeof(uint32_t) * size_t(cap) + sizeof(uint32_t) * (W + 1) + 512));
    B->cap_pairs = cap;
    drop_graph(B);
    return KZG_OK;
}

// The device step of every world, captured once per buffer layout.
// The device step of every world, in stream order.  `ev` (direct launches
// only) receives events between the phases: narrowphase, compaction and
// constraint build, schedule, integration and prepare, sweeps, write-back
// to the end.
void enqueue_step(kzb_batch* B, hipEvent_t* ev) {
    kzg_ctx* c = B->ctx;
    hipStream_t s = c->stream;
    auto mark = [&](int k) {
        if (ev) hipEventRecord(ev[k], s);
    };
    const kz_world_desc& ws = B->settings;
    const uint32_t W = uint32_t(B->worlds.size()), cap = B->cap_pairs, pts = 4 * cap, nb = B->nbodies;
    const int use_simd = ws.use_simd_solver ? 1 : 0;
    const float dt = ws.fixed_timestep;
    BatchDev d = B->dev();
    const uint32_t waves = B->cfg.narrowphase_waves ? B->cfg.narrowphase_waves : c->cu_count * 16u;
    const uint32_t threads = B->cfg.solve_threads ? B->cfg.solve_threads : 256u;
    SolveParams sp{dt, ws.warm_start_factor, ws.velocity_iterations, ws.position_iterations, ws.use_simd_solver,
                   ws.use_velocity_bias};
    mark(0);
    k_batch_begin<<<1, 256, 0, s>>>(d, c->work_count.p, c->resid.p, ws.velocity_iterations);
    launch_narrowphase_counted(s, c->d_pairs.p, cap, B->d_pair_off.p + W, c->bodies(), c->shapes(), c->manifolds(),
                               c->work.p, c->work_count.p, c->gjk.p, 0.02f, waves);
    mark(1);
    k_world_count<<<W, 256, 0, s>>>(d, c->m_count.p);
    k_world_scan<<<1, 1024, 0, s>>>(d);
    k_world_emit<<<W, 256, 0, s>>>(d, c->contacts(), c->bodies(), c->d_pairs.p, c->m_count.p);
    launch_build_counted(s, c->contacts(), c->manifolds(), c->bodies(), c->cache(), cap, B->cnt.p + kCntCons, 0);
    mark(2);
    k_world_schedule<<<W, 64, 0, s>>>(d, c->contacts(), c->bodies(), use_simd);
    k_level_scan<<<1, 1024, 0, s>>>(d, c->ws_grp.p, c->vel_grp.p);
    k_world_bucket<<<W, 256, 0, s>>>(d, c->ws_off.p, c->vel_off.p, c->ws_pos.p, c->vel_pos.p, use_simd);
    mark(3);
    launch_integrate_velocities(s, c->bodies(), nb, dt, ws.gravity);
    launch_prepare_counted(s, c->contacts(), c->bodies(), pts, B->cnt.p + kCntPts, dt, ws.restitution_threshold,
                           ws.baumgarte, ws.penetration_slop, ws.max_bias_velocity, use_simd,
                           ws.use_velocity_bias ? 1 : 0, c->vel_records(), c->ws_records(), c->vel_pos.p, c->ws_pos.p);
    mark(4);
    launch_solve_groups(s, c->vel_records(), c->ws_records(), c->bodies(), nb, c->b_packed_solver.p,
                        Schedule{c->ws_off.p, c->ws_grp.p, nullptr}, Schedule{c->vel_off.p, c->vel_grp.p, nullptr},
                        sp, W, threads, c->resid.p, UnitKernel{});
    mark(5);
    launch_writeback_counted(s, c->vel_records(), c->ws_records(), c->contacts(), c->vel_pos.p, c->ws_pos.p, pts,
                             B->cnt.p + kCntPts);
    launch_cache_store_counted(s, c->contacts(), c->bodies(), c->cache(), cap, B->cnt.p + kCntCons);
    launch_finish(s, c->bodies(), nb, dt, ws.use_velocity_bias ? 0 : 1);
    mark(6);
}

int capture(kzb_batch* B) {
    drop_graph(B);
    hipStream_t s = B->ctx->stream;
    KZB_HIP(B, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    enqueue_step(B, nullptr);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s, &g);
    if (e != hipSuccess) return bfail(B, KZG_ERR_DEVICE, std::string("graph capture: ") + hipGetErrorString(e));
    B->graph = g;
    KZB_HIP(B, hipGraphInstantiate(&B->exec, g, nullptr, nullptr, 0));
    B->graph_valid = true;
    return KZG_OK;
}

uint32_t hw_threads(const kzb_config& cfg) {
    if (cfg.broadphase_threads) return cfg.broadphase_threads;
    const uint32_t h = std::thread::hardware_concurrency();
    return std::max<uint32_t>(1, std::min<uint32_t>(h ? h : 1, 16));
}

template <class F> void parallel_for(uint32_t n, uint32_t threads, F&& f) {
    threads = std::max<uint32_t>(1, std::min(threads, n));
    if (threads == 1) {
        for (uint32_t i = 0; i < n; ++i) f(i);
        return;
    }
    std::atomic<uint32_t> next{0};
    std::vector<std::thread> pool;
    for (uint32_t t = 0; t < threads; ++t)
        pool.emplace_back([&] {
            for (uint32_t i; (i = next.fetch_add(16)) < n;)
                for (uint32_t k = i; k < std::min(n, i + 16); ++k) f(k);
        });
    for (auto& th : pool) th.join();
}

}  // namespace

extern "C" {

int kzb_create(kzb_batch** out, const kz_world_desc* settings, int device) {
    if (!out || !settings) return KZG_ERR_ARGUMENT;
    *out = nullptr;
    if (settings->allow_sleeping) return KZG_ERR_UNSUPPORTED;
    auto* B = new kzb_batch();
    B->settings = *settings;
    int rc = kzg_create(&B->ctx, settings, device);
    if (rc) {
        delete B;
        return rc;
    }
    hipEventCreate(&B->e0);
    hipEventCreate(&B->e1);
    hipEventCreate(&B->emid);
    for (auto& e : B->phase_ev) hipEventCreate(&e);
    *out = B;
    return KZG_OK;
}

void kzb_destroy(kzb_batch* B) {
    if (!B) return;
    if (B->ctx) {
        hipSetDevice(B->ctx->device);
        hipStreamSynchronize(B->ctx->stream);
    }
    drop_graph(B);
    for (auto* v : {&B->world_order, &B->d_body_off, &B->d_pair_off, &B->wc, &B->wp, &B->con_base, &B->pt_base, &B->cnt, &B->ws_level,
                    &B->vel_level, &B->depth_ws, &B->depth_vel, &B->vel_slots, &B->werr, &B->lvb_ws, &B->lvb_vel})
        v->release();
    B->lane_pt.release();
    B->h_pairs.release();
    B->h_out.release();
    if (B->e0) hipEventDestroy(B->e0);
    if (B->e1) hipEventDestroy(B->e1);
    if (B->emid) hipEventDestroy(B->emid);
    for (auto& e : B->phase_ev)
        if (e) hipEventDestroy(e);
    for (auto& W : B->worlds) kzw_destroy(W.w);
    kzg_destroy(B->ctx);
    delete B;
}

const char* kzb_last_error(const kzb_batch* B) { return B ? B->err.c_str() : "null batch"; }

int kzb_add_world(kzb_batch* B, const kz_shape_desc* shapes, uint32_t ns, const float* hull, uint32_t nh,
                  const kz_body_desc* bodies, uint32_t nb) {
    if (!B || (ns && !shapes) || (nb && !bodies)) return KZG_ERR_ARGUMENT;
    if (nb > kWorldBodies)
        return bfail(B, KZG_ERR_UNSUPPORTED, "a batched world holds at most " + std::to_string(kWorldBodies) + " bodies");
    kzb_batch::World W;
    int rc = kzw_create(&W.w, &B->settings, -1);
    if (rc) return bfail(B, rc, "kzw_create failed");
    if ((rc = kzw_add_shapes(W.w, shapes, ns, hull, nh)) || (rc = kzw_add_bodies(W.w, bodies, nb))) {
        std::string m = kzw_last_error(W.w);
        kzw_destroy(W.w);
        return bfail(B, rc, m);
    }
    W.shapes.assign(shapes, shapes + ns);
    W.bodies.assign(bodies, bodies + nb);
    B->worlds.push_back(std::move(W));
    B->tables_dirty = true;
    return int(B->worlds.size() - 1);
}

int kzb_num_worlds(kzb_batch* B) { return B ? int(B->worlds.size()) : KZG_ERR_ARGUMENT; }

int kzb_step(kzb_batch* B, float dt) {
    if (!B) return KZG_ERR_ARGUMENT;
    if (B->worlds.empty()) return bfail(B, KZG_ERR_STATE, "no worlds");
    if (dt != B->settings.fixed_timestep)
        return bfail(B, KZG_ERR_ARGUMENT, "the batched step runs at the fixed timestep of its settings");
    using clock = std::chrono::steady_clock;
    const auto t0 = clock::now();
    kzg_ctx* c = B->ctx;
    hipSetDevice(c->device);
    int rc;
    if (B->tables_dirty && (rc = upload_tables(B))) return rc;
    const uint32_t W = uint32_t(B->worlds.size());
    // 1. Broadphase of every world (UpdateBroadPhase + UpdatePairs), host threads.
    std::vector<int> brc(W, 0);
    parallel_for(W, hw_threads(B->cfg), [&](uint32_t i) {
        auto& w = B->worlds[i];
        brc[i] = kzw_update_broadphase(w.w);
        const int n = kzw_get_pairs(w.w, nullptr, 0);
        w.pairs.resize(2 * size_t(std::max(n, 0)));
        if (n > 0) kzw_get_pairs(w.w, w.pairs.data(), uint32_t(n));
    });
    for (uint32_t i = 0; i < W; ++i)
        if (brc[i]) return bfail(B, brc[i], "broadphase of world " + std::to_string(i) + ": " + kzw_last_error(B->worlds[i].w));
    B->pair_off.assign(W + 1, 0);
    for (uint32_t i = 0; i < W; ++i) B->pair_off[i + 1] = B->pair_off[i] + uint32_t(B->worlds[i].pairs.size() / 2);
    const uint32_t npairs = B->pair_off[W];
    const auto t1 = clock::now();
    // 2. Capacities (a change re-captures the graph).
    if ((rc = ensure_capacity(B, npairs))) return rc;
    const uint32_t cache_cap = c->cache_cap;
    if ((rc = ensure_cache(c, size_t(B->cache_used) + npairs))) return bfail(B, rc, kzg_last_error(c));
    if (c->cache_cap != cache_cap) drop_graph(B);
    if (!c->resid.p || c->resid.cap < std::max<uint32_t>(B->settings.velocity_iterations, 1)) {
        KZB_HIP(B, c->resid.ensure(std::max<uint32_t>(B->settings.velocity_iterations, 1)));
        drop_graph(B);
    }
    // 3. Pairs up (global slots), outside the graph.
    uint32_t* hp = B->h_pairs.as<uint32_t>();
    parallel_for(W, hw_threads(B->cfg), [&](uint32_t i) {
        const auto& pr = B->worlds[i].pairs;
        uint32_t* dst = hp + 2 * size_t(B->pair_off[i]);
        const uint32_t off = B->body_off[i];
        for (size_t k = 0; k < pr.size(); ++k) dst[k] = pr[k] + off;
    });
    uint32_t* hoff = hp + 2 * size_t(B->cap_pairs);
    std::memcpy(hoff, B->pair_off.data(), sizeof(uint32_t) * (W + 1));
    if (npairs)
        KZB_HIP(B, hipMemcpyAsync(c->d_pairs.p, hp, sizeof(uint32_t) * 2 * size_t(npairs), hipMemcpyHostToDevice,
                                  c->stream));
    KZB_HIP(B, hipMemcpyAsync(B->d_pair_off.p, hoff, sizeof(uint32_t) * (W + 1), hipMemcpyHostToDevice, c->stream));
    // 4. The captured device step.
    const bool direct = (B->cfg.flags & KZB_DIRECT_LAUNCH) != 0;
    if (!direct && !B->graph_valid && (rc = capture(B))) return rc;
    KZB_HIP(B, hipEventRecord(B->e0, c->stream));
    const auto tl0 = clock::now();
    if (direct) {
        enqueue_step(B, B->phase_ev);
        KZB_HIP(B, hipGetLastError());
    } else {
        KZB_HIP(B, hipGraphLaunch(B->exec, c->stream));
    }
    const auto tl1 = clock::now();
    KZB_HIP(B, hipEventRecord(B->emid, c->stream));
    {
        // Body states and counters back to the host (after the graph).
        const uint32_t n = B->nbodies;
        float4* o = B->h_out.as<float4>();
        hipStream_t s = c->stream;
        KZB_HIP(B, hipMemcpyAsync(o, c->b_x.p, sizeof(float4) * n, hipMemcpyDeviceToHost, s));
        KZB_HIP(B, hipMemcpyAsync(o + n, c->b_q.p, sizeof(float4) * n, hipMemcpyDeviceToHost, s));
        KZB_HIP(B, hipMemcpyAsync(o + 2 * size_t(n), c->b_v.p, sizeof(float4) * n, hipMemcpyDeviceToHost, s));
        KZB_HIP(B, hipMemcpyAsync(o + 3 * size_t(n), c->b_w.p, sizeof(float4) * n, hipMemcpyDeviceToHost, s));
        uint32_t* t = reinterpret_cast<uint32_t*>(o + 4 * size_t(n));
        KZB_HIP(B, hipMemcpyAsync(t, B->cnt.p, sizeof(uint32_t) * kCntN, hipMemcpyDeviceToHost, s));
        KZB_HIP(B, hipMemcpyAsync(t + kCntN, c->cache_used.p, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    }
    KZB_HIP(B, hipEventRecord(B->e1, c->stream));
    KZB_HIP(B, hipStreamSynchronize(c->stream));
    float dev_ms = 0.0f, all_ms = 0.0f;
    if (hipEventElapsedTime(&dev_ms, B->e0, B->emid) != hipSuccess) dev_ms = 0.0f;
    if (hipEventElapsedTime(&all_ms, B->e0, B->e1) != hipSuccess) all_ms = 0.0f;
    (void)hipGetLastError();
    const float4* out = B->h_out.as<float4>();
    const uint32_t nb = B->nbodies;
    const uint32_t* tail = reinterpret_cast<const uint32_t*>(out + 4 * size_t(nb));
    if (tail[kCntError])
        return bfail(B, KZG_ERR_UNSUPPORTED,
                     "a world exceeded the batched schedule's tables (error bits " + std::to_string(tail[kCntError]) +
                         ": 1 bodies/points, 2 first-fit batches > " + std::to_string(kWorldBatches) + ", 4 levels > " +
                         std::to_string(kWorldLevels) + ")");
    B->cache_used = tail[kCntN];
    c->ncons = tail[kCntCons];
    c->npoints = tail[kCntPts];
    c->resid_points = tail[kCntPts];
    c->npairs = npairs;
    // 5. States back into the host worlds (for their next broadphase).
    parallel_for(W, hw_threads(B->cfg), [&](uint32_t i) {
        const uint32_t b0 = B->body_off[i], n = B->body_off[i + 1] - b0;
        std::vector<kz_body_state> st(n);
        for (uint32_t k = 0; k < n; ++k) {
            const float4 x = out[b0 + k], q = out[nb + b0 + k], v = out[2 * size_t(nb) + b0 + k],
                         w = out[3 * size_t(nb) + b0 + k];
            st[k] = kz_body_state{{x.x, x.y, x.z}, {q.x, q.y, q.z, q.w}, {v.x, v.y, v.z}, {w.x, w.y, w.z}};
        }
        kzw_set_states(B->worlds[i].w, st.data(), n);
    });
    const auto t2 = clock::now();
    kzb_stats& S = B->stats;
    S.num_worlds = W;
    S.num_bodies = nb;
    S.num_pairs = npairs;
    S.num_constraints = tail[kCntCons];
    S.num_points = tail[kCntPts];
    S.max_velocity_levels = tail[kCntMaxVel];
    S.max_bias_levels = tail[kCntMaxWs];
    S.broadphase_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    S.device_ms = dev_ms;
    S.device_with_copies_ms = all_ms;
    S.step_ms = std::chrono::duration<double, std::milli>(t2 - t0).count();
    S.launch_ms = std::chrono::duration<double, std::milli>(tl1 - tl0).count();
    for (int k = 0; k < 6; ++k) {
        float ms = 0.0f;
        if (!direct || hipEventElapsedTime(&ms, B->phase_ev[k], B->phase_ev[k + 1]) != hipSuccess) ms = 0.0f;
        S.phase_ms[k] = ms;
    }
    (void)hipGetLastError();
    return KZG_OK;
}

int kzb_get_states(kzb_batch* B, uint32_t world, kz_body_state* out, uint32_t n) {
    if (!B || world >= B->worlds.size()) return KZG_ERR_ARGUMENT;
    return kzw_get_states(B->worlds[world].w, out, n);
}

int kzb_get_stats(kzb_batch* B, kzb_stats* out) {
    if (!B || !out) return KZG_ERR_ARGUMENT;
    *out = B->stats;
    return KZG_OK;
}

int kzb_set_config(kzb_batch* B, const kzb_config* cfg) {
    if (!B || !cfg) return KZG_ERR_ARGUMENT;
    const uint32_t t = cfg->solve_threads;
    if (t && (t < 64 || t > 1024 || (t & (t - 1))))
        return bfail(B, KZG_ERR_ARGUMENT, "solve_threads must be a power of two in [64, 1024]");
    if (cfg->flags & ~uint32_t(KZB_DIRECT_LAUNCH)) return bfail(B, KZG_ERR_ARGUMENT, "unknown config flags");
    B->cfg = *cfg;
    drop_graph(B);
    return KZG_OK;
}

int kzb_get_config(kzb_batch* B, kzb_config* cfg) {
    if (!B || !cfg) return KZG_ERR_ARGUMENT;
    *cfg = B->cfg;
    if (!cfg->narrowphase_waves) cfg->narrowphase_waves = B->ctx->cu_count * 16u;
    if (!cfg->solve_threads) cfg->solve_threads = 256;
    return KZG_OK;
}

int kzb_autotune_candidates(kzb_batch* B, kzb_config* out, uint32_t cap) {
    if (!B) return KZG_ERR_ARGUMENT;
    const uint32_t cu = B->ctx->cu_count;
    uint32_t n = 0;
    for (uint32_t waves : {cu * 8u, cu * 16u, cu * 32u})
        for (uint32_t threads : {64u, 128u, 256u, 512u}) {
            if (out && n < cap) out[n] = kzb_config{waves, threads, B->cfg.broadphase_threads, B->cfg.flags};
            ++n;
        }
    return int(n);
}

int kzb_autotune(kzb_batch* B, float dt, uint32_t repeats, kzb_config* best, float* times, uint32_t cap) {
    if (!B || !repeats) return KZG_ERR_ARGUMENT;
    const int n = kzb_autotune_candidates(B, nullptr, 0);
    std::vector<kzb_config> cand(n);
    kzb_autotune_candidates(B, cand.data(), uint32_t(n));
    std::vector<std::vector<float>> ms(n);
    for (uint32_t r = 0; r < repeats; ++r)
        for (int k = 0; k < n; ++k) {
            int rc = kzb_set_config(B, &cand[k]);
            if (!rc) rc = kzb_step(B, dt);
            if (rc) return rc;
            ms[k].push_back(float(B->stats.device_ms));
        }
    int bi = 0;
    std::vector<float> med(n);
    for (int k = 0; k < n; ++k) {
        std::sort(ms[k].begin(), ms[k].end());
        med[k] = ms[k][ms[k].size() / 2];
        if (med[k] < med[bi]) bi = k;
        if (times && uint32_t(k) < cap) times[k] = med[k];
    }
    int rc = kzb_set_config(B, &cand[bi]);
    if (rc) return rc;
    if (best) *best = cand[bi];
    return bi;
}

int kzb_world_schedule(kzb_batch* B, uint32_t world, uint32_t* vel_level, uint32_t* ws_level, uint32_t cap) {
    if (!B || world >= B->worlds.size()) return KZG_ERR_ARGUMENT;
    kzg_ctx* c = B->ctx;
    hipSetDevice(c->device);
    uint32_t base = 0, np = 0;
    KZB_HIP(B, hipMemcpy(&base, B->pt_base.p + world, sizeof(uint32_t), hipMemcpyDeviceToHost));
    KZB_HIP(B, hipMemcpy(&np, B->wp.p + world, sizeof(uint32_t), hipMemcpyDeviceToHost));
    const uint32_t k = std::min(np, cap);
    if (k && vel_level)
        KZB_HIP(B, hipMemcpy(vel_level, B->vel_level.p + base, sizeof(uint32_t) * k, hipMemcpyDeviceToHost));
    if (k && ws_level)
        KZB_HIP(B, hipMemcpy(ws_level, B->ws_level.p + base, sizeof(uint32_t) * k, hipMemcpyDeviceToHost));
    return int(np);
}

}  // extern "C"
