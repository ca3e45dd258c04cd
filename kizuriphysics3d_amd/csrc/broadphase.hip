// broadphase.hip - the broadphase pair query on the device, in the
// reference's pair order (DynamicAABBTree::QueryPairs + BroadPhase::
// UpdatePairs, include/Kizuri/BroadPhase/BroadPhase.h:61-104,
// src/BroadPhase/BroadPhase.cpp:368-381).
//
// The reference walks node pairs depth-first from (root, root) with a LIFO
// stack.  Popping (n1, n2):
//   n1 == n2: a leaf yields nothing; an internal node pushes (c1, c1),
//             (c2, c2), (c1, c2);
//   fat boxes apart: nothing;
//   two leaves: the pair if n1 < n2, the layer / mask test passes both ways
//             and the tight boxes overlap (a pair met as n1 > n2 is lost);
//   else the larger node is split: with a a leaf, or b internal and taller,
//             (n1, b.c1), (n1, b.c2) are pushed, else (a.c1, n2), (a.c2, n2).
// The output is therefore the pre-order of the walk's tree, children in
// reverse push order.  On the device that order is kept in two stages:
//   1. Refinement rounds.  A list of items (a pair found, or a node pair
//      still to walk) starts as {(root, root)}; each round replaces every
//      node pair whose heights sum above kWalkHeight by what popping it
//      yields, in pop order, keeping every other item in place (count,
//      scan, write).
//   2. Walks.  Every remaining node pair is walked depth-first by one lane
//      with its stack in LDS, the reference's pop for pop: a counting pass,
//      a scan of the counts, and a pass that writes the pairs at their
//      offsets.  A lane takes its next item from a shared counter as soon
//      as its stack is empty, so walks of very different lengths do not hold
//      each other up.
// The tree itself (inserts, removals, rotations) is kept on the host, as the
// reference keeps it; it comes up as an image of one 64-byte record per
// node.
#include "ctx_internal.h"

namespace kz {
namespace {

constexpr uint32_t kNullNode = 0xFFFFFFFFu;
constexpr uint32_t kPairFlag = 0x80000000u;  // item.x: a pair found (node ids < 2^31)
// Node pairs whose heights sum to at most this are walked by one thread: a
// walk from (n1, n2) descends at most h1 + h2 times and leaves at most two
// siblings per descent on the stack, so kWalkStack entries always suffice.
// 4: on a settled 32k-body pile the longest walk is 31 pops (335 at 8,
// 1,873 at 12), and a walk is a chain of dependent node loads, so the
// longest one sets the walk kernels' time.
constexpr uint32_t kWalkHeight = 4;
constexpr uint32_t kWalkStack = 2 * kWalkHeight + 4;
constexpr uint32_t kWalkThreads = 128;
constexpr uint32_t kScanThreads = 256;
// Error flags (BpCounters::flags).
constexpr uint32_t kBpOverflow = 1u, kBpBadTree = 2u;

// Node image: fat min + child1, fat max + child2, tight min + user,
// tight max + height (the ints as float bits).
struct BpTree {
    const float4* node;
    const uint32_t* layer;
    const uint32_t* mask;
    uint32_t num_nodes, num_layers, num_bodies;
};
struct BpNode {
    float4 mn, mx;
    uint32_t c1, c2;
    int32_t h;
};
__device__ __forceinline__ BpNode load_node(const BpTree& t, uint32_t i) {
    const float4 a = t.node[4 * size_t(i)], b = t.node[4 * size_t(i) + 1], d = t.node[4 * size_t(i) + 3];
    return BpNode{a, b, __float_as_uint(a.w), __float_as_uint(b.w), __float_as_int(d.w)};
}
// AABB::Overlaps, bounds inclusive.
__device__ __forceinline__ bool overlap(float4 amn, float4 amx, float4 bmn, float4 bmx) {
    return amn.x <= bmx.x && amx.x >= bmn.x && amn.y <= bmx.y && amx.y >= bmn.y && amn.z <= bmx.z && amx.z >= bmn.z;
}

// The counters of one query: items of the current list, pairs, flags.
struct BpCounters {
    uint32_t items, next_items, pairs, flags;
};

// One pop of node pair e: what it yields, in the order the reference's
// stack pops it; pairs carry kPairFlag.  Returns the count (0-3).
__device__ __forceinline__ uint32_t pop_pair(const BpTree& t, uint2 e, uint2 out[3], uint32_t* flags) {
    const uint32_t n1 = e.x, n2 = e.y;
    if (n1 >= t.num_nodes || n2 >= t.num_nodes) {
        atomicOr(flags, kBpBadTree);
        return 0;
    }
    const BpNode a = load_node(t, n1);
    if (n1 == n2) {
        if (a.c1 == kNullNode) return 0;
        out[0] = make_uint2(a.c1, a.c2);
        out[1] = make_uint2(a.c2, a.c2);
        out[2] = make_uint2(a.c1, a.c1);
        return 3;
    }
    const BpNode b = load_node(t, n2);
    if (!overlap(a.mn, a.mx, b.mn, b.mx)) return 0;
    const bool la = a.c1 == kNullNode, lb = b.c1 == kNullNode;
    if (la && lb) {
        if (n1 >= n2) return 0;
        if (n1 >= t.num_layers || n2 >= t.num_layers) {
            atomicOr(flags, kBpBadTree);
            return 0;
        }
        if ((t.layer[n1] & t.mask[n2]) == 0u || (t.layer[n2] & t.mask[n1]) == 0u) return 0;
        const float4 ta0 = t.node[4 * size_t(n1) + 2], ta1 = t.node[4 * size_t(n1) + 3];
        const float4 tb0 = t.node[4 * size_t(n2) + 2], tb1 = t.node[4 * size_t(n2) + 3];
        if (!overlap(ta0, ta1, tb0, tb1)) return 0;
        out[0] = make_uint2(n1 | kPairFlag, n2);
        return 1;
    }
    if (la || (!lb && a.h < b.h)) {
        out[0] = make_uint2(n1, b.c2);
        out[1] = make_uint2(n1, b.c1);
    } else {
        out[0] = make_uint2(a.c2, n2);
        out[1] = make_uint2(a.c1, n2);
    }
    return 2;
}

__device__ __forceinline__ bool is_pair(uint2 it) { return (it.x & kPairFlag) != 0u; }
__device__ __forceinline__ uint32_t height_sum(const BpTree& t, uint2 e) {
    if (e.x >= t.num_nodes || e.y >= t.num_nodes) return 0;  // a bad id is flagged where it is popped
    return uint32_t(max(__float_as_int(t.node[4 * size_t(e.x) + 3].w), 0)) +
           uint32_t(max(__float_as_int(t.node[4 * size_t(e.y) + 3].w), 0));
}

// Exclusive scan of one value per thread over a workgroup of blockDim.x
// (a power of two up to 1024); returns the prefix and stores the total.
__device__ __forceinline__ uint32_t block_scan(uint32_t v, uint32_t* lds, uint32_t* total) {
    const uint32_t t = threadIdx.x, n = blockDim.x;
    lds[t] = v;
    __syncthreads();
    for (uint32_t off = 1; off < n; off <<= 1) {
        const uint32_t add = t >= off ? lds[t - off] : 0u;
        __syncthreads();
        lds[t] += add;
        __syncthreads();
    }
    const uint32_t incl = lds[t];
    if (t == n - 1) *total = incl;
    __syncthreads();
    return incl - v;
}

// Block totals -> exclusive offsets, one workgroup; *grand = the sum.
__global__ void __launch_bounds__(1024) k_bp_scan_blocks(uint32_t* sums, uint32_t nblocks_cap, const uint32_t* nitems,
                                                         uint32_t per_block, uint32_t* grand) {
    __shared__ uint32_t lds[1024];
    __shared__ uint32_t tot;
    const uint32_t nb = (*nitems + per_block - 1) / per_block;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nb && base < nblocks_cap; base += blockDim.x) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t v = i < nb ? sums[i] : 0u;
        const uint32_t ex = block_scan(v, lds, &tot);
        if (i < nb) sums[i] = carry + ex;
        carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) *grand = carry;
}

// ---- refinement rounds ------------------------------------------------------
__device__ __forceinline__ uint32_t refine_count(const BpTree& t, uint2 it, uint2 out[3], uint32_t* flags) {
    if (is_pair(it) || height_sum(t, it) <= kWalkHeight) {
        out[0] = it;
        return 1;
    }
    return pop_pair(t, it, out, flags);
}
__global__ void __launch_bounds__(kScanThreads) k_bp_refine_count(BpTree t, const uint2* items, BpCounters* cnt,
                                                                  uint32_t* sums) {
    __shared__ uint32_t lds[kScanThreads];
    __shared__ uint32_t tot;
    const uint32_t n = cnt->items, i = blockIdx.x * blockDim.x + threadIdx.x;
    if (blockIdx.x * blockDim.x >= n) return;
    uint2 out[3];
    const uint32_t k = i < n ? refine_count(t, items[i], out, &cnt->flags) : 0u;
    block_scan(k, lds, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}
__global__ void __launch_bounds__(kScanThreads) k_bp_refine_emit(BpTree t, const uint2* items, BpCounters* cnt,
                                                                 const uint32_t* sums, uint2* next, uint32_t cap) {
    __shared__ uint32_t lds[kScanThreads];
    __shared__ uint32_t tot;
    const uint32_t n = cnt->items, i = blockIdx.x * blockDim.x + threadIdx.x;
    if (blockIdx.x * blockDim.x >= n) return;
    uint2 out[3];
    const uint32_t k = i < n ? refine_count(t, items[i], out, &cnt->flags) : 0u;
    const uint32_t at = sums[blockIdx.x] + block_scan(k, lds, &tot);
    if (at + k > cap) {
        if (k) atomicOr(&cnt->flags, kBpOverflow);
        return;
    }
    if (k >= 1) next[at] = out[0];
    if (k >= 2) next[at + 1] = out[1];
    if (k == 3) next[at + 2] = out[2];
}
// The new list becomes the current one (the grand total is its length).
__global__ void k_bp_swap(BpCounters* cnt, uint32_t cap) {
    if (threadIdx.x == 0) cnt->items = min(cnt->next_items, cap);
}

// ---- walks ------------------------------------------------------------------
// Walks of the list's items, lane by lane: each wave owns a contiguous run
// of the list; each lane keeps one walk in flight (its stack in LDS) and
// takes the wave's next item as soon as its stack runs empty, so one loop
// serves items of any length.  The run's next index is the same in every
// active lane: the lanes that need an item count themselves with a ballot
// and take consecutive items by rank (no atomics, which on one address
// serialise across the device).  on_pair(i, p1, p2) for every pair of item
// i in order; on_done(i, count) when item i ends.  A pair item yields
// itself.
template <class OnPair, class OnDone>
__device__ __forceinline__ void walk_items(const BpTree& t, const uint2* items, uint32_t n, uint2* stack,
                                           uint32_t* flags, OnPair&& on_pair, OnDone&& on_done) {
    const uint32_t waves = gridDim.x * (blockDim.x / 64u);
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / 64u;
    const uint32_t per = (n + waves - 1) / waves;
    uint32_t next = wave * per;  // the same in every lane of the wave
    const uint32_t end = min(next + per, n);
    const uint32_t lane = __lane_id();
    uint32_t sp = 0, item = 0, found = 0;
    for (;;) {
        const bool need = sp == 0;
        const uint64_t mask = __ballot(need);
        if (need) {
            item = next + uint32_t(__popcll(mask & ((1ull << lane) - 1ull)));
            if (item >= end) return;
        }
        next += uint32_t(__popcll(mask));
        if (need) {
            const uint2 it = items[item];
            if (is_pair(it)) {
                on_pair(item, it.x & ~kPairFlag, it.y);
                on_done(item, 1u);
                continue;
            }
            stack[sp++] = it;
            found = 0;
        }
        const uint2 e = stack[--sp];
        uint2 out[3];
        const uint32_t k = pop_pair(t, e, out, flags);
        if (k == 1 && is_pair(out[0])) {
            on_pair(item, out[0].x & ~kPairFlag, out[0].y);
            ++found;
        } else if (sp + k > kWalkStack) {  // heights that break the tree's invariant
            atomicOr(flags, kBpBadTree);
            sp = 0;
        } else {
            // Pushed in reverse pop order, so the stack pops them as listed
            // (constant indices: `out` stays in registers).
            if (k == 3) stack[sp++] = out[2];
            if (k >= 2) stack[sp++] = out[1];
            if (k >= 1) stack[sp++] = out[0];
        }
        if (sp == 0) on_done(item, found);
    }
}
__global__ void __launch_bounds__(kWalkThreads) k_bp_walk_count(BpTree t, const uint2* items, BpCounters* cnt,
                                                                uint32_t* counts) {
    __shared__ uint2 stacks[kWalkThreads * kWalkStack];
    walk_items(t, items, cnt->items, stacks + threadIdx.x * kWalkStack, &cnt->flags, [](uint32_t, uint32_t, uint32_t) {},
               [&](uint32_t i, uint32_t k) { counts[i] = k; });
}
__global__ void __launch_bounds__(kWalkThreads) k_bp_walk_emit(BpTree t, const uint2* items, BpCounters* cnt,
                                                               const uint32_t* offsets, uint2* pairs, uint32_t cap) {
    __shared__ uint2 stacks[kWalkThreads * kWalkStack];
    uint32_t at = 0, cur = 0xFFFFFFFFu;
    walk_items(
        t, items, cnt->items, stacks + threadIdx.x * kWalkStack, &cnt->flags,
        [&](uint32_t i, uint32_t p1, uint32_t p2) {
            if (i != cur) {
                cur = i;
                at = offsets[i];
            }
            if (at >= cap) {
                atomicOr(&cnt->flags, kBpOverflow);
                return;
            }
            // Pair of proxies -> pair of bodies (the proxies' user data).
            const uint32_t u1 = __float_as_uint(t.node[4 * size_t(p1) + 2].w);
            const uint32_t u2 = __float_as_uint(t.node[4 * size_t(p2) + 2].w);
            if (u1 >= t.num_bodies || u2 >= t.num_bodies) atomicOr(&cnt->flags, kBpBadTree);
            pairs[at++] = make_uint2(u1 < t.num_bodies ? u1 : 0u, u2 < t.num_bodies ? u2 : 0u);
        },
        [](uint32_t, uint32_t) {});
}
// Exclusive scan of counts[0 .. cnt->items) into offsets, in three launches
// (block sums, their scan, per-item offsets); the grand total goes to
// cnt->pairs.
constexpr uint32_t kScanBlock = 1024;
__global__ void __launch_bounds__(kScanBlock) k_bp_block_sums(const uint32_t* counts, const BpCounters* cnt,
                                                              uint32_t* sums) {
    __shared__ uint32_t lds[kScanBlock];
    __shared__ uint32_t tot;
    const uint32_t n = cnt->items, i = blockIdx.x * blockDim.x + threadIdx.x;
    if (blockIdx.x * blockDim.x >= n) return;
    block_scan(i < n ? counts[i] : 0u, lds, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}
__global__ void __launch_bounds__(kScanBlock) k_bp_offsets(const uint32_t* counts, const BpCounters* cnt,
                                                           const uint32_t* sums, uint32_t* offsets) {
    __shared__ uint32_t lds[kScanBlock];
    __shared__ uint32_t tot;
    const uint32_t n = cnt->items, i = blockIdx.x * blockDim.x + threadIdx.x;
    if (blockIdx.x * blockDim.x >= n) return;
    const uint32_t ex = block_scan(i < n ? counts[i] : 0u, lds, &tot);
    if (i < n) offsets[i] = sums[blockIdx.x] + ex;
}
__global__ void k_bp_start(uint2* items, BpCounters* cnt, uint32_t root) {
    if (threadIdx.x == 0) {
        items[0] = make_uint2(root, root);
        cnt->items = root == kNullNode ? 0u : 1u;
        cnt->next_items = 0;
        cnt->pairs = 0;
        cnt->flags = 0;
    }
}
__global__ void k_bp_finish(BpCounters* cnt, uint32_t cap) {
    if (threadIdx.x == 0) cnt->pairs = min(cnt->pairs, cap);
}

// ---- proxy moves ------------------------------------------------------------
__device__ __forceinline__ f3 xyz(const float4& v) { return f3{v.x, v.y, v.z}; }
// MoveProxy's first half for the active list's non-static bodies: the world
// bounds as Shape::GetWorldBounds forms them (the local box's centre rotated
// and moved, its half extents through |R|; the same expressions as the host's
// transform_aabb) become the leaf's tight box in the image; a leaf whose fat
// box no longer contains it is reported (moved[0] = count, moved[1] = flags,
// records of 7 words from moved[2]).
__global__ void k_bp_move(BodyBuf b, const float4* local, const uint32_t* proxy, const uint32_t* act, uint32_t n,
                          float4* node, uint32_t num_nodes, uint32_t* moved, uint32_t cap) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t x = act[i];
    if (meta_motion(b.meta[x]) == 0u) return;
    const uint32_t id = proxy[x];
    if (id >= num_nodes) {
        atomicOr(moved + 1, kBpBadTree);
        return;
    }
    const f3 lmn = xyz(local[2 * size_t(x)]), lmx = xyz(local[2 * size_t(x) + 1]);
    const f3 c = mul(add(lmn, lmx), 0.5f);
    const f3 e = mul(sub(lmx, lmn), 0.5f);
    const float4 q4 = b.q[x];
    const f4q q{q4.x, q4.y, q4.z, q4.w};
    const m3 m = to_mat3(q);
    const f3 nc = add(rotate(q, c), xyz(b.x[x]));
    const f3 ne{fabs_(m.c0.x) * e.x + fabs_(m.c1.x) * e.y + fabs_(m.c2.x) * e.z,
                fabs_(m.c0.y) * e.x + fabs_(m.c1.y) * e.y + fabs_(m.c2.y) * e.z,
                fabs_(m.c0.z) * e.x + fabs_(m.c1.z) * e.y + fabs_(m.c2.z) * e.z};
    const f3 tmn = sub(nc, ne), tmx = add(nc, ne);
    float4* rec = node + 4 * size_t(id);
    const float4 f0 = rec[0], f1 = rec[1];
    float4 r2 = rec[2], r3 = rec[3];
    r2.x = tmn.x; r2.y = tmn.y; r2.z = tmn.z;
    r3.x = tmx.x; r3.y = tmx.y; r3.z = tmx.z;
    rec[2] = r2;
    rec[3] = r3;
    const bool inside = f0.x <= tmn.x && f1.x >= tmx.x && f0.y <= tmn.y && f1.y >= tmx.y && f0.z <= tmn.z &&
                        f1.z >= tmx.z;
    if (inside) return;
    const uint32_t k = atomicAdd(moved, 1u);
    if (k >= cap) return;
    uint32_t* o = moved + 2 + 7 * size_t(k);
    o[0] = i;
    o[1] = __float_as_uint(tmn.x); o[2] = __float_as_uint(tmn.y); o[3] = __float_as_uint(tmn.z);
    o[4] = __float_as_uint(tmx.x); o[5] = __float_as_uint(tmx.y); o[6] = __float_as_uint(tmx.z);
}
__global__ void k_bp_move_reset(uint32_t* moved) {
    if (threadIdx.x < 2) moved[threadIdx.x] = 0u;
}
// Replaced node records into the image.
__global__ void k_bp_scatter(const uint32_t* ids, const float4* rec, uint32_t n, float4* node) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4* d = node + 4 * size_t(ids[i]);
    for (int k = 0; k < 4; ++k) d[k] = rec[4 * size_t(i) + k];
}

}  // namespace

// Query of one tree image: rounds = how many refinement rounds the root's
// height needs.  items_a / items_b: list buffers of items_cap entries each;
// counts, offsets: items_cap words each; sums: ceil(items_cap /
// kScanThreads) words; pairs: pairs_cap entries.  cnt
// (device) ends with items, pairs and the error flags.
void launch_bp_query(hipStream_t s, const BpTree& t, uint32_t root, uint32_t rounds, uint2* items_a, uint2* items_b,
                     uint32_t items_cap, uint32_t* counts, uint32_t* offsets, uint32_t* sums, uint32_t sums_cap,
                     uint2* pairs, uint32_t pairs_cap, BpCounters* cnt, uint32_t walk_blocks) {
    k_bp_start<<<1, 64, 0, s>>>(items_a, cnt, root);
    uint2 *cur = items_a, *nxt = items_b;
    const uint32_t rb = (items_cap + kScanThreads - 1) / kScanThreads;
    for (uint32_t r = 0; r < rounds; ++r) {
        k_bp_refine_count<<<rb, kScanThreads, 0, s>>>(t, cur, cnt, sums);
        k_bp_scan_blocks<<<1, 1024, 0, s>>>(sums, sums_cap, &cnt->items, kScanThreads, &cnt->next_items);
        k_bp_refine_emit<<<rb, kScanThreads, 0, s>>>(t, cur, cnt, sums, nxt, items_cap);
        k_bp_swap<<<1, 64, 0, s>>>(cnt, items_cap);
        std::swap(cur, nxt);
    }
    k_bp_walk_count<<<walk_blocks, kWalkThreads, 0, s>>>(t, cur, cnt, counts);
    const uint32_t sb = (items_cap + kScanBlock - 1) / kScanBlock;
    k_bp_block_sums<<<sb, kScanBlock, 0, s>>>(counts, cnt, sums);
    k_bp_scan_blocks<<<1, 1024, 0, s>>>(sums, sums_cap, &cnt->items, kScanBlock, &cnt->pairs);
    k_bp_offsets<<<sb, kScanBlock, 0, s>>>(counts, cnt, sums, offsets);
    k_bp_walk_emit<<<walk_blocks, kWalkThreads, 0, s>>>(t, cur, cnt, offsets, pairs, pairs_cap);
    k_bp_finish<<<1, 64, 0, s>>>(cnt, pairs_cap);
}

}  // namespace kz

namespace {
// The query over the image on the device (num_nodes records, num_layers
// layers / masks), with the rounds the root's height needs.
int query_image(kzg_ctx* c, uint32_t num_nodes, uint32_t num_layers, uint32_t root, int32_t root_height,
                uint32_t* num_pairs) {
    const int32_t h = root == kNullNode ? 0 : std::max(root_height, 0);
    const uint32_t rounds = uint32_t(std::max<int64_t>(0, 2 * int64_t(h) - int64_t(kWalkHeight)));
    KZG_HIP(c, c->bp_counters.ensure(4));
    KZG_HIP(c, c->pin_bp.ensure(4 * sizeof(uint32_t)));
    const BpTree t{c->bp_nodes.p, c->bp_layer.p, c->bp_mask.p, num_nodes, num_layers, c->nbodies};
    if (!c->bp_items_cap) c->bp_items_cap = std::max<uint32_t>(1u << 16, 16u * num_nodes);
    if (!c->bp_pairs_cap) c->bp_pairs_cap = std::max<uint32_t>(1u << 16, 8u * num_nodes);
    for (int attempt = 0;; ++attempt) {
        const uint32_t icap = c->bp_items_cap, pcap = c->bp_pairs_cap;
        const uint32_t scap = (icap + kScanThreads - 1) / kScanThreads;
        KZG_HIP(c, c->bp_items_a.ensure(icap));
        KZG_HIP(c, c->bp_items_b.ensure(icap));
        KZG_HIP(c, c->bp_counts.ensure(2 * size_t(icap)));  // counts, then offsets
        KZG_HIP(c, c->bp_sums.ensure(scap));
        KZG_HIP(c, c->d_pairs.ensure(pcap));
        BpCounters* cnt = reinterpret_cast<BpCounters*>(c->bp_counters.p);
        launch_bp_query(c->stream, t, root, rounds, c->bp_items_a.p, c->bp_items_b.p, icap, c->bp_counts.p,
                        c->bp_counts.p + icap, c->bp_sums.p, scap, c->d_pairs.p, pcap, cnt, c->cu_count * 7u);
        KZG_HIP(c, hipGetLastError());
        uint32_t* pin = c->pin_bp.as<uint32_t>();
        KZG_HIP(c, hipMemcpyAsync(pin, cnt, sizeof(BpCounters), hipMemcpyDeviceToHost, c->stream));
        KZG_HIP(c, hipStreamSynchronize(c->stream));
        const BpCounters got{pin[0], pin[1], pin[2], pin[3]};
        if (got.flags & kBpBadTree) {
            c->bp_num_nodes = 0;  // the image is not to be reused
            return fail(c, KZG_ERR_ARGUMENT, "kzg_broadphase_pairs: malformed tree image (a node id, a height or "
                                             "a proxy's body outside its range)");
        }
        if (got.flags & kBpOverflow) {
            if (attempt >= 8) return fail(c, KZG_ERR_DEVICE, "kzg_broadphase_pairs: lists keep overflowing");
            // The pair count is exact once the walk counted; the item lists
            // may have been cut short, so both grow.
            c->bp_items_cap = icap * 2;
            c->bp_pairs_cap = std::max(pcap * 2, got.pairs + got.pairs / 4);
            continue;
        }
        c->npairs = got.pairs;
        c->pairs_resident = true;
        c->bp_root = root;
        c->bp_root_height = h;
        if (num_pairs) *num_pairs = got.pairs;
        return KZG_OK;
    }
}
}  // namespace

extern "C" {

int kzg_broadphase_pairs(kzg_ctx* c, const kzg_bp_node* nodes, uint32_t num_nodes, uint32_t root,
                         const uint32_t* layers, const uint32_t* masks, uint32_t num_layers, uint32_t* num_pairs) {
    if (!c || (num_nodes && !nodes) || (num_layers && (!layers || !masks)))
        return fail(c, KZG_ERR_ARGUMENT, "kzg_broadphase_pairs: null array");
    if (num_nodes >= kPairFlag || num_layers > num_nodes)
        return fail(c, KZG_ERR_ARGUMENT, "kzg_broadphase_pairs: bad node / layer count");
    if (root != kNullNode && root >= num_nodes) return fail(c, KZG_ERR_ARGUMENT, "kzg_broadphase_pairs: bad root");
    join_schedule(c);
    hipSetDevice(c->device);
    // The image goes up through one pinned buffer, with no wait until the
    // query's counters come back.
    int rc;
    KZG_HIP(c, c->stage_bp.ensure(16 * size_t(num_nodes) * sizeof(float) + 2 * sizeof(uint32_t) * num_layers + 768));
    size_t off = 0;
    if ((rc = upload_async(c, c->bp_nodes, c->stage_bp, off, reinterpret_cast<const float4*>(nodes),
                           4 * size_t(num_nodes))) ||
        (rc = upload_async(c, c->bp_layer, c->stage_bp, off, layers, num_layers)) ||
        (rc = upload_async(c, c->bp_mask, c->stage_bp, off, masks, num_layers)))
        return rc;
    c->bp_num_nodes = num_nodes;
    c->bp_num_layers = num_layers;
    return query_image(c, num_nodes, num_layers, root, root == kNullNode ? 0 : nodes[root].height, num_pairs);
}

int kzg_broadphase_pairs_update(kzg_ctx* c, const uint32_t* ids, const kzg_bp_node* records, uint32_t n,
                                uint32_t num_nodes, uint32_t root, uint32_t* num_pairs) {
    if (!c || (n && (!ids || !records))) return fail(c, KZG_ERR_ARGUMENT, "kzg_broadphase_pairs_update: null array");
    if (!c->bp_num_nodes || num_nodes != c->bp_num_nodes)
        return fail(c, KZG_ERR_STATE, "kzg_broadphase_pairs_update: no image of that size on the device "
                                      "(kzg_broadphase_pairs first)");
    if (root != kNullNode && root >= num_nodes)
        return fail(c, KZG_ERR_ARGUMENT, "kzg_broadphase_pairs_update: bad root");
    int32_t h = root == c->bp_root ? c->bp_root_height : -1;
    for (uint32_t k = 0; k < n; ++k) {
        if (ids[k] >= num_nodes) return fail(c, KZG_ERR_ARGUMENT, "kzg_broadphase_pairs_update: bad node id");
        if (ids[k] == root) h = records[k].height;
    }
    if (root != kNullNode && h < 0)
        return fail(c, KZG_ERR_ARGUMENT, "kzg_broadphase_pairs_update: a new root must come with its record");
    join_schedule(c);
    hipSetDevice(c->device);
    if (n) {
        int rc;
        KZG_HIP(c, c->stage_bp.ensure(n * (sizeof(uint32_t) + sizeof(kzg_bp_node)) + 512));
        size_t off = 0;
        if ((rc = upload_async(c, c->bp_upd_ids, c->stage_bp, off, ids, n)) ||
            (rc = upload_async(c, c->bp_upd_rec, c->stage_bp, off, reinterpret_cast<const float4*>(records),
                               4 * size_t(n))))
            return rc;
        k_bp_scatter<<<(n + 255) / 256, 256, 0, c->stream>>>(c->bp_upd_ids.p, c->bp_upd_rec.p, n, c->bp_nodes.p);
        KZG_HIP(c, hipGetLastError());
    }
    return query_image(c, num_nodes, c->bp_num_layers, root, h, num_pairs);
}

int kzg_broadphase_set_proxies(kzg_ctx* c, const float* local, const uint32_t* proxies, uint32_t num_bodies) {
    if (!c || (num_bodies && (!local || !proxies))) return fail(c, KZG_ERR_ARGUMENT, "null array");
    if (num_bodies != c->nbodies) return fail(c, KZG_ERR_ARGUMENT, "one record per body of the body table");
    std::vector<float4> lb(2 * size_t(num_bodies));
    for (uint32_t i = 0; i < num_bodies; ++i) {
        lb[2 * size_t(i)] = make_float4(local[6 * size_t(i)], local[6 * size_t(i) + 1], local[6 * size_t(i) + 2], 0.0f);
        lb[2 * size_t(i) + 1] =
            make_float4(local[6 * size_t(i) + 3], local[6 * size_t(i) + 4], local[6 * size_t(i) + 5], 0.0f);
    }
    join_schedule(c);
    hipSetDevice(c->device);
    int rc;
    if ((rc = upload(c, c->bp_local, lb.data(), lb.size())) || (rc = upload(c, c->bp_proxy, proxies, num_bodies)))
        return rc;
    c->bp_proxies = num_bodies;
    return KZG_OK;
}

int kzg_broadphase_move(kzg_ctx* c, float* out, uint32_t cap, uint32_t* count) {
    if (!c || (cap && !out) || !count) return KZG_ERR_ARGUMENT;
    if (!c->bp_num_nodes) return fail(c, KZG_ERR_STATE, "kzg_broadphase_move: no tree image on the device");
    if (c->bp_proxies != c->nbodies || !c->nbodies)
        return fail(c, KZG_ERR_STATE, "kzg_broadphase_move: proxies not registered (kzg_broadphase_set_proxies)");
    join_schedule(c);
    hipSetDevice(c->device);
    const uint32_t n = c->h_act_n;
    KZG_HIP(c, c->bp_moved.ensure(2 + 7 * size_t(std::max<uint32_t>(n, 1))));
    KZG_HIP(c, c->pin_bp.ensure(4 * sizeof(uint32_t)));
    k_bp_move_reset<<<1, 64, 0, c->stream>>>(c->bp_moved.p);
    if (n)
        k_bp_move<<<(n + 255) / 256, 256, 0, c->stream>>>(c->bodies(), c->bp_local.p, c->bp_proxy.p, c->d_act.p, n,
                                                          c->bp_nodes.p, c->bp_num_nodes, c->bp_moved.p, n);
    KZG_HIP(c, hipGetLastError());
    uint32_t* pin = c->pin_bp.as<uint32_t>();
    KZG_HIP(c, hipMemcpyAsync(pin, c->bp_moved.p, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    KZG_HIP(c, hipStreamSynchronize(c->stream));
    if (pin[1] & kBpBadTree) return fail(c, KZG_ERR_ARGUMENT, "kzg_broadphase_move: a proxy id outside the image");
    const uint32_t m = pin[0];
    *count = m;
    const uint32_t w = std::min(m, cap);
    if (w) {
        KZG_HIP(c, c->stage2.ensure(7 * sizeof(uint32_t) * size_t(w)));
        KZG_HIP(c, hipMemcpyAsync(c->stage2.p, c->bp_moved.p + 2, 7 * sizeof(uint32_t) * size_t(w),
                                  hipMemcpyDeviceToHost, c->stream));
        KZG_HIP(c, hipStreamSynchronize(c->stream));
        std::memcpy(out, c->stage2.p, 7 * sizeof(uint32_t) * size_t(w));
    }
    return KZG_OK;
}

int kzg_broadphase_download(kzg_ctx* c, kzg_bp_node* out, uint32_t num_nodes) {
    if (!c || (num_nodes && !out)) return KZG_ERR_ARGUMENT;
    if (num_nodes != c->bp_num_nodes) return fail(c, KZG_ERR_ARGUMENT, "kzg_broadphase_download: image size");
    hipSetDevice(c->device);
    return download(c, reinterpret_cast<float4*>(out), c->bp_nodes, 4 * size_t(num_nodes));
}

int kzg_download_pairs(kzg_ctx* c, uint32_t* out, uint32_t n) {
    if (!c || (n && !out)) return KZG_ERR_ARGUMENT;
    if (n > c->npairs) return fail(c, KZG_ERR_ARGUMENT, "kzg_download_pairs: more pairs than the context holds");
    hipSetDevice(c->device);
    return download(c, reinterpret_cast<uint2*>(out), c->d_pairs, n);
}

}  // extern "C"
