// schedule.cpp - islands, first-fit SIMD batches and level decomposition.
//
// Reference behaviour reproduced:
//   * BuildIslands (src/World/PhysicsWorld.cpp:587-664): compact indices for
//     dynamic bodies in slot order, union by rank with path halving over
//     dynamic-dynamic contacts, islands numbered by first root appearance,
//     a contact belongs to the island of A, else of B.
//   * The first-fit batch builder of SolveIsland (:677-717): for each point
//     in island contact order (sensors skipped), the first batch with fewer
//     than four lanes none of whose lanes shares body A or B (static bodies
//     included) receives it, else a new batch is opened.  The reference scans
//     every batch from the first; here the non-full batches sit in ordered
//     sets (one global, one per non-dynamic body listing the non-full batches
//     without that body), which yields the same first fit without the
//     quadratic scan.
// The device needs levels: walking a sequence in order, a point's level is
// 1 + the largest last level of its dynamic bodies.  Points of one level
// share no dynamic body; for every dynamic body the order of its updates is
// the sequence order, which is what makes the sweep bit-identical.
#include "schedule.h"

#include <algorithm>
#include <queue>
#include <unordered_map>

namespace kz {

namespace {

struct Batch {
    uint32_t a[4], b[4], pt[4];
    uint32_t n = 0;
    bool has(uint32_t x) const {
        for (uint32_t k = 0; k < n; ++k)
            if (a[k] == x || b[k] == x) return true;
        return false;
    }
};

// Bucket (group, level) -> flat arrays.
void flatten(const std::vector<uint32_t>& pts, const std::vector<uint32_t>& grp_of, const std::vector<uint32_t>& lvl_of,
             uint32_t groups, uint32_t num_points, LevelSchedule& out) {
    std::vector<uint32_t> depth(groups, 0);
    for (size_t i = 0; i < pts.size(); ++i) depth[grp_of[i]] = std::max(depth[grp_of[i]], lvl_of[i]);
    out.grp.assign(groups, U2{0, 0});
    std::vector<uint32_t> base(groups, 0);
    uint32_t total = 0;
    out.max_depth = 0;
    for (uint32_t g = 0; g < groups; ++g) {
        base[g] = total;
        out.grp[g] = U2{total, total + depth[g]};
        total += depth[g];
        out.max_depth = std::max(out.max_depth, depth[g]);
    }
    out.levels = total;
    out.off.assign(total + 1, 0);
    for (size_t i = 0; i < pts.size(); ++i) out.off[base[grp_of[i]] + lvl_of[i] - 1 + 1] += 1;
    out.max_width = 0;
    for (uint32_t l = 0; l < total; ++l) {
        out.max_width = std::max(out.max_width, out.off[l + 1]);
        out.off[l + 1] += out.off[l];
    }
    out.idx.assign(pts.size(), 0);
    std::vector<uint32_t> fill(out.off.begin(), out.off.end() - 1);
    out.pos.assign(num_points, 0xFFFFFFFFu);
    for (size_t i = 0; i < pts.size(); ++i) {
        const uint32_t e = fill[base[grp_of[i]] + lvl_of[i] - 1]++;
        out.idx[e] = pts[i];
        out.pos[pts[i]] = e;
    }
}

}  // namespace

void ScheduleBuilder::build(const ScheduleInput& in, ScheduleOutput& out) {
    const uint32_t nb = in.num_bodies, nc = in.num_constraints;
    uint32_t num_points = 0;
    for (uint32_t c = 0; c < nc; ++c) num_points += in.info[c].w & 0xFFu;
    const uint8_t* dyn = in.dynamic;
    // ---- islands -----------------------------------------------------------
    dyn_index_.assign(nb, 0xFFFFFFFFu);
    uint32_t nd = 0;
    for (uint32_t i = 0; i < nb; ++i)
        if (dyn[i]) dyn_index_[i] = nd++;
    parent_.resize(nd);
    rank_.assign(nd, 0);
    for (uint32_t i = 0; i < nd; ++i) parent_[i] = i;
    auto find = [&](uint32_t x) {
        while (parent_[x] != x) {
            parent_[x] = parent_[parent_[x]];
            x = parent_[x];
        }
        return x;
    };
    for (uint32_t c = 0; c < nc; ++c) {
        const U4& ci = in.info[c];
        if (!dyn[ci.x] || !dyn[ci.y]) continue;
        uint32_t a = find(dyn_index_[ci.x]), b = find(dyn_index_[ci.y]);
        if (a == b) continue;
        if (rank_[a] < rank_[b]) std::swap(a, b);
        parent_[b] = a;
        if (rank_[a] == rank_[b]) ++rank_[a];
    }
    root_island_.assign(nd, 0xFFFFFFFFu);
    uint32_t ni = 0;
    for (uint32_t i = 0; i < nd; ++i) {
        uint32_t r = find(i);
        if (root_island_[r] == 0xFFFFFFFFu) root_island_[r] = ni++;
    }
    island_of_con_.assign(nc, 0xFFFFFFFFu);
    island_points_.assign(ni, 0);
    if (island_cons_.size() < ni) island_cons_.resize(ni);
    for (uint32_t i = 0; i < ni; ++i) island_cons_[i].clear();
    for (uint32_t c = 0; c < nc; ++c) {
        const U4& ci = in.info[c];
        uint32_t is = 0xFFFFFFFFu;
        if (dyn[ci.x]) is = root_island_[find(dyn_index_[ci.x])];
        else if (dyn[ci.y]) is = root_island_[find(dyn_index_[ci.y])];
        island_of_con_[c] = is;
        if (is != 0xFFFFFFFFu) {
            island_cons_[is].push_back(c);
            island_points_[is] += ci.w & 0xFFu;
        }
    }
    out.islands = ni;
    // ---- groups: largest islands first, each to the least-loaded group ------
    const uint32_t groups = std::max<uint32_t>(1, std::min<uint32_t>(ni, in.max_groups));
    out.groups = groups;
    group_of_island_.assign(ni, 0);
    {
        std::vector<uint32_t> order(ni);
        for (uint32_t i = 0; i < ni; ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(),
                         [&](uint32_t a, uint32_t b) { return island_points_[a] > island_points_[b]; });
        using Load = std::pair<uint64_t, uint32_t>;
        std::priority_queue<Load, std::vector<Load>, std::greater<Load>> heap;
        for (uint32_t g = 0; g < groups; ++g) heap.push({0, g});
        for (uint32_t is : order) {
            Load l = heap.top();
            heap.pop();
            group_of_island_[is] = l.second;
            l.first += island_points_[is];
            heap.push(l);
        }
    }
    // ---- warm start / bias sequence: constraint order ----------------------
    std::vector<uint32_t> pts, grp, lvl;
    last_.assign(nb, 0);
    for (uint32_t c = 0; c < nc; ++c) {
        const U4& ci = in.info[c];
        if (island_of_con_[c] == 0xFFFFFFFFu) continue;
        const uint32_t g = group_of_island_[island_of_con_[c]];
        const uint32_t np = ci.w & 0xFFu;
        for (uint32_t k = 0; k < np; ++k) {
            uint32_t l = 1 + std::max(dyn[ci.x] ? last_[ci.x] : 0u, dyn[ci.y] ? last_[ci.y] : 0u);
            if (dyn[ci.x]) last_[ci.x] = l;
            if (dyn[ci.y]) last_[ci.y] = l;
            pts.push_back(ci.z + k);
            grp.push_back(g);
            lvl.push_back(l);
        }
    }
    flatten(pts, grp, lvl, groups, num_points, out.ws);
    // ---- velocity sequence: first-fit batches per island --------------------
    pts.clear(); grp.clear(); lvl.clear();
    if (in.use_simd) {
        last_.assign(nb, 0);
        std::vector<Batch> batches;
        // Non-full batches as ordered lists with lazy deletion: `open` holds
        // every batch in creation order (full ones are skipped and dropped
        // from the front); open_wo[h] holds, for a non-dynamic body h, the
        // batches created without h.  Batches h later joins or that fill up
        // are skipped on the way and compacted away when they dominate.
        struct OpenList {
            std::vector<uint32_t> ids;
            size_t head = 0;
        };
        OpenList open;
        // One list per non-dynamic body met in the island (a handful: the
        // static floor and walls), searched linearly.
        std::vector<std::pair<uint32_t, OpenList>> open_wo;
        for (uint32_t is = 0; is < ni; ++is) {
            batches.clear();
            open.ids.clear();
            open.head = 0;
            open_wo.clear();
            auto open_without = [&](uint32_t h) -> OpenList& {
                for (auto& kv : open_wo)
                    if (kv.first == h) return kv.second;
                open_wo.emplace_back(h, OpenList{});
                OpenList& l = open_wo.back().second;
                for (size_t k = open.head; k < open.ids.size(); ++k) {
                    const Batch& B = batches[open.ids[k]];
                    if (B.n < 4 && !B.has(h)) l.ids.push_back(open.ids[k]);
                }
                return l;
            };
            // First usable batch of a list for (a, b); drops dead entries.
            auto first_fit = [&](OpenList& l, uint32_t a, uint32_t b) -> uint32_t {
                while (l.head < l.ids.size() && batches[l.ids[l.head]].n >= 4) ++l.head;
                size_t dead = 0;
                uint32_t found = 0xFFFFFFFFu;
                for (size_t k = l.head; k < l.ids.size(); ++k) {
                    const Batch& B = batches[l.ids[k]];
                    if (B.n >= 4) { ++dead; continue; }
                    if (B.has(a) || B.has(b)) continue;
                    found = l.ids[k];
                    break;
                }
                if (dead > 64 && dead * 2 > l.ids.size() - l.head) {
                    size_t w = 0;
                    for (size_t k = l.head; k < l.ids.size(); ++k)
                        if (batches[l.ids[k]].n < 4) l.ids[w++] = l.ids[k];
                    l.ids.resize(w);
                    l.head = 0;
                }
                return found;
            };
            for (uint32_t c : island_cons_[is]) {
                const U4& ci = in.info[c];
                if ((ci.w >> 8) & 1u) continue;  // sensor
                const uint32_t a = ci.x, b = ci.y, np = ci.w & 0xFFu;
                for (uint32_t k = 0; k < np; ++k) {
                    OpenList* cand = &open;
                    if (!dyn[a]) cand = &open_without(a);
                    else if (!dyn[b]) cand = &open_without(b);
                    uint32_t placed = first_fit(*cand, a, b);
                    if (placed == 0xFFFFFFFFu) {
                        Batch nb2;
                        nb2.a[0] = a; nb2.b[0] = b; nb2.pt[0] = ci.z + k; nb2.n = 1;
                        const uint32_t id = uint32_t(batches.size());
                        batches.push_back(nb2);
                        open.ids.push_back(id);
                        for (auto& kv : open_wo)
                            if (kv.first != a && kv.first != b) kv.second.ids.push_back(id);
                    } else {
                        Batch& B = batches[placed];
                        B.a[B.n] = a; B.b[B.n] = b; B.pt[B.n] = ci.z + k; ++B.n;
                        // Entries of open_wo[a] / open_wo[b] now containing a or
                        // b are skipped by the conflict test on later visits.
                    }
                }
            }
            const uint32_t g = group_of_island_[is];
            for (const Batch& B : batches)
                for (uint32_t j = 0; j < B.n; ++j) {
                    const uint32_t a = B.a[j], b = B.b[j];
                    uint32_t l = 1 + std::max(dyn[a] ? last_[a] : 0u, dyn[b] ? last_[b] : 0u);
                    if (dyn[a]) last_[a] = l;
                    if (dyn[b]) last_[b] = l;
                    pts.push_back(B.pt[j]);
                    grp.push_back(g);
                    lvl.push_back(l);
                }
        }
    }
    if (in.use_simd) flatten(pts, grp, lvl, groups, num_points, out.vel);
    else out.vel = out.ws;  // the scalar path sweeps in constraint order, sensors included
}

}  // namespace kz
