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
//     lists (one global, one per non-dynamic body listing the non-full
//     batches without that body), which yields the same first fit without
//     the quadratic scan.
// The device needs levels: walking a sequence in order, a point's level is
// 1 + the largest last level of its dynamic bodies.  Points of one level
// share no dynamic body; for every dynamic body the order of its updates is
// the sequence order, which is what makes the sweep bit-identical.
// A point whose two dynamic bodies were both last updated by the previous
// point of the same pair follows that point in one thread (a unit) instead
// of opening the next level: a manifold's points do so in the warm-start
// order and mostly in the first-fit order.  Units of two points shorten the
// level chain by ~30%.
//
// The warm-start/bias sequence (constraint order) and the velocity sequence
// (first-fit order) are independent once the islands are known; above a few
// thousand points the first runs on a helper thread beside the second.  A
// deferred build swaps them: the caller builds the warm-start sequence and
// gets it back while the velocity sequence, the longer one, is still being
// built on the helper, so the device can sweep the warm start meanwhile.
#include "schedule.h"

#include <algorithm>
#include <queue>

namespace kz {

namespace {

constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr uint32_t kParallelPoints = 16384;  // below this the helper thread costs more than it saves

inline bool holds(const FitBatch& B, uint32_t x) {
    uint32_t hit = 0;
    for (int k = 0; k < 8; ++k) hit |= uint32_t(B.body[k] == x);
    return hit != 0;
}
inline bool holds_either(const FitBatch& B, uint32_t a, uint32_t b) {
    uint32_t hit = 0;
    for (int k = 0; k < 8; ++k) hit |= uint32_t(B.body[k] == a) | uint32_t(B.body[k] == b);
    return hit != 0;
}

}  // namespace

ScheduleBuilder::~ScheduleBuilder() {
    if (helper_.joinable()) {
        {
            std::lock_guard<std::mutex> lk(m_);
            stop_ = true;
        }
        cv_.notify_all();
        helper_.join();
    }
}

void ScheduleBuilder::run_helper(std::function<void()> job) {
    if (!helper_.joinable())
        helper_ = std::thread([this] {
            std::unique_lock<std::mutex> lk(m_);
            for (;;) {
                cv_.wait(lk, [this] { return stop_ || job_; });
                if (stop_) return;
                std::function<void()> j = std::move(job_);
                job_ = nullptr;
                lk.unlock();
                j();
                lk.lock();
                busy_ = false;
                cv_.notify_all();
            }
        });
    {
        std::lock_guard<std::mutex> lk(m_);
        job_ = std::move(job);
        busy_ = true;
    }
    cv_.notify_all();
}

void ScheduleBuilder::wait_helper() {
    std::unique_lock<std::mutex> lk(m_);
    cv_.wait(lk, [this] { return !busy_; });
}

void ScheduleBuilder::Seq::reserve(uint32_t points) {
    for (auto* v : {&pts, &uid, &lvl}) v->resize(points);
    n = nu = 0;
}

// Groups' level ranges (the shared head of both flatten variants).
uint32_t ScheduleBuilder::level_bases(Seq& s, uint32_t groups, LevelSchedule& out) {
    out.largest_unit = s.nu ? s.largest : 0;
    out.grp.assign(groups, U2{0, 0});
    s.base.assign(groups, 0);
    uint32_t total = 0;
    out.max_depth = 0;
    for (uint32_t g = 0; g < groups; ++g) {
        s.base[g] = total;
        out.grp[g] = U2{total, total + s.depth[g]};
        total += s.depth[g];
        out.max_depth = std::max(out.max_depth, s.depth[g]);
    }
    out.levels = total;
    out.units = s.nu;
    out.off.assign(total + 1, 0);
    return total;
}

void ScheduleBuilder::prefix_levels(Seq& s, uint32_t total, LevelSchedule& out) {
    out.max_width = 0;
    for (uint32_t l = 0; l < total; ++l) {
        out.max_width = std::max(out.max_width, out.off[l + 1]);
        out.off[l + 1] += out.off[l];
    }
    s.fill.assign(out.off.begin(), out.off.end() - 1);
}

void ScheduleBuilder::place_points(Seq& s, uint32_t num_points, LevelSchedule& out) {
    for (uint32_t k = 0; k < s.nu; ++k) out.unit[k + 1] += out.unit[k];
    out.device_bucketed = false;
    out.g0_points = out.grp.empty() ? 0u : out.unit[out.off[out.grp[0].y]] - out.unit[out.off[out.grp[0].x]];
    out.idx.resize(s.n);
    out.pos.assign(num_points, kNone);
    for (uint32_t i = 0; i < s.n; ++i) {
        const uint32_t e = out.unit[s.slot[s.uid[i] >> 2]] + (s.uid[i] & 3u);
        out.idx[e] = s.pts[i];
        out.pos[s.pts[i]] = e;
    }
}

// Bucket units by (group, level) -> flat arrays, from the per-unit tables;
// a unit's points keep their sequence order.
void ScheduleBuilder::flatten(Seq& s, uint32_t groups, uint32_t num_points, LevelSchedule& out) {
    const uint32_t total = level_bases(s, groups, out);
    s.slot.resize(s.nu);
    for (uint32_t u = 0; u < s.nu; ++u) {
        const uint32_t l = s.base[s.ugrp[u]] + s.ulvl[u];
        s.slot[u] = l;  // the level (+1) for now, the slot below
        out.off[l] += 1;
    }
    prefix_levels(s, total, out);
    out.unit.resize(s.nu + 1);
    out.unit[0] = 0;
    for (uint32_t u = 0; u < s.nu; ++u) {
        const uint32_t k = s.fill[s.slot[u] - 1]++;
        s.slot[u] = k;
        out.unit[k + 1] = s.usize[u];
    }
    place_points(s, num_points, out);
}

// The same from the per-point entries and the group runs (no unit tables):
// a unit's first point (index 0 in it) places the unit, its last sets the
// size.  Units are numbered in the order of their first points.
void ScheduleBuilder::flatten_points(Seq& s, uint32_t groups, uint32_t num_points, LevelSchedule& out) {
    // Deepest level of each group and the largest unit.
    s.depth.assign(groups, 0);
    uint32_t maxj = 0, i0 = 0;
    for (const U2& r : s.runs) {
        uint32_t d = s.depth[r.y];
        for (uint32_t i = i0; i < r.x; ++i) {
            d = std::max(d, s.lvl[i]);
            maxj = std::max(maxj, s.uid[i] & 3u);
        }
        s.depth[r.y] = d;
        i0 = r.x;
    }
    s.largest = maxj + 1;
    const uint32_t total = level_bases(s, groups, out);
    uint32_t i = 0;
    for (const U2& r : s.runs) {
        uint32_t* off = out.off.data() + s.base[r.y];
        for (; i < r.x; ++i) off[s.lvl[i]] += (s.uid[i] & 3u) == 0 ? 1u : 0u;
    }
    prefix_levels(s, total, out);
    s.slot.resize(s.nu);
    out.unit.assign(s.nu + 1, 0);
    i = 0;
    for (const U2& r : s.runs) {
        uint32_t* fill = s.fill.data() + s.base[r.y] - 1;
        for (; i < r.x; ++i) {
            const uint32_t u = s.uid[i] >> 2, j = s.uid[i] & 3u, start = j == 0 ? 1u : 0u;
            uint32_t& f = fill[s.lvl[i]];
            const uint32_t k = start ? f : s.slot[u];
            f += start;
            s.slot[u] = k;
            out.unit[k + 1] = j + 1;
        }
    }
    place_points(s, num_points, out);
}

void ScheduleBuilder::islands(const ScheduleInput& in, ScheduleOutput& out) {
    const uint32_t nc = in.num_constraints;
    uint32_t ni = 0;
    if (in.island_of_con) {
        ni = in.num_islands;
        island_of_con_.assign(in.island_of_con, in.island_of_con + nc);
    } else {
        ni = union_find(in);
    }
    // The constraints of each island in constraint order (counting sort).
    island_points_.assign(ni, 0);
    island_off_.assign(ni + 1, 0);
    for (uint32_t c = 0; c < nc; ++c) {
        const uint32_t is = island_of_con_[c];
        if (is != kNone) {
            island_off_[is + 1] += 1;
            island_points_[is] += in.info[c].w & 0xFFu;
        }
    }
    for (uint32_t i = 0; i < ni; ++i) island_off_[i + 1] += island_off_[i];
    island_cons_.resize(island_off_[ni]);
    order_.assign(island_off_.begin(), island_off_.end() - 1);
    for (uint32_t c = 0; c < nc; ++c)
        if (island_of_con_[c] != kNone) island_cons_[order_[island_of_con_[c]]++] = c;
    out.islands = ni;
    // Groups: largest islands first, each to the least-loaded group.
    const uint32_t groups = std::max<uint32_t>(1, std::min<uint32_t>(ni, in.max_groups));
    out.groups = groups;
    groups_ = groups;
    group_of_island_.assign(ni, 0);
    order_.resize(ni);
    for (uint32_t i = 0; i < ni; ++i) order_[i] = i;
    std::stable_sort(order_.begin(), order_.end(),
                     [&](uint32_t a, uint32_t b) { return island_points_[a] > island_points_[b]; });
    using Load = std::pair<uint64_t, uint32_t>;
    std::priority_queue<Load, std::vector<Load>, std::greater<Load>> heap;
    for (uint32_t g = 0; g < groups; ++g) heap.push({0, g});
    for (uint32_t is : order_) {
        Load l = heap.top();
        heap.pop();
        group_of_island_[is] = l.second;
        l.first += island_points_[is];
        heap.push(l);
    }
}

// Union by rank with path halving over the dynamic-dynamic contacts; islands
// numbered by the first appearance of their root in slot order.  Fills
// island_of_con_ and returns the island count.
uint32_t ScheduleBuilder::union_find(const ScheduleInput& in) {
    const uint32_t nb = in.num_bodies, nc = in.num_constraints;
    const uint8_t* dyn = in.dynamic;
    dyn_index_.assign(nb, kNone);
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
    root_island_.assign(nd, kNone);
    uint32_t ni = 0;
    for (uint32_t i = 0; i < nd; ++i) {
        const uint32_t r = find(i);
        if (root_island_[r] == kNone) root_island_[r] = ni++;
    }
    for (uint32_t i = 0; i < nd; ++i) root_island_[i] = root_island_[find(i)];
    island_of_con_.resize(nc);
    for (uint32_t c = 0; c < nc; ++c) {
        const U4& ci = in.info[c];
        uint32_t is = kNone;
        if (dyn[ci.x]) is = root_island_[dyn_index_[ci.x]];
        else if (dyn[ci.y]) is = root_island_[dyn_index_[ci.y]];
        island_of_con_[c] = is;
    }
    return ni;
}

// Levels of the warm-start/bias sequence (constraint order, every
// constraint with a dynamic body).  Needs only the body types, so it can
// start before the islands are known; the groups are filled in afterwards.
void ScheduleBuilder::ws_levels(const ScheduleInput& in) {
    // The points of a constraint follow each other in this order, so its
    // units are runs of up to `cap` consecutive points: one level per run.
    const uint8_t* dyn = in.dynamic;
    Seq& s = ws_;
    ws_body_level_.assign(in.num_bodies, 0);
    uint32_t* bl = ws_body_level_.data();
    const uint32_t cap = std::min(std::max(in.use_simd ? in.max_unit_ws : in.max_unit_vel, 1u), kMaxUnit);
    uint32_t n = 0, nu = 0, np_max = 0;
    for (uint32_t c = 0; c < in.num_constraints; ++c) {
        const U4& ci = in.info[c];
        const bool da = dyn[ci.x] != 0, db = dyn[ci.y] != 0;
        if (!da && !db) continue;
        // With sleeping bodies a constraint can be in no island: the
        // reference never solves it.
        if (in.island_of_con && in.island_of_con[c] == kNone) continue;
        const uint32_t np = ci.w & 0xFFu;
        for (uint32_t k0 = 0; k0 < np; k0 += cap) {
            const uint32_t l = 1 + std::max(da ? bl[ci.x] : 0u, db ? bl[ci.y] : 0u);
            if (da) bl[ci.x] = l;
            if (db) bl[ci.y] = l;
            const uint32_t m = std::min(cap, np - k0);
            np_max = std::max(np_max, m);
            s.ulvl[nu] = l;
            s.usize[nu] = m;
            ws_con_[nu] = c;
            for (uint32_t j = 0; j < m; ++j) {
                s.pts[n] = ci.z + k0 + j;
                s.uid[n] = (nu << 2) | j;
                ++n;
            }
            ++nu;
        }
    }
    s.n = n;
    s.nu = nu;
    s.largest = np_max;
}

// First-fit batches of every island, in island order, into one batch array;
// island i's batches end at fit_island_end_[i].
void ScheduleBuilder::vel_fit(const ScheduleInput& in) {
    const uint8_t* dyn = in.dynamic;
    const uint32_t cap = std::min(std::max(in.max_unit_vel, 1u), kMaxUnit);
    // Non-full batches as ordered lists with lazy deletion: `open` holds
    // every batch in creation order (full ones are skipped and dropped from
    // the front); open_wo[h] holds, for a non-dynamic body h, the batches
    // created without h.  Batches h later joins or that fill up are skipped
    // on the way and compacted away when they dominate.
    struct OpenList {
        std::vector<uint32_t> ids;
        size_t head = 0;
    };
    OpenList open;
    // One list per non-dynamic body met in the island (a handful: the static
    // floor and walls), searched linearly.
    std::vector<std::pair<uint32_t, OpenList>> open_wo;
    size_t wo_used = 0;
    std::vector<FitBatch>& batches = batches_;
    const uint32_t ni = uint32_t(island_points_.size());
    batches.clear();
    for (uint32_t is = 0; is < ni; ++is) {
        open.ids.clear();
        open.head = 0;
        wo_used = 0;
        auto open_without = [&](uint32_t h) -> OpenList& {
            for (size_t k = 0; k < wo_used; ++k)
                if (open_wo[k].first == h) return open_wo[k].second;
            if (wo_used == open_wo.size()) open_wo.emplace_back();
            auto& kv = open_wo[wo_used++];
            kv.first = h;
            OpenList& l = kv.second;
            l.ids.clear();
            l.head = 0;
            for (size_t k = open.head; k < open.ids.size(); ++k) {
                const FitBatch& B = batches[open.ids[k]];
                if (B.n < 4 && !holds(B, h)) l.ids.push_back(open.ids[k]);
            }
            return l;
        };
        // First usable batch of a list for (a, b); drops dead entries.  An
        // entry is dead when its batch is full or, in the list kept for a
        // non-dynamic body h, holds h: every point that searches that list
        // touches h, so neither batch can take it again.  (Without the
        // second rule every floor point rescans the batches earlier floor
        // points joined.)
        //   `from` carries a constraint's search from point to point: every
        // entry before the batch its last point took was dead or held a or b,
        // and still is, and that batch now holds both, so the next point's
        // search starts after it (after the list's end when the point opened
        // a new batch: no entry could take it).
        auto first_fit = [&](OpenList& l, uint32_t h, uint32_t a, uint32_t b, size_t& from) -> uint32_t {
            auto dead = [&](const FitBatch& B) { return B.n >= 4 || (h != kNone && holds(B, h)); };
            while (l.head < l.ids.size() && dead(batches[l.ids[l.head]])) ++l.head;
            size_t ndead = 0;
            uint32_t found = kNone;
            size_t k = std::max(from, l.head);
            for (; k < l.ids.size(); ++k) {
                const FitBatch& B = batches[l.ids[k]];
                if (dead(B)) {
                    ++ndead;
                    continue;
                }
                if (holds_either(B, a, b)) continue;
                found = l.ids[k];
                break;
            }
            from = found == kNone ? l.ids.size() : k + 1;
            if (ndead > 64 && ndead * 2 > l.ids.size() - l.head) {
                size_t w = 0;
                for (size_t q = l.head; q < l.ids.size(); ++q)
                    if (!dead(batches[l.ids[q]])) l.ids[w++] = l.ids[q];
                l.ids.resize(w);
                l.head = 0;
                from = 0;
            }
            return found;
        };
        // A point placed in the batch right after the one holding the
        // previous point of its constraint follows that point with no update
        // of either body between them (the lanes of a batch share no body):
        // it joins that point's unit while the unit has room.
        for (uint32_t j = island_off_[is]; j < island_off_[is + 1]; ++j) {
            const U4& ci = in.info[island_cons_[j]];
            if ((ci.w >> 8) & 1u) continue;  // sensor
            const uint32_t a = ci.x, b = ci.y, np = ci.w & 0xFFu;
            OpenList* cand = &open;
            uint32_t h = kNone;
            if (!dyn[a]) cand = &open_without(h = a);
            else if (!dyn[b]) cand = &open_without(h = b);
            uint32_t prev = kNone, prev_lane = 0, run = 0;
            size_t from = 0;
            for (uint32_t k = 0; k < np; ++k) {
                const uint32_t placed = first_fit(*cand, h, a, b, from);
                const uint32_t id = placed == kNone ? uint32_t(batches.size()) : placed;
                const bool join = (prev != kNone) & (id == prev + 1) & (run < cap);
                run = join ? run + 1 : 1;
                const uint32_t jbits = join ? 1u | (prev_lane << 8) : 0u;
                if (placed == kNone) {
                    FitBatch nb;
                    nb.body[0] = a;
                    nb.body[4] = b;
                    for (int q = 1; q < 4; ++q) nb.body[q] = nb.body[4 + q] = kNone;
                    nb.pt[0] = ci.z + k;
                    nb.n = 1;
                    nb.join = jbits;
                    batches.push_back(nb);
                    open.ids.push_back(id);
                    for (size_t q = 0; q < wo_used; ++q)
                        if (open_wo[q].first != a && open_wo[q].first != b) open_wo[q].second.ids.push_back(id);
                    prev_lane = 0;
                } else {
                    // Entries of open_wo[a] / open_wo[b] now holding a or b
                    // are skipped by the conflict test on later visits.
                    FitBatch& B = batches[placed];
                    B.body[B.n] = a;
                    B.body[4 + B.n] = b;
                    B.pt[B.n] = ci.z + k;
                    B.join |= (jbits & 1u) << B.n | (jbits >> 8) << (8 + 2 * B.n);
                    prev_lane = B.n;
                    ++B.n;
                }
                prev = id;
            }
        }
        fit_island_end_[is] = uint32_t(batches.size());
    }
}


// Levels of the batch sequence, unit by unit.  A joining lane takes the
// unit of the lane it names in the previous batch (read back from uid).
void ScheduleBuilder::vel_levels(const ScheduleInput& in, bool to_sink) {
    const uint8_t* dyn = in.dynamic;
    Seq& s = vel_;
    const uint32_t ni = uint32_t(island_points_.size());
    const FitBatch* batches = batches_.data();
    uint32_t* last = vel_body_level_.data();
    sunk_ = to_sink;
    uint32_t *pts = s.pts.data(), *uid = s.uid.data(), *lvl = s.lvl.data();
    if (to_sink) {
        pts = sink_;
        uid = sink_ + sink_cap_;
        lvl = sink_ + 2 * sink_cap_;
    }
    uint32_t n = 0, nu = 0, x = 0;
    // Per group: the deepest level; group 0: units per level (its levels
    // cannot outnumber the batches); every other group counts into a scratch
    // row that is never read.
    vel_depth_.assign(groups_, 0);
    vel_g0_count_.assign(size_t(batches_.size()) + 2, 0);
    std::vector<uint32_t>& scratch = vel_group_base_;  // reused below as the group bases
    scratch.assign(size_t(batches_.size()) + 2, 0);
    uint32_t jmax = 0;
    if (!to_sink) vel_rank_.resize(s.pts.size());
    uint32_t* rank = to_sink ? sink_ + 3 * sink_cap_ : vel_rank_.data();
    for (uint32_t is = 0; is < ni; ++is) {
        const uint32_t n0 = n;
        uint32_t pb = n;  // first point of the previous batch
        const uint32_t grp = group_of_island_[is];
        uint32_t* count = grp == 0 ? vel_g0_count_.data() : scratch.data();
        const uint32_t rank_mask = grp == 0 ? 0u : kNone;  // group 0's units take their rank
        uint32_t deepest = 0;
        for (; x < fit_island_end_[is]; ++x) {
            const FitBatch& B = batches[x];
            const uint32_t cb = n;
            for (uint32_t q = 0; q < B.n; ++q) {
                const uint32_t a = B.body[q], b = B.body[4 + q];
                const bool da = dyn[a] != 0, db = dyn[b] != 0, join = (B.join >> q) & 1u;
                // Masks, not branches: the join is taken about a fifth of the
                // time at random.  (For a first batch pb + lane is the current
                // point, whose entry is read and not used.)
                const uint32_t top = std::max(da ? last[a] : 0u, db ? last[b] : 0u);
                const uint32_t jm = 0u - uint32_t(join);
                const uint32_t l = top + 1 - uint32_t(join);
                const uint32_t ux = ((uid[pb + ((B.join >> (8 + 2 * q)) & 3u)] + 1) & jm) | ((nu << 2) & ~jm);
                nu += 1 - uint32_t(join);
                if (da) last[a] = l;
                if (db) last[b] = l;
                pts[n] = B.pt[q];
                uid[n] = ux;
                lvl[n] = l;
                // A unit's place among its level's units (group 0): the count
                // of the level before it.
                const uint32_t before = count[l];
                count[l] = before + 1 - uint32_t(join);
                rank[n] = before | rank_mask;
                deepest = deepest > l ? deepest : l;
                jmax = jmax > (ux & 3u) ? jmax : (ux & 3u);
                ++n;
            }
            pb = cb;
        }
        if (n > n0) s.runs.push_back(U2{n, grp});
        vel_depth_[grp] = std::max(vel_depth_[grp], deepest);
    }
    s.n = n;
    s.nu = nu;
    vel_jmax_ = jmax;
}

// The host's share of a velocity sequence bucketed on the device: group
// level ranges, group 0's level offsets (for the launch plan), counts.
void ScheduleBuilder::vel_lite(uint32_t num_points, LevelSchedule& out) {
    (void)num_points;
    Seq& s = vel_;
    out.device_bucketed = true;
    out.largest_unit = s.nu ? vel_jmax_ + 1 : 0;
    out.grp.assign(groups_, U2{0, 0});
    vel_group_base_.assign(groups_, 0);
    uint32_t total = 0;
    out.max_depth = 0;
    for (uint32_t g = 0; g < groups_; ++g) {
        vel_group_base_[g] = total;
        out.grp[g] = U2{total, total + vel_depth_[g]};
        total += vel_depth_[g];
        out.max_depth = std::max(out.max_depth, vel_depth_[g]);
    }
    out.levels = total;
    out.units = s.nu;
    const uint32_t d0 = groups_ ? vel_depth_[0] : 0u;
    out.off.assign(d0 + 1, 0);
    out.max_width = 0;
    for (uint32_t l = 0; l < d0; ++l) {
        const uint32_t c = vel_g0_count_[l + 1];
        out.max_width = std::max(out.max_width, c);
        out.off[l + 1] = out.off[l] + c;
    }
    uint32_t g0 = 0;
    for (uint32_t is = 0; is < uint32_t(island_points_.size()); ++is)
        if (group_of_island_[is] == 0) g0 += island_points_[is];
    out.g0_points = g0;
    out.idx.clear();
    out.pos.clear();
    out.unit.clear();
}

VelSequence ScheduleBuilder::velocity_sequence() const {
    VelSequence v;
    v.pts = vel_.pts.data();
    v.uid = vel_.uid.data();
    v.lvl = vel_.lvl.data();
    v.rank = vel_rank_.data();
    if (sunk_) {
        v.pts = sink_;
        v.uid = sink_ + sink_cap_;
        v.lvl = sink_ + 2 * sink_cap_;
        v.rank = sink_ + 3 * sink_cap_;
        v.in_sink = true;
        v.stride = sink_cap_;
    }
    v.n = vel_.n;
    v.units = vel_.nu;
    v.runs = vel_.runs.data();
    v.num_runs = uint32_t(vel_.runs.size());
    v.group_base = vel_group_base_.data();
    v.groups = groups_;
    v.g0_count = vel_g0_count_.data() + 1;
    v.g0_levels = groups_ ? vel_depth_[0] : 0u;
    v.levels = 0;
    for (uint32_t g = 0; g < groups_; ++g) v.levels += vel_depth_[g];
    return v;
}

void ScheduleBuilder::build(const ScheduleInput& in, ScheduleOutput& out) {
    start(in, out, kDeferNever);
    finish();
}

bool ScheduleBuilder::start(const ScheduleInput& in_arg, ScheduleOutput& out, uint32_t defer) {
    finish();
    in_ = in_arg;
    const ScheduleInput& in = in_;
    const uint32_t nc = in.num_constraints;
    uint32_t num_points = 0;
    for (uint32_t c = 0; c < nc; ++c) num_points += in.info[c].w & 0xFFu;
    for (Seq* s : {&ws_, &vel_}) s->reserve(num_points);
    for (auto* v : {&ws_.ulvl, &ws_.ugrp, &ws_.usize}) v->resize(num_points);
    ws_con_.resize(num_points);
    // The warm-start sequence: levels (body types only), then, once the
    // islands are known, groups and buckets.
    groups_ready_.store(false, std::memory_order_relaxed);
    using clock = std::chrono::steady_clock;
    auto ws_chain = [this, &in, &out, num_points](bool wait) {
        const auto t0 = clock::now();
        ws_levels(in);
        if (wait)
            while (!groups_ready_.load(std::memory_order_acquire)) std::this_thread::yield();
        ws_.depth.assign(groups_, 0);
        for (uint32_t u = 0; u < ws_.nu; ++u) {
            const uint32_t g = group_of_island_[island_of_con_[ws_con_[u]]];
            ws_.ugrp[u] = g;
            ws_.depth[g] = std::max(ws_.depth[g], ws_.ulvl[u]);
        }
        flatten(ws_, groups_, num_points, out.ws);
        ws_ms = std::chrono::duration<double, std::milli>(clock::now() - t0).count();
    };
    auto vel_chain = [this, &in, &out, num_points](bool may_defer_bucketing) {
        vel_body_level_.assign(in.num_bodies, 0);
        vel_.runs.clear();
        fit_island_end_.assign(island_points_.size(), 0);
        vel_fit(in);
        // Device bucketing pays where one group dominates (the launch plan
        // and thread counts are taken from group 0's levels).
        uint32_t g0 = 0;
        for (uint32_t is = 0; is < uint32_t(island_points_.size()); ++is)
            if (group_of_island_[is] == 0) g0 += island_points_[is];
        const bool device = may_defer_bucketing && device_bucket_ && 2u * g0 >= num_points;
        vel_levels(in, device && sink_ && sink_cap_ >= num_points);
        if (device)
            vel_lite(num_points, out.vel);
        else
            flatten_points(vel_, groups_, num_points, out.vel);
    };
    const bool parallel = num_points >= kParallelPoints;
    const bool simd = in.use_simd;
    vel_after_ms = wait_ms = 0.0;
    if (simd && num_points && (defer == kDeferAlways || (defer == kDeferLarge && parallel))) {
        // The islands first (both sequences need the groups), then the
        // velocity sequence on the helper and the warm-start one here.
        islands(in, out);
        run_helper([this, vel_chain] {
            vel_chain(true);
            vel_done_ = clock::now();
        });
        ws_chain(false);
        ws_ready_ = clock::now();
        pending_ = true;
        return true;
    }
    if (parallel) run_helper([&ws_chain] { ws_chain(true); });
    islands(in, out);
    groups_ready_.store(true, std::memory_order_release);
    if (!parallel) ws_chain(false);
    if (simd) vel_chain(false);
    const auto tw = clock::now();
    if (parallel) wait_helper();
    wait_ms = std::chrono::duration<double, std::milli>(clock::now() - tw).count();
    if (!in.use_simd) out.vel = out.ws;  // the scalar path sweeps in constraint order, sensors included
    return false;
}

void ScheduleBuilder::finish() {
    if (!pending_) return;
    const auto tw = std::chrono::steady_clock::now();
    wait_helper();
    pending_ = false;
    wait_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tw).count();
    vel_after_ms = std::max(0.0, std::chrono::duration<double, std::milli>(vel_done_ - ws_ready_).count());
}

}  // namespace kz
