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
//
// The warm-start/bias sequence (constraint order) and the velocity sequence
// (first-fit order) are independent once the islands are known; above a few
// thousand points the first runs on a helper thread beside the second.
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

// Bucket (group, level) -> flat arrays.
void ScheduleBuilder::flatten(Seq& s, uint32_t groups, uint32_t num_points, LevelSchedule& out) {
    s.depth.assign(groups, 0);
    for (uint32_t i = 0; i < s.n; ++i) s.depth[s.grp[i]] = std::max(s.depth[s.grp[i]], s.lvl[i]);
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
    out.off.assign(total + 1, 0);
    for (uint32_t i = 0; i < s.n; ++i) out.off[s.base[s.grp[i]] + s.lvl[i]] += 1;
    out.max_width = 0;
    for (uint32_t l = 0; l < total; ++l) {
        out.max_width = std::max(out.max_width, out.off[l + 1]);
        out.off[l + 1] += out.off[l];
    }
    out.idx.resize(s.n);
    s.fill.assign(out.off.begin(), out.off.end() - 1);
    out.pos.assign(num_points, kNone);
    for (uint32_t i = 0; i < s.n; ++i) {
        const uint32_t e = s.fill[s.base[s.grp[i]] + s.lvl[i] - 1]++;
        out.idx[e] = s.pts[i];
        out.pos[s.pts[i]] = e;
    }
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
    const uint8_t* dyn = in.dynamic;
    Seq& s = ws_;
    s.last.assign(in.num_bodies, 0);
    uint32_t n = 0;
    for (uint32_t c = 0; c < in.num_constraints; ++c) {
        const U4& ci = in.info[c];
        const bool da = dyn[ci.x] != 0, db = dyn[ci.y] != 0;
        if (!da && !db) continue;
        const uint32_t np = ci.w & 0xFFu;
        for (uint32_t k = 0; k < np; ++k) {
            const uint32_t l = 1 + std::max(da ? s.last[ci.x] : 0u, db ? s.last[ci.y] : 0u);
            if (da) s.last[ci.x] = l;
            if (db) s.last[ci.y] = l;
            s.pts[n] = ci.z + k;
            ws_con_[n] = c;
            s.lvl[n] = l;
            ++n;
        }
    }
    s.n = n;
}

// First-fit batches per island, then the levels of the batch sequence.
void ScheduleBuilder::vel_sequence(const ScheduleInput& in) {
    const uint8_t* dyn = in.dynamic;
    Seq& s = vel_;
    s.last.assign(in.num_bodies, 0);
    uint32_t n = 0;
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
    for (uint32_t is = 0; is < ni; ++is) {
        batches.clear();
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
        // First usable batch of a list for (a, b); drops dead entries.
        auto first_fit = [&](OpenList& l, uint32_t a, uint32_t b) -> uint32_t {
            while (l.head < l.ids.size() && batches[l.ids[l.head]].n >= 4) ++l.head;
            size_t dead = 0;
            uint32_t found = kNone;
            for (size_t k = l.head; k < l.ids.size(); ++k) {
                const FitBatch& B = batches[l.ids[k]];
                if (B.n >= 4) {
                    ++dead;
                    continue;
                }
                if (holds_either(B, a, b)) continue;
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
        for (uint32_t j = island_off_[is]; j < island_off_[is + 1]; ++j) {
            const U4& ci = in.info[island_cons_[j]];
            if ((ci.w >> 8) & 1u) continue;  // sensor
            const uint32_t a = ci.x, b = ci.y, np = ci.w & 0xFFu;
            OpenList* cand = &open;
            if (!dyn[a]) cand = &open_without(a);
            else if (!dyn[b]) cand = &open_without(b);
            for (uint32_t k = 0; k < np; ++k) {
                const uint32_t placed = first_fit(*cand, a, b);
                if (placed == kNone) {
                    FitBatch nb;
                    nb.body[0] = a;
                    nb.body[4] = b;
                    for (int q = 1; q < 4; ++q) nb.body[q] = nb.body[4 + q] = kNone;
                    nb.pt[0] = ci.z + k;
                    nb.n = 1;
                    const uint32_t id = uint32_t(batches.size());
                    batches.push_back(nb);
                    open.ids.push_back(id);
                    for (size_t q = 0; q < wo_used; ++q)
                        if (open_wo[q].first != a && open_wo[q].first != b) open_wo[q].second.ids.push_back(id);
                } else {
                    // Entries of open_wo[a] / open_wo[b] now holding a or b
                    // are skipped by the conflict test on later visits.
                    FitBatch& B = batches[placed];
                    B.body[B.n] = a;
                    B.body[4 + B.n] = b;
                    B.pt[B.n] = ci.z + k;
                    ++B.n;
                }
            }
        }
        const uint32_t g = group_of_island_[is];
        for (const FitBatch& B : batches)
            for (uint32_t q = 0; q < B.n; ++q) {
                const uint32_t a = B.body[q], b = B.body[4 + q];
                const bool da = dyn[a] != 0, db = dyn[b] != 0;
                const uint32_t l = 1 + std::max(da ? s.last[a] : 0u, db ? s.last[b] : 0u);
                if (da) s.last[a] = l;
                if (db) s.last[b] = l;
                s.pts[n] = B.pt[q];
                s.grp[n] = g;
                s.lvl[n] = l;
                ++n;
            }
    }
    s.n = n;
}

void ScheduleBuilder::build(const ScheduleInput& in, ScheduleOutput& out) {
    const uint32_t nc = in.num_constraints;
    uint32_t num_points = 0;
    for (uint32_t c = 0; c < nc; ++c) num_points += in.info[c].w & 0xFFu;
    for (Seq* s : {&ws_, &vel_}) {
        s->pts.resize(num_points);
        s->grp.resize(num_points);
        s->lvl.resize(num_points);
    }
    ws_con_.resize(num_points);
    // The warm-start sequence: levels (body types only), then, once the
    // islands are known, groups and buckets.
    groups_ready_.store(false, std::memory_order_relaxed);
    auto ws_chain = [this, &in, &out, num_points](bool wait) {
        ws_levels(in);
        if (wait)
            while (!groups_ready_.load(std::memory_order_acquire)) std::this_thread::yield();
        for (uint32_t i = 0; i < ws_.n; ++i) ws_.grp[i] = group_of_island_[island_of_con_[ws_con_[i]]];
        flatten(ws_, groups_, num_points, out.ws);
    };
    const bool parallel = num_points >= kParallelPoints;
    if (parallel) run_helper([&ws_chain] { ws_chain(true); });
    islands(in, out);
    groups_ready_.store(true, std::memory_order_release);
    if (!parallel) ws_chain(false);
    if (in.use_simd) {
        vel_sequence(in);
        flatten(vel_, groups_, num_points, out.vel);
    }
    if (parallel) wait_helper();
    if (!in.use_simd) out.vel = out.ws;  // the scalar path sweeps in constraint order, sensors included
}

}  // namespace kz
