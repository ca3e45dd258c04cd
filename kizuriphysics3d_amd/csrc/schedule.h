// schedule.h - host-side solve schedule: islands, the reference's first-fit
// SIMD batch order, and the level decomposition that the device sweeps.
#pragma once

#include <stdint.h>

#include <atomic>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace kz {

struct U4 {
    uint32_t x, y, z, w;
};
struct U2 {
    uint32_t x, y;
};

// One level-scheduled sequence: idx holds point indices grouped by level,
// off[l]..off[l+1] is level l, grp[g] = [first level, end level) of group g,
// pos[p] = position of point p in idx (0xFFFFFFFF when absent).
struct LevelSchedule {
    std::vector<uint32_t> idx;
    std::vector<uint32_t> pos;
    std::vector<uint32_t> off;
    std::vector<U2> grp;
    uint32_t levels = 0;      // total levels over all groups
    uint32_t max_depth = 0;   // deepest group
    uint32_t max_width = 0;   // widest level
};

struct ScheduleInput {
    uint32_t num_bodies = 0;
    const uint8_t* dynamic = nullptr;  // per body: motion type == Dynamic
    uint32_t num_constraints = 0;
    const U4* info = nullptr;          // body A, body B, first point, num points | sensor << 8
    bool use_simd = true;
    uint32_t max_groups = 1024;
    // Optional: the island of each constraint, numbered as BuildIslands
    // numbers them, and the island count (every dynamic body is in one).
    // When given, the builder takes them instead of running union-find.
    const uint32_t* island_of_con = nullptr;
    uint32_t num_islands = 0;
};

struct ScheduleOutput {
    LevelSchedule ws;   // warm start and bias passes: points in constraint order
    LevelSchedule vel;  // velocity passes: first-fit batch order (SIMD solver)
    uint32_t islands = 0;
    uint32_t groups = 0;
};

// A first-fit batch: bodies A of lanes 0-3, then bodies B (unused lanes hold
// 0xFFFFFFFF), the points, the lane count.
struct FitBatch {
    uint32_t body[8];
    uint32_t pt[4];
    uint32_t n;
};

class ScheduleBuilder {
public:
    ScheduleBuilder() = default;
    ScheduleBuilder(const ScheduleBuilder&) = delete;
    ScheduleBuilder& operator=(const ScheduleBuilder&) = delete;
    ~ScheduleBuilder();
    void build(const ScheduleInput& in, ScheduleOutput& out);

private:
    // A sequence of points with their group and level, before bucketing.
    struct Seq {
        std::vector<uint32_t> pts, grp, lvl, last, depth, base, fill;
        uint32_t n = 0;
    };
    void islands(const ScheduleInput& in, ScheduleOutput& out);
    uint32_t union_find(const ScheduleInput& in);
    void ws_levels(const ScheduleInput& in);
    void vel_sequence(const ScheduleInput& in);
    void flatten(Seq& s, uint32_t groups, uint32_t num_points, LevelSchedule& out);
    void run_helper(std::function<void()> job);
    void wait_helper();

    std::vector<uint32_t> dyn_index_, parent_, rank_, root_island_, island_of_con_, island_points_;
    std::vector<uint32_t> island_off_, island_cons_, group_of_island_, order_;
    std::vector<uint32_t> ws_con_;  // per point of the warm-start sequence: its constraint
    Seq ws_, vel_;
    std::vector<FitBatch> batches_;
    uint32_t groups_ = 0;
    // One helper thread runs the warm-start sequence beside the velocity one.
    std::thread helper_;
    std::mutex m_;
    std::condition_variable cv_;
    std::function<void()> job_;
    bool stop_ = false, busy_ = false;
    std::atomic<bool> groups_ready_{false};
};

}  // namespace kz
