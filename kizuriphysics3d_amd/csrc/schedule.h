// schedule.h - host-side solve schedule: islands, the reference's first-fit
// SIMD batch order, and the level decomposition that the device sweeps.
#pragma once

#include <stdint.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

#include "sweep_limits.h"

namespace kz {

struct U4 {
    uint32_t x, y, z, w;
};
struct U2 {
    uint32_t x, y;
};

// One level-scheduled sequence.  The sweep's work item is a unit: a run of
// at most kMaxUnit points of one body pair that follow each other in the
// sequence with no other update of either body between them, so one thread
// applies them in order.  idx holds the points unit by unit, units grouped
// by level; unit[u]..unit[u+1] are the positions in idx of unit u;
// off[l]..off[l+1] are the units of level l; grp[g] = [first level, end
// level) of group g; pos[p] = position of point p in idx (0xFFFFFFFF when
// absent).
struct LevelSchedule {
    std::vector<uint32_t> idx;
    std::vector<uint32_t> pos;
    std::vector<uint32_t> unit;
    std::vector<uint32_t> off;
    std::vector<U2> grp;
    uint32_t units = 0;
    uint32_t largest_unit = 0;
    uint32_t levels = 0;      // total levels over all groups
    uint32_t max_depth = 0;   // deepest group
    uint32_t max_width = 0;   // widest level, in units
    uint32_t g0_points = 0;   // points of group 0
    // Left for the device to bucket (ScheduleBuilder::set_device_bucketing):
    // idx, pos and unit are empty, off holds group 0's levels only, and
    // max_width is group 0's widest level.
    bool device_bucketed = false;
};

// The velocity sequence as the level walk leaves it, for bucketing on the
// device: per point in sequence order its point index, unit << 2 | index in
// the unit (units numbered in the order of their first points) and level
// (1-based, within its group); the runs of consecutive points of one island
// (end point, group); and each group's first global level.
struct VelSequence {
    const uint32_t* pts = nullptr;
    const uint32_t* uid = nullptr;
    const uint32_t* lvl = nullptr;
    uint32_t n = 0, units = 0;
    const U2* runs = nullptr;
    uint32_t num_runs = 0;
    const uint32_t* group_base = nullptr;  // [groups]
    uint32_t groups = 0, levels = 0;
    // Per point: for the first point of a unit of group 0, the unit's place
    // among its level's units in sequence order; 0xFFFFFFFF for other groups
    // (the device assigns those).  Group 0's units per level [g0_levels].
    const uint32_t* rank = nullptr;
    const uint32_t* g0_count = nullptr;
    uint32_t g0_levels = 0;
    // The four per-point arrays sit in the sink (ScheduleBuilder::
    // set_vel_sink) `stride` entries apart, from its start.
    bool in_sink = false;
    size_t stride = 0;
};

struct ScheduleInput {
    uint32_t num_bodies = 0;
    const uint8_t* dynamic = nullptr;  // per body: motion type == Dynamic
    uint32_t num_constraints = 0;
    const U4* info = nullptr;          // body A, body B, first point, num points | sensor << 8
    bool use_simd = true;
    // Most points one unit of the velocity / warm-start sequence may hold
    // (1: every unit is one point).  Without the SIMD order the velocity
    // passes sweep the warm-start sequence, which then takes max_unit_vel.
    uint32_t max_unit_vel = kDefaultUnitVel;
    uint32_t max_unit_ws = kDefaultUnitWs;
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
// 0xFFFFFFFF), the points, the lane count, and the lanes that join a unit
// of the previous batch (bit q: lane q joins; bits 8 + 2q: which lane).
struct FitBatch {
    uint32_t body[8];
    uint32_t pt[4];
    uint32_t n;
    uint32_t join;
};

class ScheduleBuilder {
public:
    ScheduleBuilder() = default;
    ScheduleBuilder(const ScheduleBuilder&) = delete;
    ScheduleBuilder& operator=(const ScheduleBuilder&) = delete;
    ~ScheduleBuilder();
    void build(const ScheduleInput& in, ScheduleOutput& out);
    // A build in two halves.  start() returns once the islands, the groups
    // and the warm-start sequence (out.ws) are complete.  With the SIMD
    // solver and `defer` = kDeferLarge (above the two-thread size) or
    // kDeferAlways, the velocity sequence is still being built on the helper
    // thread then; it reads the arrays `in` points to and writes out.vel
    // until finish() returns, so neither may be touched before.  start()
    // returns whether the velocity sequence is pending.
    static constexpr uint32_t kDeferNever = 0, kDeferLarge = 1, kDeferAlways = 2;
    bool start(const ScheduleInput& in, ScheduleOutput& out, uint32_t defer);
    void finish();
    bool pending() const { return pending_; }
    // With on, a deferred velocity sequence whose group 0 holds at least
    // half the points stops after the level walk (out.vel.device_bucketed):
    // the caller buckets it on the device from velocity_sequence().
    void set_device_bucketing(bool on) { device_bucket_ = on; }
    // Where the level walk of such a sequence writes its per-point arrays
    // (point, unit, level, rank; `cap` entries each, one after the other):
    // memory the caller uploads from, e.g. pinned, so the arrays reach the
    // device without a copy on another thread, whose reads would also cost
    // the next walk a transfer of every line back.  Null, or less room than
    // the points of a build: the builder's own arrays.  The walk writes
    // there until finish() returns.
    void set_vel_sink(uint32_t* base, size_t cap) {
        sink_ = base;
        sink_cap_ = cap;
    }
    VelSequence velocity_sequence() const;
    // Of the last build: the warm-start sequence's time, the time spent
    // waiting for the other thread's sequence, and (deferred builds) the
    // velocity sequence's time after start() returned.
    double ws_ms = 0.0, wait_ms = 0.0, vel_after_ms = 0.0;

private:
    // A sequence of points and their units, before bucketing.  The two
    // sequences are built on two threads: each Seq gets its own cache lines.
    struct alignas(64) Seq {
        // Per point in sequence order: point, unit << 2 | index in the unit,
        // level.
        std::vector<uint32_t> pts, uid, lvl;
        std::vector<uint32_t> ulvl, ugrp, usize;  // per unit: level, group, points (warm-start order)
        std::vector<U2> runs;                     // (end point, group) of consecutive points of one group
        std::vector<uint32_t> depth;  // per group: deepest level (filled by the sequence)
        std::vector<uint32_t> base, fill, slot;
        uint32_t n = 0, nu = 0, largest = 0;  // points, units, largest unit
        void reserve(uint32_t points);
    };
    void islands(const ScheduleInput& in, ScheduleOutput& out);
    uint32_t union_find(const ScheduleInput& in);
    void ws_levels(const ScheduleInput& in);
    void vel_fit(const ScheduleInput& in);
    void vel_levels(const ScheduleInput& in, bool to_sink = false);
    uint32_t level_bases(Seq& s, uint32_t groups, LevelSchedule& out);
    void prefix_levels(Seq& s, uint32_t total, LevelSchedule& out);
    void place_points(Seq& s, uint32_t num_points, LevelSchedule& out);
    void flatten(Seq& s, uint32_t groups, uint32_t num_points, LevelSchedule& out);
    void flatten_points(Seq& s, uint32_t groups, uint32_t num_points, LevelSchedule& out);
    void run_helper(std::function<void()> job);
    void wait_helper();

    std::vector<uint32_t> dyn_index_, parent_, rank_, root_island_, island_of_con_, island_points_;
    std::vector<uint32_t> island_off_, island_cons_, group_of_island_, order_;
    std::vector<uint32_t> ws_con_;  // per unit of the warm-start sequence: its constraint
    std::vector<uint32_t> ws_body_level_, vel_body_level_;
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
    bool pending_ = false;
    ScheduleInput in_{};
    std::chrono::steady_clock::time_point ws_ready_, vel_done_;
    std::vector<uint32_t> fit_island_end_;  // per island: end of its batches
    // Level walk of the velocity sequence: deepest level of each group, the
    // largest unit index, and group 0's units per level (index level).
    std::vector<uint32_t> vel_depth_, vel_g0_count_, vel_group_base_, vel_rank_;
    uint32_t vel_jmax_ = 0;
    bool device_bucket_ = false;
    uint32_t* sink_ = nullptr;
    size_t sink_cap_ = 0;
    bool sunk_ = false;  // the last walk wrote to the sink
    void vel_lite(uint32_t num_points, LevelSchedule& out);
};

}  // namespace kz
