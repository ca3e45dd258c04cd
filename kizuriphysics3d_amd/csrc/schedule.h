// schedule.h - host-side solve schedule: islands, the reference's first-fit
// SIMD batch order, and the level decomposition that the device sweeps.
#pragma once

#include <stdint.h>

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
};

struct ScheduleOutput {
    LevelSchedule ws;   // warm start and bias passes: points in constraint order
    LevelSchedule vel;  // velocity passes: first-fit batch order (SIMD solver)
    uint32_t islands = 0;
    uint32_t groups = 0;
};

class ScheduleBuilder {
public:
    void build(const ScheduleInput& in, ScheduleOutput& out);

private:
    std::vector<uint32_t> dyn_index_, parent_, rank_, root_island_, island_of_con_, island_points_;
    std::vector<uint32_t> group_of_island_, last_;
    std::vector<uint32_t> pt_group_, pt_level_;
    std::vector<std::vector<uint32_t>> island_cons_;
};

}  // namespace kz
