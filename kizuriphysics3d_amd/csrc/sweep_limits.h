// sweep_limits.h - sizes shared by the host solve schedule and the device
// sweeps.
#pragma once

#include <stdint.h>

namespace kz {

// Most points one sweep unit holds (the points of one contact manifold),
// and the default caps of the single-world solve.  Each level costs a
// device-wide barrier and a load round trip (~2.2 us on the 32k pile) plus
// the arithmetic of its longest unit (~0.8 us per velocity point): units of
// 2 take 13% off the sweeps, units of 4 lose it again; warm-start units of 4
// shorten that chain further but need 512-thread workgroups, which the
// velocity sweeps then pay for.
constexpr uint32_t kMaxUnit = 4;
constexpr uint32_t kDefaultUnitVel = 2;
constexpr uint32_t kDefaultUnitWs = 2;
// The unit size a sweep is unrolled for when its largest unit holds
// `largest` points, and the instantiation of the solve kernels (warm-start /
// bias and velocity unrolling) that serves a schedule.
constexpr uint32_t unit_size(uint32_t largest) { return largest <= 1 ? 1u : largest <= 2 ? 2u : kMaxUnit; }
struct UnitKernel {
    uint32_t ws = 1, vel = 1;
};
constexpr UnitKernel unit_kernel(uint32_t largest_ws, uint32_t largest_vel) {
    return UnitKernel{unit_size(largest_ws), unit_size(largest_vel)};
}
// Most threads per workgroup of a solve kernel whose warm-start / bias units
// hold up to UW points and velocity units up to UV: a unit's records sit in
// registers, and four records of either kind need the register budget of
// 512 threads (2 waves per SIMD).
constexpr uint32_t solve_threads(uint32_t UW, uint32_t UV) { return UV <= 2 && UW <= 2 ? 1024u : 512u; }
// Words of a cooperative sweep's sync block: the launch's arrivals, the XCC
// chosen for a one-XCD sweep, each XCC's arrivals, the level counter and
// the give-up flag.
constexpr uint32_t kSyncWords = 32;
enum : uint32_t { kSyncArrive = 0, kSyncTarget = 1, kSyncXcc = 2, kSyncLevel = 18, kSyncGiveUp = 19 };
// XCDs of an MI355X (workgroups are dealt to them in turn); the one-XCD
// sweep launches this many workgroups per participant.
constexpr uint32_t kXcds = 8;
// Default bound on the polls of one wait of a cooperative sweep (a wait
// sleeps between polls: about a second in all).
constexpr uint32_t kDefaultSpinLimit = 1u << 21;

}  // namespace kz
