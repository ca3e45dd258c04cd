// context.hip - the kzg_* C-ABI: device buffers, uploads and the per-step
// orchestration of the contact path on one HIP stream.
#include <mutex>

#include "ctx_internal.h"

namespace {

// ---------------------------------------------------------------------------
// Scratch of the allocation-free standalone narrowphase
// (kzg_collide_batch_device): every device buffer the narrowphase needs
// beyond the caller's inputs and outputs, for at most kScratchPairs pairs,
// in one file-scope device array.  Each GPU holds its own copy for the life
// of the process; host code reaches it only through the address
// hipGetSymbolAddress returns on that device (kzg_create).
constexpr uint32_t kScratchPairs = KZG_NO_ALLOC_MAX_PAIRS;
constexpr size_t scratch_align(size_t x) { return (x + 255) & ~size_t(255); }
struct ScratchLayout {
    size_t x, q, meta, packed, pairs, work, counters, gjk, mcount, mpa, mpb, mn, total;
};
// counters: 3 work-list counters, the pair count, the error flag.
constexpr ScratchLayout make_scratch_layout(uint32_t n) {
    ScratchLayout L{};
    size_t o = 0;
    auto take = [&](size_t& field, size_t bytes) {
        field = o;
        o = scratch_align(o + bytes);
    };
    take(L.x, 2 * size_t(n) * sizeof(float4));
    take(L.q, 2 * size_t(n) * sizeof(float4));
    take(L.meta, 2 * size_t(n) * sizeof(uint32_t));
    take(L.packed, 2 * size_t(n) * sizeof(uint32_t));
    take(L.pairs, size_t(n) * sizeof(uint2));
    take(L.work, (3 * size_t(n) + size_t(kWorkClasses) * ((n + kWorkChunk - 1) / kWorkChunk)) * sizeof(uint32_t));
    take(L.counters, 8 * sizeof(uint32_t));
    take(L.gjk, size_t(n) * kGjkRecBytes);
    take(L.mcount, size_t(n) * sizeof(uint32_t));
    take(L.mpa, size_t(n) * kPairPoints * sizeof(float4));
    take(L.mpb, size_t(n) * kPairPoints * sizeof(float4));
    take(L.mn, size_t(n) * kPairPoints * sizeof(float4));
    L.total = o;
    return L;
}
constexpr ScratchLayout kScratch = make_scratch_layout(kScratchPairs);
static __device__ __align__(256) unsigned char scratch[kScratch.total];

// Per device: the device address of `scratch`, resolved once with that
// device current.
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_dev[kMaxScratchDevices] = {};

hipError_t resolve_scratch(int device, unsigned char** out) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (device < 0 || device >= kMaxScratchDevices) return hipErrorInvalidDevice;
    if (!g_scratch_dev[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_dev[device] = static_cast<unsigned char*>(p);
    }
    *out = g_scratch_dev[device];
    return hipSuccess;
}

__global__ void k_fill_zero(double* p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}
void launch_fill_zero(hipStream_t s, double* p, uint32_t n) {
    if (n) k_fill_zero<<<(n + 255) / 256, 256, 0, s>>>(p, n);
}

// Clears the work-list counters and the error flag and stores the pair count
// (a kernel, not a memset: a caller may capture the launches in a graph).
__global__ void k_scratch_reset(uint32_t* counters, uint32_t n) {
    if (threadIdx.x == 0) {
        counters[0] = counters[1] = counters[2] = 0u;
        counters[3] = n;
        counters[4] = 0u;
    }
}

// Pair i of the caller's arrays -> two scratch bodies (slots i and n + i)
// and the pair (i, n + i).  Both bodies are "dynamic" with packed ids 0 < 1,
// so the world-mode narrowphase collides A against B in the caller's order,
// as NarrowPhase::Collide(shapeA, xfA, shapeB, xfB) does.  A shape index
// outside the table raises the error flag and both bodies become static, so
// the narrowphase skips the pair (count 0) and reads no shape record.
__global__ void k_scratch_setup(const uint32_t* shape_a, const uint32_t* shape_b, const float* xf_a, const float* xf_b,
                                uint32_t n, uint32_t num_shapes, BodyBuf b, uint2* pairs, uint32_t* counters) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t sa = shape_a[i], sb = shape_b[i];
    const bool ok = sa < num_shapes && sb < num_shapes;
    if (!ok) atomicOr(counters + 4, 1u);
    const float* a = xf_a + 7 * size_t(i);
    const float* c = xf_b + 7 * size_t(i);
    b.q[i] = make_float4(a[0], a[1], a[2], a[3]);
    b.x[i] = make_float4(a[4], a[5], a[6], 0.0f);
    b.q[n + i] = make_float4(c[0], c[1], c[2], c[3]);
    b.x[n + i] = make_float4(c[4], c[5], c[6], 0.0f);
    const uint32_t motion = ok ? (2u << 24) : 0u;
    b.meta[i] = (ok ? sa : 0u) | motion;
    b.meta[n + i] = (ok ? sb : 0u) | motion;
    b.packed[i] = 0u;
    b.packed[n + i] = 1u;
    pairs[i] = make_uint2(i, n + i);
}

// Manifold slots -> the caller's counts[n] and points[n * 4] in the layout
// kzg_collide_batch returns (slots past a pair's count: zeros, normal +Y).
__global__ void k_scratch_output(ManifoldBuf m, uint32_t n, uint32_t* counts, kz_contact_point* points) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * uint32_t(kPairPoints)) return;
    const uint32_t i = t / kPairPoints, j = t % kPairPoints;
    const uint32_t k = m.count[i];
    if (j == 0) counts[i] = k;
    const bool live = j < k;
    const float4 pa = live ? m.pa[t] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const float4 pb = live ? m.pb[t] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const float4 nn = live ? m.n[t] : make_float4(0.0f, 1.0f, 0.0f, 0.0f);
    float* o = reinterpret_cast<float*>(points + t);
    o[0] = pa.x; o[1] = pa.y; o[2] = pa.z;
    o[3] = pb.x; o[4] = pb.y; o[5] = pb.z;
    o[6] = nn.x; o[7] = nn.y; o[8] = nn.z;
    o[9] = pa.w;
    o[10] = pb.w;  // feature id bits
}


// Face normal of a hull triangle in the hull's frame, as GetShapeFace
// computes it: NormalizedOr((b - a) x (c - a), +Y) (CollideShape.cpp:66-78).
float4 hull_face_normal(const float* v, uint32_t a, uint32_t b, uint32_t cc) {
    f3 A{v[3 * a], v[3 * a + 1], v[3 * a + 2]}, B{v[3 * b], v[3 * b + 1], v[3 * b + 2]},
        C{v[3 * cc], v[3 * cc + 1], v[3 * cc + 2]};
    f3 n = normalized_or(cross(sub(B, A), sub(C, A)), f3{0.0f, 1.0f, 0.0f});
    return make_float4(n.x, n.y, n.z, 0.0f);
}

constexpr uint32_t kMaxCoopBlocks = 64;
constexpr uint32_t kAutoCoopBlocks = 16;  // auto mode's ceiling (fastest measured on the 32k pile: 16 > 12 > 8 > 24)
constexpr uint32_t kCoopSlice = 48;       // auto mode: units per workgroup and level, on average
constexpr uint32_t kXcdCoopBlocks = 32;   // one-XCD placement: at most one workgroup per CU of the XCD
constexpr uint32_t kAutoXcdBlocks = 24;   // its auto ceiling (32k pile: 24 > 16 = 32 > 8, 3.29 / 3.43 / 4.18 ms)

struct CoopPlan {
    uint32_t blocks = 1, threads = 64;
    uint32_t ws_points = 0, vel_points = 0;  // group 0's
};

// Workgroups and threads for group 0 (the group holding the largest
// island).  `request` 0 sizes the split from the group's mean level width;
// `max_threads` is the kernel's workgroup limit; a velocity unit takes
// `vel_lanes` threads.  With `with_vel` false the velocity schedule is not
// read at all: the builder's helper thread may still be writing it.
CoopPlan plan_cooperative(const ScheduleOutput& s, uint32_t request, uint32_t max_threads, uint32_t vel_lanes,
                          bool one_xcd, bool with_vel = true) {
    CoopPlan p;
    if (request == 1 || s.groups == 0 || s.ws.grp.empty()) return p;
    uint32_t units = 0, levels = 0, widest = 0;
    const LevelSchedule* both[2] = {&s.ws, &s.vel};
    for (uint32_t k = 0; k < (with_vel ? 2u : 1u); ++k) {
        const LevelSchedule* ls = both[k];
        if (ls->grp.empty()) continue;
        const U2 g = ls->grp[0];
        const uint32_t lanes = ls == &s.vel ? vel_lanes : 1u;
        for (uint32_t l = g.x; l < g.y; ++l) widest = std::max(widest, lanes * (ls->off[l + 1] - ls->off[l]));
        units += ls->off[g.y] - ls->off[g.x];
        levels += g.y - g.x;
        (ls == &s.ws ? p.ws_points : p.vel_points) = ls->g0_points;
    }
    if (!levels) return CoopPlan{};
    p.blocks = request ? request
                       : std::min(one_xcd ? kAutoXcdBlocks : kAutoCoopBlocks, std::max(1u, units / levels / kCoopSlice));
    if (one_xcd) p.blocks = std::min(p.blocks, kXcdCoopBlocks);
    if (p.blocks <= 1) return CoopPlan{};
    const uint32_t slice = (widest + p.blocks - 1) / p.blocks;
    while (p.threads < slice && p.threads < max_threads) p.threads *= 2;
    return p;
}

// The cooperative sweep's launch shape and sync block: the block is
// cleared once per solve (the give-up flag with it) before the first
// cooperative launch.
int begin_cooperative(kzg_ctx* c) {
    KZG_HIP(c, c->coop_sync.ensure(kSyncWords));
    if (!c->coop_launched) launch_sync_reset(c->stream, c->coop_sync.p);
    c->coop_launched = true;
    return KZG_OK;
}
// A cooperative sweep that gave up waiting (a workgroup never arrived within
// the spin limit) left its flag in the step's pinned totals: report it from
// the next call that waits for the stream.
int check_sweep(kzg_ctx* c) {
    uint32_t* tot = c->pin_totals.as<uint32_t>();
    if (tot && tot[5]) {
        tot[5] = 0;
        return fail(c, KZG_ERR_DEVICE, "the cooperative sweep gave up waiting for a workgroup (spin limit); "
                                       "that step's velocities are not the solve's");
    }
    return KZG_OK;
}

// ---- Active list (src/Body/BodyManager.cpp:138-160) -------------------------
constexpr uint32_t kNoBody = 0xFFFFFFFFu;

// The listed bodies active in this order, every other body asleep with zero
// velocities; timers zero, island indices unset.
int set_active_list(kzg_ctx* c, const uint32_t* list, uint32_t n) {
    const uint32_t nb = c->nbodies;
    std::vector<uint8_t> active(nb, 0);
    c->h_act.assign(std::max<uint32_t>(nb, 1), kNoBody);
    c->h_act_pos.assign(nb, kNoBody);
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t x = list[i];
        if (x >= nb || active[x])
            return fail(c, KZG_ERR_ARGUMENT, "active list: entry " + std::to_string(i) + " is out of range or repeated");
        active[x] = 1;
        c->h_act[i] = x;
        c->h_act_pos[x] = i;
    }
    c->h_act_n = n;
    const size_t nb1 = std::max<uint32_t>(nb, 1);
    for (auto* d : {&c->d_act, &c->d_island, &c->d_wake, &c->sl_parent, &c->sl_min}) KZG_HIP(c, d->ensure(nb1));
    KZG_HIP(c, c->d_timer.ensure(nb1));
    KZG_HIP(c, c->sl_q.ensure(nb1));
    KZG_HIP(c, c->pin_sleep.ensure(nb1 + 64));
    const uint32_t head[4] = {n, 0u, 0u, 0u};
    int rc;
    if ((rc = upload(c, c->d_act, list, n)) || (rc = upload(c, c->d_sleep, head, 4)) ||
        (rc = upload(c, c->sl_q, active.data(), nb)))
        return rc;
    launch_set_activity(c->stream, c->bodies(), nb, c->sl_q.p, c->d_timer.p, c->d_island.p);
    KZG_HIP(c, hipGetLastError());
    KZG_HIP(c, hipStreamSynchronize(c->stream));
    return KZG_OK;
}

// After the build's first wait: the bodies its wake-up pass put on the
// device list go onto the host's copy, in the same order.
int take_woken(kzg_ctx* c) {
    const uint32_t* sl = c->pin_sleep.as<uint32_t>();
    if (!c->wake_launched) return KZG_OK;
    if (sl[3]) return fail(c, KZG_ERR_DEVICE, "the wake-up pass did not settle within its round limit");
    const uint32_t w = sl[2], n0 = c->h_act_n;
    if (sl[0] != n0 + w || n0 + w > c->nbodies) return fail(c, KZG_ERR_DEVICE, "active list out of step with the device");
    if (!w) return KZG_OK;
    KZG_HIP(c, c->stage2.ensure(sizeof(uint32_t) * w));
    KZG_HIP(c, hipMemcpyAsync(c->stage2.p, c->d_act.p + n0, sizeof(uint32_t) * w, hipMemcpyDeviceToHost, c->stream));
    KZG_HIP(c, hipStreamSynchronize(c->stream));
    const uint32_t* woken = c->stage2.as<uint32_t>();
    for (uint32_t k = 0; k < w; ++k) {
        const uint32_t x = woken[k];
        if (x >= c->nbodies || c->h_act_pos[x] != kNoBody) return fail(c, KZG_ERR_DEVICE, "bad woken body");
        c->h_act[n0 + k] = x;
        c->h_act_pos[x] = n0 + k;
    }
    c->h_act_n = n0 + w;
    return KZG_OK;
}

// The reference solves islands one after another (PhysicsWorld.cpp:563-566
// without worker threads); the sweeps run them side by side, which is the
// same while no dynamic body is in two islands.  A sleeping body reached
// through a stale index can be (a contact goes to its partner's island
// when the body's own index is out of range).
bool shares_bodies(kzg_ctx* c) {
    c->h_owner.assign(c->nbodies, kNoBody);
    for (uint32_t ci = 0; ci < c->ncons; ++ci) {
        const uint32_t is = c->h_con_island[ci];
        if (is == kNoBody) continue;
        for (const uint32_t x : {c->h_info[ci].x, c->h_info[ci].y}) {
            if (!c->h_dynamic[x]) continue;
            uint32_t& o = c->h_owner[x];
            if (o == kNoBody) o = is;
            else if (o != is) return true;
        }
    }
    return false;
}

// Generations: island I gets 1 + the largest generation of an earlier
// island it shares a dynamic body with (0 if none).  Islands of one
// generation share no body and keep, body by body, the island order, so
// solving the generations one after another, each as one solve over its
// islands, updates every body in the reference's order.  Each generation's
// schedule is built as for a whole step (its islands only) and the
// schedules go up side by side: records at disjoint positions, levels,
// groups and units at per-generation offsets.
int schedule_generations(kzg_ctx* c, ScheduleInput si, uint32_t islands) {
    const uint32_t nc = c->ncons, np = c->npoints;
    std::vector<uint32_t> gen(islands, 0), used(c->nbodies, 0), off(islands + 1, 0), cons(nc);
    for (uint32_t ci = 0; ci < nc; ++ci)
        if (c->h_con_island[ci] != kNoBody) ++off[c->h_con_island[ci] + 1];
    for (uint32_t i = 0; i < islands; ++i) off[i + 1] += off[i];
    std::vector<uint32_t> fill(off.begin(), off.end() - 1);
    for (uint32_t ci = 0; ci < nc; ++ci)
        if (c->h_con_island[ci] != kNoBody) cons[fill[c->h_con_island[ci]]++] = ci;
    uint32_t ng = 1;
    for (uint32_t is = 0; is < islands; ++is) {
        uint32_t g = 0;
        for (uint32_t k = off[is]; k < off[is + 1]; ++k)
            for (const uint32_t x : {c->h_info[cons[k]].x, c->h_info[cons[k]].y})
                if (c->h_dynamic[x]) g = std::max(g, used[x]);
        gen[is] = g;
        ng = std::max(ng, g + 1);
        for (uint32_t k = off[is]; k < off[is + 1]; ++k)
            for (const uint32_t x : {c->h_info[cons[k]].x, c->h_info[cons[k]].y})
                if (c->h_dynamic[x]) used[x] = g + 1;
    }
    std::vector<uint32_t> ws_off, ws_unit, vel_off, vel_unit, ws_pos(np, kNoBody), vel_pos(np, kNoBody);
    std::vector<U2> ws_grp, vel_grp;
    c->gens.assign(ng, kzg_ctx::GenPlan{});
    c->h_gen_mask.resize(nc);
    uint32_t ws_rec = 0, vel_rec = 0, ws_ud = 0, vel_ud = 0;
    const uint32_t vel_lanes = c->world.use_simd_solver ? 2u : 1u;
    for (uint32_t g = 0; g < ng; ++g) {
        for (uint32_t ci = 0; ci < nc; ++ci) {
            const uint32_t is = c->h_con_island[ci];
            c->h_gen_mask[ci] = (is != kNoBody && gen[is] == g) ? is : kNoBody;
        }
        si.island_of_con = c->h_gen_mask.data();
        ScheduleOutput out;
        c->sched_builder.build(si, out);
        kzg_ctx::GenPlan& G = c->gens[g];
        G.ws_loff = uint32_t(ws_off.size());
        G.ws_goff = uint32_t(ws_grp.size());
        G.ws_uoff = uint32_t(ws_unit.size());
        G.ws_udoff = ws_ud;
        G.ws_units = out.ws.units;
        G.vel_loff = uint32_t(vel_off.size());
        G.vel_goff = uint32_t(vel_grp.size());
        G.vel_uoff = uint32_t(vel_unit.size());
        G.vel_udoff = vel_ud;
        G.vel_units = out.vel.units;
        G.groups = out.groups;
        G.width = std::max(out.ws.max_width, vel_lanes * out.vel.max_width);
        G.largest_ws = out.ws.largest_unit;
        G.largest_vel = out.vel.largest_unit;
        ws_off.insert(ws_off.end(), out.ws.off.begin(), out.ws.off.end());
        ws_grp.insert(ws_grp.end(), out.ws.grp.begin(), out.ws.grp.end());
        for (uint32_t u : out.ws.unit) ws_unit.push_back(u + ws_rec);
        vel_off.insert(vel_off.end(), out.vel.off.begin(), out.vel.off.end());
        vel_grp.insert(vel_grp.end(), out.vel.grp.begin(), out.vel.grp.end());
        for (uint32_t u : out.vel.unit) vel_unit.push_back(u + vel_rec);
        for (uint32_t p = 0; p < np; ++p) {
            if (out.ws.pos[p] != kNoBody) ws_pos[p] = out.ws.pos[p] + ws_rec;
            if (out.vel.pos[p] != kNoBody) vel_pos[p] = out.vel.pos[p] + vel_rec;
        }
        ws_rec += uint32_t(out.ws.idx.size());
        vel_rec += uint32_t(out.vel.idx.size());
        ws_ud += out.ws.units;
        vel_ud += out.vel.units;
        if (g == 0) c->sched = std::move(out);  // the step's reported schedule
    }
    KZG_HIP(c, c->ws_udesc.ensure(std::max<uint32_t>(ws_ud, 1)));
    KZG_HIP(c, c->vel_udesc.ensure(std::max<uint32_t>(vel_ud, 1)));
    const size_t bytes = sizeof(uint32_t) * (ws_off.size() + vel_off.size() + ws_unit.size() + vel_unit.size() +
                                             2 * size_t(np)) +
                         sizeof(U2) * (ws_grp.size() + vel_grp.size()) + 10 * 256;
    KZG_HIP(c, c->stage_b.ensure(bytes));
    size_t o = 0;
    int rc;
    if ((rc = upload_async(c, c->ws_off, c->stage_b, o, ws_off.data(), ws_off.size())) ||
        (rc = upload_async(c, c->ws_grp, c->stage_b, o, reinterpret_cast<const uint2*>(ws_grp.data()), ws_grp.size())) ||
        (rc = upload_async(c, c->ws_pos, c->stage_b, o, ws_pos.data(), ws_pos.size())) ||
        (rc = upload_async(c, c->vel_off, c->stage_b, o, vel_off.data(), vel_off.size())) ||
        (rc = upload_async(c, c->vel_grp, c->stage_b, o, reinterpret_cast<const uint2*>(vel_grp.data()), vel_grp.size())) ||
        (rc = upload_async(c, c->vel_pos, c->stage_b, o, vel_pos.data(), vel_pos.size())) ||
        (rc = upload_async(c, c->ws_unit, c->stage_b, o, ws_unit.data(), ws_unit.size())) ||
        (rc = upload_async(c, c->vel_unit, c->stage_b, o, vel_unit.data(), vel_unit.size())))
        return rc;
    return KZG_OK;
}

// The solve of a step scheduled in generations: one group-kernel solve per
// generation (warm start, velocity and bias sweeps), in order.
int solve_generations(kzg_ctx* c, const SolveParams& sp) {
    KZG_HIP(c, c->b_packed_solver.ensure(8 * size_t(std::max<uint32_t>(c->nbodies, 1))));
    for (const kzg_ctx::GenPlan& G : c->gens) {
        if (!G.ws_units && !G.vel_units) continue;
        launch_unit_desc(c->stream, c->ws_unit.p + G.ws_uoff, c->ws_records().bd, G.ws_units,
                         c->ws_udesc.p + G.ws_udoff);
        launch_unit_desc(c->stream, c->vel_unit.p + G.vel_uoff, c->vel_records().bd, G.vel_units,
                         c->vel_udesc.p + G.vel_udoff);
        const Schedule ws{c->ws_off.p + G.ws_loff, c->ws_grp.p + G.ws_goff, c->ws_udesc.p + G.ws_udoff};
        const Schedule vel{c->vel_off.p + G.vel_loff, c->vel_grp.p + G.vel_goff, c->vel_udesc.p + G.vel_udoff};
        const UnitKernel uk = unit_kernel(G.largest_ws, G.largest_vel);
        uint32_t threads = 64;
        while (threads < G.width && threads < solve_threads(uk.ws, uk.vel)) threads *= 2;
        launch_solve_groups(c->stream, c->vel_records(), c->ws_records(), c->bodies(), c->nbodies,
                            c->b_packed_solver.p, ws, vel, sp, G.groups, threads, c->resid.p, uk,
                            kPackBodies | kUnpackBodies);
        KZG_HIP(c, hipGetLastError());
    }
    return KZG_OK;
}

}  // namespace

extern "C" {

int kzg_update_sleeping(kzg_ctx* c, float dt) {
    if (!c) return KZG_ERR_ARGUMENT;
    if (!c->world.allow_sleeping) return KZG_OK;  // PhysicsWorld.cpp:867
    if (c->phase != 0) return fail(c, KZG_ERR_STATE, "kzg_update_sleeping follows kzg_integrate_positions");
    hipSetDevice(c->device);
    const uint32_t n = c->h_act_n;
    if (!n) return KZG_OK;
    launch_sleep_decide(c->stream, c->contacts(), c->ncons, c->compact_tot.p, c->bodies(), c->d_act.p, n,
                        c->d_timer.p, c->d_island.p, c->sl_parent.p, c->sl_min.p, c->sl_q.p, dt,
                        c->sleep_lin * c->sleep_lin, c->sleep_ang * c->sleep_ang, c->sleep_time);
    KZG_HIP(c, hipGetLastError());
    KZG_HIP(c, hipMemcpyAsync(c->pin_sleep.p, c->sl_q.p, n, hipMemcpyDeviceToHost, c->stream));
    KZG_HIP(c, hipStreamSynchronize(c->stream));
    const uint8_t* q = c->pin_sleep.as<uint8_t>();
    // The reference's loop (:932-941) with DeactivateBody's swap-remove;
    // position i holds, when the loop reaches it, the body that was there
    // when the loop started.
    c->h_orig.assign(c->h_act.begin(), c->h_act.begin() + n);
    uint32_t size = n;
    bool any = false;
    for (uint32_t i = 0; i < n; ++i) {
        if (!q[i]) continue;
        const uint32_t x = c->h_orig[i], p = c->h_act_pos[x];
        if (p == kNoBody || p >= size) return fail(c, KZG_ERR_DEVICE, "active list out of step with the device");
        const uint32_t last = c->h_act[size - 1];
        c->h_act[p] = last;
        c->h_act_pos[last] = p;
        c->h_act_pos[x] = kNoBody;
        --size;
        any = true;
    }
    if (!any) return KZG_OK;
    c->h_act_n = size;
    int rc;
    if ((rc = upload(c, c->d_act, c->h_act.data(), size)) || (rc = upload(c, c->d_sleep, &size, 1))) return rc;
    return KZG_OK;
}

int kzg_get_active_bodies(kzg_ctx* c, uint32_t* out, uint32_t cap, uint32_t* count) {
    if (!c || (cap && !out)) return KZG_ERR_ARGUMENT;
    join_schedule(c);
    const uint32_t n = c->h_act_n;
    if (count) *count = n;
    if (cap) std::memcpy(out, c->h_act.data(), sizeof(uint32_t) * std::min(n, cap));
    return KZG_OK;
}

int kzg_set_active_bodies(kzg_ctx* c, const uint32_t* list, uint32_t n) {
    if (!c || (n && !list)) return KZG_ERR_ARGUMENT;
    if (n > c->nbodies) return fail(c, KZG_ERR_ARGUMENT, "active list longer than the body table");
    join_schedule(c);
    hipSetDevice(c->device);
    int rc = set_active_list(c, list, n);
    if (rc) return rc;
    c->sleep_model = true;  // the build's wake-ups and islands follow the list from now on
    return KZG_OK;
}

int kzg_set_sleep_thresholds(kzg_ctx* c, float time, float linear, float angular) {
    if (!c) return KZG_ERR_ARGUMENT;
    if (!(time >= 0.0f) || !(linear >= 0.0f) || !(angular >= 0.0f) || !std::isfinite(time) ||
        !std::isfinite(linear) || !std::isfinite(angular))
        return fail(c, KZG_ERR_ARGUMENT, "sleep thresholds must be finite and non-negative");
    c->sleep_time = time;
    c->sleep_lin = linear;
    c->sleep_ang = angular;
    return KZG_OK;
}

int kzg_create(kzg_ctx** out, const kz_world_desc* world, int device) {
    if (!out || !world) return KZG_ERR_ARGUMENT;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= device || device < 0) return KZG_ERR_DEVICE;
    kzg_ctx* c = new kzg_ctx();
    c->device = device;
    c->world = *world;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return KZG_ERR_DEVICE;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->cu_count = uint32_t(prop.multiProcessorCount);
    for (auto& e : c->ev) hipEventCreate(&e);
    if (resolve_scratch(device, &c->scratch) != hipSuccess || c->pin_flag.ensure(4096) != hipSuccess ||
        c->pin_totals.ensure(8 * sizeof(uint32_t)) != hipSuccess || c->work_count.ensure(3) != hipSuccess ||
        ensure_cache(c, 512) != KZG_OK) {
        kzg_destroy(c);
        return KZG_ERR_DEVICE;
    }
    std::memset(c->pin_totals.p, 0, 8 * sizeof(uint32_t));  // slot 5: a sweep's give-up flag
    *out = c;
    return KZG_OK;
}

void kzg_destroy(kzg_ctx* c) {
    if (!c) return;
    join_schedule(c);
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    for (auto* d : {&c->b_x, &c->b_q, &c->b_v, &c->b_w, &c->b_pv, &c->b_pw, &c->b_iw, &c->b_il, &c->b_mat, &c->b_lim,
                    &c->s_params, &c->s_hverts, &c->s_hfn, &c->m_pa, &c->m_pb, &c->m_n, &c->p_pa, &c->p_pb, &c->p_n,
                    &c->p_t1, &c->p_t2, &c->p_ra, &c->p_rb, &c->p_la, &c->p_lb, &c->p_imp, &c->cache_p, &c->x_x,
                    &c->x_q, &c->vr_ra, &c->vr_rb, &c->vr_n, &c->vr_t1, &c->vr_t2, &c->vr_imp, &c->wr_ra, &c->wr_rb,
                    &c->wr_qa, &c->wr_qb, &c->wr_n, &c->wr_t1, &c->wr_t2, &c->wr_imp})
        d->release();
    for (auto* d : {&c->b_meta, &c->b_packed, &c->s_type, &c->s_htris, &c->m_count, &c->work, &c->work_count,
                    &c->c_pair, &c->ws_off, &c->vel_off, &c->ws_pos, &c->vel_pos, &c->ws_unit, &c->vel_unit,
                    &c->cache_count, &c->cache_used, &c->x_meta, &c->x_packed})
        d->release();
    c->s_hull.release();
    c->gjk.release();
    c->c_info.release();
    c->p_body.release();
    c->vr_bd.release();
    c->wr_bd.release();
    c->vr_fb.release();
    c->p_nvel.release();
    c->ws_grp.release();
    c->vel_grp.release();
    c->ws_udesc.release();
    c->vel_udesc.release();
    c->d_pairs.release();
    c->x_pairs.release();
    c->cache_key.release();
    c->cache_t.release();
    c->resid.release();
    c->natural.release();
    c->b_packed_solver.release();
    for (auto* d : {&c->mlp_master, &c->mlp_m, &c->mlp_v, &c->mlp_norm, &c->mlp_partial, &c->mlp_out, &c->mlp_loss})
        d->release();
    for (auto* d : {&c->bp_layer, &c->bp_mask, &c->bp_counts, &c->bp_sums, &c->bp_counters}) d->release();
    c->bp_nodes.release();
    c->bp_items_a.release();
    c->bp_items_b.release();
    c->pin_bp.release();
    c->stage_bp.release();
    c->mlp_w1b.release();
    c->mlp_w2b.release();
    c->mlp_w3b.release();
    c->feat.release();
    c->base.release();
    c->samples.release();
    c->x_samples.release();
    c->x_out.release();
    c->stage.release();
    c->stage2.release();
    c->stage_a.release();
    c->stage_b.release();
    c->stage_v.release();
    c->stage_seq.release();
    c->pin_totals.release();
    c->pin_info.release();
    c->pin_flag.release();
    c->coop_sync.release();
    c->body_load.release();
    for (auto* d : {&c->vb_pts, &c->vb_uid, &c->vb_lvl, &c->vb_rank, &c->vb_seq, &c->vb_gl, &c->vb_base, &c->vb_count,
                    &c->vb_fill, &c->vb_uslot, &c->vb_usize, &c->vb_sums})
        d->release();
    c->vb_runs.release();
    c->compact_tile.release();
    c->compact_tot.release();
    c->uf_parent.release();
    c->uf_island.release();
    c->con_island.release();
    for (auto* d : {&c->d_act, &c->d_island, &c->d_wake, &c->d_sleep, &c->sl_parent, &c->sl_min}) d->release();
    c->d_timer.release();
    c->sl_q.release();
    c->pin_sleep.release();
    for (auto& e : c->ev)
        if (e) hipEventDestroy(e);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

const char* kzg_last_error(const kzg_ctx* c) { return c ? c->err.c_str() : "null context"; }

int kzg_upload_shapes(kzg_ctx* c, const kz_shape_desc* shapes, uint32_t n, const kzg_hull_range* hulls,
                      const float* hv, uint32_t nv, const uint32_t* ht, uint32_t nt) {
    if (!c || (n && !shapes) || (n && !hulls)) return fail(c, KZG_ERR_ARGUMENT, "null shape arrays");
    hipSetDevice(c->device);
    std::vector<uint32_t> type(n);
    std::vector<float4> prm(n);
    std::vector<uint4> hr(n);
    std::vector<float4> verts(nv);
    std::vector<uint32_t> tris(nt);
    std::vector<float4> fn(nt);
    for (uint32_t i = 0; i < nv; ++i) verts[i] = f4(hv[3 * i], hv[3 * i + 1], hv[3 * i + 2], 0.0f);
    for (uint32_t i = 0; i < n; ++i) {
        const kz_shape_desc& s = shapes[i];
        if (s.type != kSphere && s.type != kBox && s.type != kCapsule && s.type != kHull && s.type != kPlane)
            return fail(c, KZG_ERR_UNSUPPORTED, "shape " + std::to_string(i) + ": type " + std::to_string(s.type) +
                                                    " is outside the device path (sphere, box, capsule, hull, plane)");
        type[i] = s.type;
        prm[i] = f4(s.params[0], s.params[1], s.params[2], s.params[3]);
        hr[i] = make_uint4(0, 0, 0, 0);
        if (s.type == kHull) {
            const kzg_hull_range& h = hulls[i];
            if (h.vert_count < 4 || h.vert_count > uint32_t(kMaxHullVerts) || h.tri_count < 4 ||
                h.tri_count > uint32_t(kMaxHullTris) || size_t(h.vert_offset) + h.vert_count > nv ||
                size_t(h.tri_offset) + h.tri_count > nt)
                return fail(c, KZG_ERR_ARGUMENT, "hull " + std::to_string(i) + ": bad vertex/triangle range");
            hr[i] = make_uint4(h.vert_offset, h.vert_count, h.tri_offset, h.tri_count);
            for (uint32_t t = 0; t < h.tri_count; ++t) {
                const uint32_t* tri = ht + 3 * size_t(h.tri_offset + t);
                if (tri[0] >= h.vert_count || tri[1] >= h.vert_count || tri[2] >= h.vert_count)
                    return fail(c, KZG_ERR_ARGUMENT, "hull " + std::to_string(i) + ": triangle index out of range");
                tris[h.tri_offset + t] = tri[0] | (tri[1] << 8) | (tri[2] << 16);
                fn[h.tri_offset + t] = hull_face_normal(hv + 3 * size_t(h.vert_offset), tri[0], tri[1], tri[2]);
            }
        }
    }
    int rc;
    if ((rc = upload(c, c->s_type, type.data(), n)) || (rc = upload(c, c->s_params, prm.data(), n)) ||
        (rc = upload(c, c->s_hull, hr.data(), n)) || (rc = upload(c, c->s_hverts, verts.data(), nv)) ||
        (rc = upload(c, c->s_htris, tris.data(), nt)) || (rc = upload(c, c->s_hfn, fn.data(), nt)))
        return rc;
    c->nshapes = n;
    c->h_shape_type = type;
    return KZG_OK;
}

int kzg_upload_bodies(kzg_ctx* c, const kzg_body* b, uint32_t n) {
    if (!c || (n && !b)) return fail(c, KZG_ERR_ARGUMENT, "null body array");
    if (n >= (1u << 24)) return fail(c, KZG_ERR_ARGUMENT, "too many bodies");
    join_schedule(c);
    hipSetDevice(c->device);
    std::vector<float4> x(n), q(n), v(n), w(n), z(n, f4(0, 0, 0, 0)), il(3 * size_t(n)), mat(n), lim(n);
    std::vector<uint32_t> meta(n), packed(n);
    c->h_packed.resize(n);
    c->h_dynamic.resize(n);
    c->h_sensor.resize(n);
    c->h_motion.resize(n);
    for (uint32_t i = 0; i < n; ++i) {
        const kzg_body& d = b[i];
        if (d.shape >= c->nshapes) return fail(c, KZG_ERR_ARGUMENT, "body " + std::to_string(i) + ": bad shape index");
        if (d.motion_type > 2) return fail(c, KZG_ERR_ARGUMENT, "body " + std::to_string(i) + ": bad motion type");
        x[i] = f4(d.position[0], d.position[1], d.position[2], d.inverse_mass);
        q[i] = f4(d.rotation[0], d.rotation[1], d.rotation[2], d.rotation[3]);
        v[i] = f4(d.linear_velocity[0], d.linear_velocity[1], d.linear_velocity[2], 0.0f);
        w[i] = f4(d.angular_velocity[0], d.angular_velocity[1], d.angular_velocity[2], 0.0f);
        for (int col = 0; col < 3; ++col)
            il[3 * size_t(i) + col] = f4(d.inverse_inertia_local[3 * col], d.inverse_inertia_local[3 * col + 1],
                                         d.inverse_inertia_local[3 * col + 2], 0.0f);
        mat[i] = f4(d.friction, d.restitution, d.linear_damping, d.angular_damping);
        lim[i] = f4(d.gravity_factor, d.max_linear_velocity, d.max_angular_velocity, 0.0f);
        meta[i] = d.shape | (d.motion_type << 24) | ((d.is_sensor ? 1u : 0u) << 26) |
                  (d.allow_sleeping ? kMetaMaySleep : 0u);
        packed[i] = d.packed_id;
        c->h_packed[i] = d.packed_id;
        c->h_dynamic[i] = d.motion_type == 2;
        c->h_sensor[i] = d.is_sensor != 0;
        c->h_motion[i] = uint8_t(d.motion_type);
    }
    int rc;
    if ((rc = upload(c, c->b_x, x.data(), n)) || (rc = upload(c, c->b_q, q.data(), n)) ||
        (rc = upload(c, c->b_v, v.data(), n)) || (rc = upload(c, c->b_w, w.data(), n)) ||
        (rc = upload(c, c->b_pv, z.data(), n)) || (rc = upload(c, c->b_pw, z.data(), n)) ||
        (rc = upload(c, c->b_il, il.data(), 3 * size_t(n))) || (rc = upload(c, c->b_mat, mat.data(), n)) ||
        (rc = upload(c, c->b_lim, lim.data(), n)) || (rc = upload(c, c->b_meta, meta.data(), n)) ||
        (rc = upload(c, c->b_packed, packed.data(), n)))
        return rc;
    KZG_HIP(c, c->b_iw.ensure(3 * size_t(std::max<uint32_t>(n, 1))));
    c->nbodies = n;
    launch_world_inertia(c->stream, c->bodies(), n);
    KZG_HIP(c, hipGetLastError());
    KZG_HIP(c, hipStreamSynchronize(c->stream));
    c->phase = 0;
    c->ncons = 0;
    // Every body on the active list in slot order (CreateBody,
    // src/Body/BodyManager.cpp:59); island indices unset (Body.h:281).
    std::vector<uint32_t> all(n);
    for (uint32_t i = 0; i < n; ++i) all[i] = i;
    if ((rc = set_active_list(c, all.data(), n))) return rc;
    c->sleep_model = c->world.allow_sleeping != 0;
    return KZG_OK;
}

int kzg_upload_states(kzg_ctx* c, const kz_body_state* s, uint32_t n) {
    if (!c || !s || n != c->nbodies) return fail(c, KZG_ERR_ARGUMENT, "state count does not match the body table");
    hipSetDevice(c->device);
    std::vector<float4> x(n), q(n), v(n), w(n);
    std::vector<float4> cur(n);
    int rc = download(c, cur.data(), c->b_x, n);
    if (rc) return rc;
    for (uint32_t i = 0; i < n; ++i) {
        x[i] = f4(s[i].position[0], s[i].position[1], s[i].position[2], cur[i].w);
        q[i] = f4(s[i].rotation[0], s[i].rotation[1], s[i].rotation[2], s[i].rotation[3]);
        v[i] = f4(s[i].linear_velocity[0], s[i].linear_velocity[1], s[i].linear_velocity[2], 0.0f);
        w[i] = f4(s[i].angular_velocity[0], s[i].angular_velocity[1], s[i].angular_velocity[2], 0.0f);
    }
    if ((rc = upload(c, c->b_x, x.data(), n)) || (rc = upload(c, c->b_q, q.data(), n)) ||
        (rc = upload(c, c->b_v, v.data(), n)) || (rc = upload(c, c->b_w, w.data(), n)))
        return rc;
    launch_world_inertia(c->stream, c->bodies(), n);
    KZG_HIP(c, hipGetLastError());
    KZG_HIP(c, hipStreamSynchronize(c->stream));
    return KZG_OK;
}

namespace {
// Step 2 over the n pairs in d_pairs (already on the device or being copied
// there on the stream).
int narrowphase_on_device(kzg_ctx* c, uint32_t n);
}  // namespace

int kzg_narrowphase(kzg_ctx* c, const uint32_t* pairs, uint32_t n) {
    if (!c || (n && !pairs)) return fail(c, KZG_ERR_ARGUMENT, "null pair array");
    if (n > kWorkPairMask) return fail(c, KZG_ERR_UNSUPPORTED, "more than 2^27 - 1 broadphase pairs");
    join_schedule(c);
    hipSetDevice(c->device);
    for (uint32_t i = 0; i < 2 * n; ++i)
        if (pairs[i] >= c->nbodies) return fail(c, KZG_ERR_ARGUMENT, "pair " + std::to_string(i / 2) + ": bad body index");
    c->recorded = 0;
    c->pairs_resident = false;
    int rc;
    if ((rc = upload(c, c->d_pairs, reinterpret_cast<const uint2*>(pairs), n))) return rc;
    return narrowphase_on_device(c, n);
}

int kzg_narrowphase_resident(kzg_ctx* c) {
    if (!c) return KZG_ERR_ARGUMENT;
    if (!c->pairs_resident) return fail(c, KZG_ERR_STATE, "kzg_narrowphase_resident needs a kzg_broadphase_pairs");
    if (c->npairs > kWorkPairMask) return fail(c, KZG_ERR_UNSUPPORTED, "more than 2^27 - 1 broadphase pairs");
    join_schedule(c);
    hipSetDevice(c->device);
    c->recorded = 0;
    return narrowphase_on_device(c, c->npairs);
}

}  // extern "C"

namespace {
int narrowphase_on_device(kzg_ctx* c, uint32_t n) {
    c->npairs = n;
    KZG_HIP(c, c->m_count.ensure(std::max<uint32_t>(n, 1)));
    KZG_HIP(c, c->m_pa.ensure(size_t(std::max<uint32_t>(n, 1)) * kPairPoints));
    KZG_HIP(c, c->m_pb.ensure(size_t(std::max<uint32_t>(n, 1)) * kPairPoints));
    KZG_HIP(c, c->m_n.ensure(size_t(std::max<uint32_t>(n, 1)) * kPairPoints));
    KZG_HIP(c, c->work.ensure(work_entries(std::max<uint32_t>(n, 1))));
    KZG_HIP(c, c->gjk.ensure(size_t(std::max<uint32_t>(n, 1)) * kGjkRecBytes));
    KZG_HIP(c, c->c_info.ensure(std::max<uint32_t>(n, 1)));
    KZG_HIP(c, c->c_pair.ensure(std::max<uint32_t>(n, 1)));
    KZG_HIP(c, c->compact_tile.ensure((std::max(std::max<uint32_t>(n, 1), c->nbodies) + 1023) / 1024));
    KZG_HIP(c, c->compact_tot.ensure(4));
    KZG_HIP(c, c->pin_totals.ensure(8 * sizeof(uint32_t)));
    KZG_HIP(c, c->uf_parent.ensure(std::max<uint32_t>(c->nbodies, 1)));
    KZG_HIP(c, c->uf_island.ensure(std::max<uint32_t>(c->nbodies, 1)));
    KZG_HIP(c, c->con_island.ensure(std::max<uint32_t>(n, 1)));
    KZG_HIP(c, record(c, 0));
    launch_narrowphase(c->stream, c->d_pairs.p, n, c->bodies(), c->shapes(), c->manifolds(), c->work.p,
                       c->work_count.p, c->gjk.p, 0.02f, 1,
                       c->np_waves ? c->np_waves : std::min<uint32_t>(std::max<uint32_t>(n, 1), c->cu_count * 16u));
    KZG_HIP(c, hipGetLastError());
    KZG_HIP(c, record(c, 1));
    // The constraint table (BuildContactConstraints' compaction, :454-471)
    // and its islands (BuildIslands, :587-664) are built on the device
    // right behind the narrowphase.
    launch_compact(c->stream, c->m_count.p, n, c->d_pairs.p, c->bodies(), c->contacts(), c->compact_tile.p,
                   c->compact_tot.p);
    if (c->sleep_model) {
        // The build's wake-ups (only while some body sleeps), then the
        // islands over the active list's compact indices.
        c->wake_launched = c->h_act_n < c->nbodies;
        if (c->wake_launched)
            launch_wake(c->stream, c->contacts(), c->compact_tot.p, c->bodies(), c->nbodies, c->d_wake.p,
                        c->d_act.p, c->d_sleep.p, c->d_timer.p, c->nbodies + 2u);
        launch_island_index(c->stream, c->d_act.p, c->d_sleep.p, c->bodies(), c->d_island.p);
        launch_islands_indexed(c->stream, c->contacts(), n, c->bodies(), c->nbodies, c->d_island.p, c->d_sleep.p,
                               c->uf_parent.p, c->uf_island.p, c->compact_tile.p, c->compact_tot.p,
                               c->con_island.p);
        KZG_HIP(c, hipMemcpyAsync(c->pin_sleep.p, c->d_sleep.p, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost,
                                  c->stream));
    } else {
        launch_islands(c->stream, c->contacts(), n, c->bodies(), c->nbodies, c->uf_parent.p, c->uf_island.p,
                       c->compact_tile.p, c->compact_tot.p, c->con_island.p, c->island_lds_limit);
    }
    KZG_HIP(c, hipGetLastError());
    KZG_HIP(c, record(c, 13));
    uint32_t* tot = c->pin_totals.as<uint32_t>();
    KZG_HIP(c, hipMemcpyAsync(tot, c->compact_tot.p, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    KZG_HIP(c, hipMemcpyAsync(tot + 4, c->cache_used.p, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    c->phase = 1;
    return KZG_OK;
}
}  // namespace

extern "C" {

int kzg_build_constraints(kzg_ctx* c, int warm) {
    if (!c) return KZG_ERR_ARGUMENT;
    if (c->phase != 1) return fail(c, KZG_ERR_STATE, "kzg_build_constraints needs a preceding kzg_narrowphase");
    const int mode = warm & 0xFF;
    if ((warm & ~(0xFF | KZG_WARM_NO_STORE | KZG_WARM_NO_COLLECT)) || mode > KZG_WARM_LEARNED_MISS)
        return fail(c, KZG_ERR_ARGUMENT, "unknown warm-start mode " + std::to_string(warm));
    if ((mode == KZG_WARM_LEARNED || mode == KZG_WARM_LEARNED_MISS) && !c->mlp_ready)
        return fail(c, KZG_ERR_STATE, "learned warm start needs network weights (kzg_mlp_init or kzg_mlp_set_params)");
    c->cur_mode = mode;
    c->no_store = (warm & KZG_WARM_NO_STORE) != 0;
    c->no_collect = (warm & KZG_WARM_NO_COLLECT) != 0;
    join_schedule(c);
    hipSetDevice(c->device);
    // Constraint and point totals and the cache occupancy (the wait is the
    // narrowphase and the compaction still running).
    KZG_HIP(c, hipStreamSynchronize(c->stream));
    int rc0;
    if ((rc0 = check_sweep(c))) return rc0;
    const uint32_t* tot = c->pin_totals.as<uint32_t>();
    c->ncons = tot[0];
    c->npoints = tot[1];
    const uint32_t islands = tot[2];
    if (c->sleep_model && (rc0 = take_woken(c))) return rc0;
    const uint32_t cache_used = tot[4];
    const uint32_t np = c->npoints;
    int rc;
    // The store at the end of this step adds at most ncons keys.
    if ((rc = ensure_cache(c, size_t(cache_used) + c->ncons))) return rc;
    const size_t np1 = std::max<uint32_t>(np, 1);
    KZG_HIP(c, c->p_body.ensure(np1));
    for (auto* d : {&c->p_pa, &c->p_pb, &c->p_n, &c->p_t1, &c->p_t2, &c->p_ra, &c->p_rb, &c->p_la, &c->p_lb, &c->p_imp,
                    &c->vr_ra, &c->vr_rb, &c->vr_n, &c->vr_t1, &c->vr_t2, &c->vr_imp, &c->wr_ra, &c->wr_rb,
                    &c->wr_qa, &c->wr_qb, &c->wr_n, &c->wr_t1, &c->wr_t2, &c->wr_imp})
        KZG_HIP(c, d->ensure(np1));
    KZG_HIP(c, c->vr_bd.ensure(np1));
    KZG_HIP(c, c->wr_bd.ensure(np1));
    KZG_HIP(c, c->vr_fb.ensure(np1));
    KZG_HIP(c, c->p_nvel.ensure(np1));
    // The constraint table comes down while the build kernel runs; the host
    // derives the solve schedule from it.
    KZG_HIP(c, c->pin_info.ensure((sizeof(U4) + sizeof(uint32_t)) * std::max<uint32_t>(c->ncons, 1)));
    c->h_info = c->pin_info.as<U4>();
    c->h_con_island = reinterpret_cast<const uint32_t*>(c->h_info + c->ncons);
    if (c->ncons) {
        KZG_HIP(c, hipMemcpyAsync(c->pin_info.p, c->c_info.p, sizeof(U4) * c->ncons, hipMemcpyDeviceToHost, c->stream));
        KZG_HIP(c, hipMemcpyAsync(const_cast<uint32_t*>(c->h_con_island), c->con_island.p,
                                  sizeof(uint32_t) * c->ncons, hipMemcpyDeviceToHost, c->stream));
    }
    KZG_HIP(c, hipEventRecord(c->ev[12], c->stream));
    KZG_HIP(c, record(c, 2));
    launch_build(c->stream, c->contacts(), c->manifolds(), c->bodies(), c->cache(), c->ncons,
                 mode == KZG_WARM_COLD ? 1 : 0);
    KZG_HIP(c, hipGetLastError());
    KZG_HIP(c, record(c, 3));
    KZG_HIP(c, hipEventSynchronize(c->ev[12]));
    // Host time of the build proper starts once the table is here.
    const auto t0 = std::chrono::steady_clock::now();
    const bool shared = c->sleep_model && shares_bodies(c);
    c->gens.clear();
    ScheduleInput si;
    si.num_bodies = c->nbodies;
    si.dynamic = c->h_dynamic.data();
    si.num_constraints = c->ncons;
    si.info = c->h_info;
    si.island_of_con = c->h_con_island;
    si.num_islands = islands;
    si.use_simd = c->world.use_simd_solver != 0;
    si.max_unit_vel = c->unit_cap_vel ? c->unit_cap_vel : kDefaultUnitVel;
    si.max_unit_ws = c->unit_cap_ws ? c->unit_cap_ws : kDefaultUnitWs;
    si.max_groups = c->cu_count * 4u;
    // With the SIMD solver and the cache warm start the velocity schedule may
    // finish on the builder's helper thread: kzg_solve sweeps the warm start
    // meanwhile.  (The learned warm start writes velocity records before the
    // sweeps, so it waits for the whole schedule.)
    const bool may_defer = c->world.use_simd_solver && mode != KZG_WARM_LEARNED && mode != KZG_WARM_LEARNED_MISS &&
                           c->early_ws_mode != 1;
    const uint32_t defer = !may_defer               ? ScheduleBuilder::kDeferNever
                           : c->early_ws_mode == 2 ? ScheduleBuilder::kDeferAlways
                                                   : ScheduleBuilder::kDeferLarge;
    const auto ts = std::chrono::steady_clock::now();
    if (shared) {
        c->early_ws = false;
        c->vel_after_ms = 0.0;
        if ((rc = schedule_generations(c, si, islands))) return rc;
        const auto te = std::chrono::steady_clock::now();
        c->build_host_ms = std::chrono::duration<double, std::milli>(te - t0).count();
        c->schedule_host_ms = std::chrono::duration<double, std::milli>(te - ts).count();
        c->phase = 2;
        return KZG_OK;
    }
    c->sched_builder.set_device_bucketing(c->device_bucketing == 0);
    // A velocity schedule bucketed on the device leaves the level walk's
    // arrays in pinned memory; the last upload from there has completed (the
    // stream has passed the table download above).
    if (c->device_bucketing == 0 && defer != ScheduleBuilder::kDeferNever) {
        const size_t cap = std::max<uint32_t>(c->npoints, 1);
        KZG_HIP(c, c->stage_seq.ensure(4 * sizeof(uint32_t) * cap));
        c->sched_builder.set_vel_sink(c->stage_seq.as<uint32_t>(), cap);
    } else {
        c->sched_builder.set_vel_sink(nullptr, 0);
    }
    c->early_ws = c->sched_builder.start(si, c->sched, defer);
    c->vel_after_ms = 0.0;
    // Host compute of the build ends here; what follows stages the
    // schedule for the device.
    const auto te = std::chrono::steady_clock::now();
    c->build_host_ms = std::chrono::duration<double, std::milli>(te - t0).count();
    c->schedule_host_ms = std::chrono::duration<double, std::milli>(te - ts).count();
    const LevelSchedule& ws = c->sched.ws;
    if (c->early_ws) {
        // The velocity schedule is still being built: only the warm-start one goes up.
        const size_t bytes = sizeof(uint32_t) * (ws.off.size() + ws.unit.size() + size_t(np)) +
                             sizeof(U2) * ws.grp.size() + 5 * 256;
        KZG_HIP(c, c->stage_b.ensure(bytes));
        size_t off = 0;
        if ((rc = upload_async(c, c->ws_off, c->stage_b, off, ws.off.data(), ws.off.size())) ||
            (rc = upload_async(c, c->ws_grp, c->stage_b, off, reinterpret_cast<const uint2*>(ws.grp.data()),
                               ws.grp.size())) ||
            (rc = upload_async(c, c->ws_pos, c->stage_b, off, ws.pos.data(), ws.pos.size())) ||
            (rc = upload_async(c, c->ws_unit, c->stage_b, off, ws.unit.data(), ws.unit.size())))
            return rc;
        c->phase = 2;
        return KZG_OK;
    }
    const LevelSchedule& vel = c->sched.vel;
    const size_t bytes = sizeof(uint32_t) * (ws.off.size() + vel.off.size() + ws.unit.size() + vel.unit.size() +
                                             2 * size_t(np)) +
                         sizeof(U2) * (ws.grp.size() + vel.grp.size()) + 10 * 256;
    KZG_HIP(c, c->stage_b.ensure(bytes));
    size_t off = 0;
    if ((rc = upload_async(c, c->ws_off, c->stage_b, off, ws.off.data(), ws.off.size())) ||
        (rc = upload_async(c, c->ws_grp, c->stage_b, off, reinterpret_cast<const uint2*>(ws.grp.data()), ws.grp.size())) ||
        (rc = upload_async(c, c->ws_pos, c->stage_b, off, ws.pos.data(), ws.pos.size())) ||
        (rc = upload_async(c, c->vel_off, c->stage_b, off, vel.off.data(), vel.off.size())) ||
        (rc = upload_async(c, c->vel_grp, c->stage_b, off, reinterpret_cast<const uint2*>(vel.grp.data()), vel.grp.size())) ||
        (rc = upload_async(c, c->vel_pos, c->stage_b, off, vel.pos.data(), vel.pos.size())) ||
        (rc = upload_async(c, c->ws_unit, c->stage_b, off, ws.unit.data(), ws.unit.size())) ||
        (rc = upload_async(c, c->vel_unit, c->stage_b, off, vel.unit.data(), vel.unit.size())))
        return rc;
    c->phase = 2;
    return KZG_OK;
}

int kzg_integrate_velocities(kzg_ctx* c, float dt) {
    if (!c) return KZG_ERR_ARGUMENT;
    hipSetDevice(c->device);
    KZG_HIP(c, record(c, 8));
    launch_integrate_velocities(c->stream, c->bodies(), c->nbodies, dt, c->world.gravity);
    KZG_HIP(c, hipGetLastError());
    KZG_HIP(c, record(c, 4));
    return KZG_OK;
}

int kzg_solve(kzg_ctx* c, float dt) {
    if (!c) return KZG_ERR_ARGUMENT;
    if (c->phase != 2) return fail(c, KZG_ERR_STATE, "kzg_solve needs a preceding kzg_build_constraints");
    hipSetDevice(c->device);
    const kz_world_desc& w = c->world;
    const bool learned = c->cur_mode == KZG_WARM_LEARNED || c->cur_mode == KZG_WARM_LEARNED_MISS;
    float4 *feat = nullptr, *base = nullptr;
    const bool collect = c->sample_cap && !c->no_collect;
    if ((learned || collect) && c->npoints) {
        KZG_HIP(c, c->feat.ensure(4 * size_t(c->npoints)));
        KZG_HIP(c, c->base.ensure(c->npoints));
        feat = c->feat.p;
        base = c->base.p;
    }
    const float g = std::sqrt(w.gravity[0] * w.gravity[0] + w.gravity[1] * w.gravity[1] + w.gravity[2] * w.gravity[2]);
    const bool early = c->early_ws;
    unsigned long long* load = nullptr;
    if (feat && (c->mlp_features & KZG_FEATURES_LOAD)) {
        KZG_HIP(c, c->body_load.ensure(std::max<uint32_t>(c->nbodies, 1)));
        load = c->body_load.p;
    }
    // Early warm start: the velocity records wait for their schedule.
    launch_prepare(c->stream, c->contacts(), c->bodies(), c->npoints, dt, w.restitution_threshold, w.baumgarte,
                   w.penetration_slop, w.max_bias_velocity, w.use_simd_solver ? 1 : 0, w.use_velocity_bias ? 1 : 0,
                   c->vel_records(), c->ws_records(), early ? nullptr : c->vel_pos.p, c->ws_pos.p,
                   (g > 0.0f ? g : 1.0f) * dt, feat, base, load, c->nbodies);
    KZG_HIP(c, hipGetLastError());
    if (learned)
        launch_mlp_warm(c->stream, c->mlp(), feat, base, c->npoints, c->cur_mode == KZG_WARM_LEARNED_MISS ? 1 : 0,
                        c->p_imp.p, c->p_la.p, c->vr_imp.p, c->wr_imp.p, c->vel_pos.p, c->ws_pos.p);
    KZG_HIP(c, c->resid.ensure(std::max<uint32_t>(w.velocity_iterations, 1)));
    KZG_HIP(c, hipMemsetAsync(c->resid.p, 0, sizeof(float) * std::max<uint32_t>(w.velocity_iterations, 1), c->stream));
    SolveParams sp{dt, w.warm_start_factor, w.velocity_iterations, w.position_iterations, w.use_simd_solver,
                   w.use_velocity_bias};
    if (!c->gens.empty()) {
        int rc = solve_generations(c, sp);
        if (rc) return rc;
    } else {
    KZG_HIP(c, c->ws_udesc.ensure(std::max<uint32_t>(c->sched.ws.units, 1)));
    launch_unit_desc(c->stream, c->ws_unit.p, c->ws_records().bd, c->sched.ws.units, c->ws_udesc.p);
    Schedule ws{c->ws_off.p, c->ws_grp.p, c->ws_udesc.p};
    // The SIMD velocity sweeps take two threads per unit (a lane pair).
    const uint32_t vel_lanes = w.use_simd_solver ? 2u : 1u;
    if (early && c->ncons) {
        // The warm-start sweep, while the host finishes the velocity schedule.
        KZG_HIP(c, c->b_packed_solver.ensure(8 * size_t(std::max<uint32_t>(c->nbodies, 1))));
        const UnitKernel uk = unit_kernel(c->sched.ws.largest_unit, 1u);
        uint32_t threads = 64;
        while (threads < c->sched.ws.max_width && threads < solve_threads(uk.ws, uk.vel)) threads *= 2;
        const CoopPlan coop = plan_cooperative(c->sched, c->coop_blocks, solve_threads(uk.ws, uk.vel), 1u,
                                               c->coop_placement == KZG_COOP_ONE_XCD, false);
        // The bias sweeps need only the warm-start schedule too: they run
        // here as well, and the velocity launch carries no bias levels.
        SolveParams spw = sp;
        spw.phases = kPhaseWarm | kPhaseBiasAlone;
        const Schedule none{nullptr, nullptr, nullptr};
        KZG_HIP(c, record(c, 14));
        if (coop.blocks > 1) {
            int rc;
            if ((rc = begin_cooperative(c))) return rc;
            KZG_HIP(c, launch_solve_split(c->stream, c->vel_records(), c->ws_records(), c->bodies(), c->nbodies,
                                          c->b_packed_solver.p, ws, none, spw, c->sched.groups, threads, c->resid.p,
                                          uk, coop.blocks, coop.threads, c->coop_sync.p, c->spin_limit_or_default(),
                                          c->coop_placement == KZG_COOP_ONE_XCD, c->ev[17], c->ev[18], kPackBodies));
            c->recorded |= 3u << 17;
        } else {
            launch_solve_groups(c->stream, c->vel_records(), c->ws_records(), c->bodies(), c->nbodies,
                                c->b_packed_solver.p, ws, none, spw, c->sched.groups, threads, c->resid.p, uk,
                                kPackBodies);
        }
        KZG_HIP(c, hipGetLastError());
        KZG_HIP(c, record(c, 15));
        sp.phases = kPhaseVelocity | kPhaseBiasDone;
    }
    if (early) {
        // The velocity schedule: wait for it, send it up, then its records.
        c->sched_builder.finish();
        c->vel_after_ms = c->sched_builder.vel_after_ms;
        c->build_host_ms += c->vel_after_ms;
        c->schedule_host_ms += c->vel_after_ms;
        const LevelSchedule& vs = c->sched.vel;
        int rc;
        if (vs.device_bucketed) {
            // The walk's per-point arrays go up; the device buckets them.
            const VelSequence q = c->sched_builder.velocity_sequence();
            const size_t bytes = sizeof(uint32_t) * ((q.in_sink ? 0 : 4 * size_t(q.n)) + q.groups + q.g0_levels) +
                                 sizeof(U2) * vs.grp.size() + sizeof(uint2) * q.num_runs + 9 * 256;
            KZG_HIP(c, c->stage_v.ensure(bytes));
            size_t off = 0;
            const uint32_t *d_pts, *d_uid, *d_lvl, *d_rank;
            if (q.in_sink) {
                // The walk wrote the four arrays into stage_seq: one transfer.
                KZG_HIP(c, c->vb_seq.ensure(4 * q.stride));
                KZG_HIP(c, hipMemcpyAsync(c->vb_seq.p, c->stage_seq.p, 4 * sizeof(uint32_t) * q.stride,
                                          hipMemcpyHostToDevice, c->stream));
                d_pts = c->vb_seq.p;
                d_uid = d_pts + q.stride;
                d_lvl = d_pts + 2 * q.stride;
                d_rank = d_pts + 3 * q.stride;
            } else {
                if ((rc = upload_async(c, c->vb_pts, c->stage_v, off, q.pts, q.n)) ||
                    (rc = upload_async(c, c->vb_uid, c->stage_v, off, q.uid, q.n)) ||
                    (rc = upload_async(c, c->vb_lvl, c->stage_v, off, q.lvl, q.n)) ||
                    (rc = upload_async(c, c->vb_rank, c->stage_v, off, q.rank, q.n)))
                    return rc;
                d_pts = c->vb_pts.p;
                d_uid = c->vb_uid.p;
                d_lvl = c->vb_lvl.p;
                d_rank = c->vb_rank.p;
            }
            if ((rc = upload_async(c, c->vb_runs, c->stage_v, off, reinterpret_cast<const uint2*>(q.runs),
                                   q.num_runs)) ||
                (rc = upload_async(c, c->vb_base, c->stage_v, off, q.group_base, q.groups)) ||
                (rc = upload_async(c, c->vel_grp, c->stage_v, off, reinterpret_cast<const uint2*>(vs.grp.data()),
                                   vs.grp.size())))
                return rc;
            const size_t n1 = std::max<uint32_t>(q.n, 1), l1 = size_t(vs.levels) + 1, u1 = size_t(vs.units) + 1;
            for (auto* d : {&c->vb_gl}) KZG_HIP(c, d->ensure(n1));
            for (auto* d : {&c->vb_count, &c->vb_fill}) KZG_HIP(c, d->ensure(l1));
            // Group 0's level counts straight into the head of the count array
            // (sized above: the upload then does not reallocate it).
            if ((rc = upload_async(c, c->vb_count, c->stage_v, off, q.g0_count, q.g0_levels))) return rc;
            for (auto* d : {&c->vb_uslot, &c->vb_usize}) KZG_HIP(c, d->ensure(u1));
            KZG_HIP(c, c->vel_off.ensure(l1));
            KZG_HIP(c, c->vel_unit.ensure(u1));
            KZG_HIP(c, c->vel_pos.ensure(std::max<uint32_t>(c->npoints, 1)));
            KZG_HIP(c, c->vb_sums.ensure(vel_bucket_sums_words(vs.levels, vs.units)));
            const VelBucketArgs a{d_pts, d_uid, d_lvl, q.n, c->vb_runs.p, q.num_runs,
                                  c->vb_base.p, d_rank, q.g0_levels, vs.levels, vs.units, c->npoints,
                                  c->vb_gl.p, c->vb_count.p,
                                  c->vb_fill.p, c->vb_uslot.p, c->vb_usize.p, c->vel_off.p, c->vel_unit.p,
                                  c->vel_pos.p, c->vb_sums.p};
            // The upload ends here; the bucketing is the solve's device work.
            KZG_HIP(c, record(c, 16));
            launch_vel_bucket(c->stream, a);
            KZG_HIP(c, hipGetLastError());
        } else {
            const size_t bytes = sizeof(uint32_t) * (vs.off.size() + vs.unit.size() + size_t(c->npoints)) +
                                 sizeof(U2) * vs.grp.size() + 5 * 256;
            KZG_HIP(c, c->stage_v.ensure(bytes));
            size_t off = 0;
            if ((rc = upload_async(c, c->vel_off, c->stage_v, off, vs.off.data(), vs.off.size())) ||
                (rc = upload_async(c, c->vel_grp, c->stage_v, off, reinterpret_cast<const uint2*>(vs.grp.data()),
                                   vs.grp.size())) ||
                (rc = upload_async(c, c->vel_pos, c->stage_v, off, vs.pos.data(), vs.pos.size())) ||
                (rc = upload_async(c, c->vel_unit, c->stage_v, off, vs.unit.data(), vs.unit.size())))
                return rc;
            KZG_HIP(c, record(c, 16));
        }
        launch_velocity_records(c->stream, c->contacts(), c->ws_records(), c->vel_records(), c->vel_pos.p,
                                c->ws_pos.p, c->npoints, w.use_simd_solver ? 1 : 0, w.use_velocity_bias ? 1 : 0);
    }
    KZG_HIP(c, c->vel_udesc.ensure(std::max<uint32_t>(c->sched.vel.units, 1)));
    launch_unit_desc(c->stream, c->vel_unit.p, c->vel_records().bd, c->sched.vel.units, c->vel_udesc.p);
    Schedule vel{c->vel_off.p, c->vel_grp.p, c->vel_udesc.p};
    const uint32_t width = std::max(c->sched.ws.max_width, vel_lanes * c->sched.vel.max_width);
    const UnitKernel uk = unit_kernel(c->sched.ws.largest_unit, c->sched.vel.largest_unit);
    uint32_t threads = 64;
    while (threads < width && threads < solve_threads(uk.ws, uk.vel)) threads *= 2;
    const CoopPlan coop = plan_cooperative(c->sched, c->coop_blocks, solve_threads(uk.ws, uk.vel), vel_lanes,
                                           c->coop_placement == KZG_COOP_ONE_XCD);
    c->coop_used = coop.blocks;
    c->coop_ws_points = coop.ws_points;
    c->coop_vel_points = coop.vel_points;
    KZG_HIP(c, record(c, 5));
    if (c->ncons) {
        KZG_HIP(c, c->b_packed_solver.ensure(8 * size_t(std::max<uint32_t>(c->nbodies, 1))));
        if (coop.blocks > 1) {
            int rc;
            if ((rc = begin_cooperative(c))) return rc;
            KZG_HIP(c, launch_solve_split(c->stream, c->vel_records(), c->ws_records(), c->bodies(), c->nbodies,
                                          c->b_packed_solver.p, ws, vel, sp, c->sched.groups, threads, c->resid.p,
                                          uk, coop.blocks, coop.threads, c->coop_sync.p, c->spin_limit_or_default(),
                                          c->coop_placement == KZG_COOP_ONE_XCD, c->ev[10], c->ev[11],
                                          early ? kUnpackBodies : kPackBodies | kUnpackBodies));
            c->recorded |= 3u << 10;
        } else {
            launch_solve_groups(c->stream, c->vel_records(), c->ws_records(), c->bodies(), c->nbodies,
                                c->b_packed_solver.p, ws, vel, sp, c->sched.groups, threads, c->resid.p, uk,
                                early ? kUnpackBodies : kPackBodies | kUnpackBodies);
        }
    }
    }  // one solve over all islands
    KZG_HIP(c, hipGetLastError());
    if (c->coop_launched) {
        KZG_HIP(c, hipMemcpyAsync(c->pin_totals.as<uint32_t>() + 5, c->coop_sync.p + kSyncGiveUp, sizeof(uint32_t),
                                  hipMemcpyDeviceToHost, c->stream));
        c->coop_launched = false;
    }
    KZG_HIP(c, record(c, 6));
    launch_writeback(c->stream, c->vel_records(), c->ws_records(), c->contacts(), c->vel_pos.p, c->ws_pos.p,
                     c->npoints);
    if (collect && c->npoints) {
        // Training samples: this step's features with its final impulses.
        const uint32_t n = std::min(c->npoints, c->sample_cap);
        launch_mlp_collect(c->stream, feat, base, c->p_imp.p, n, c->samples.p, c->sample_cap,
                           uint32_t(c->sample_count % c->sample_cap));
        c->sample_count += n;
    }
    if (!c->no_store) launch_cache_store(c->stream, c->contacts(), c->bodies(), c->cache(), c->ncons);
    KZG_HIP(c, hipGetLastError());
    KZG_HIP(c, record(c, 9));
    c->resid_points = c->npoints;
    c->phase = 3;
    return KZG_OK;
}

int kzg_integrate_positions(kzg_ctx* c, float dt) {
    if (!c) return KZG_ERR_ARGUMENT;
    hipSetDevice(c->device);
    launch_finish(c->stream, c->bodies(), c->nbodies, dt, c->world.use_velocity_bias ? 0 : 1);
    KZG_HIP(c, hipGetLastError());
    KZG_HIP(c, record(c, 7));
    c->phase = 0;
    return KZG_OK;
}

int kzg_step(kzg_ctx* c, const uint32_t* pairs, uint32_t n, float dt) {
    int rc;
    if ((rc = kzg_narrowphase(c, pairs, n))) return rc;
    if ((rc = kzg_build_constraints(c, c->warm_mode))) return rc;
    if ((rc = kzg_integrate_velocities(c, dt))) return rc;
    if ((rc = kzg_solve(c, dt))) return rc;
    if ((rc = kzg_integrate_positions(c, dt))) return rc;
    return kzg_update_sleeping(c, dt);
}

int kzg_query_sizes(kzg_ctx* c, uint32_t* np, uint32_t* nc, uint32_t* npts) {
    if (!c) return KZG_ERR_ARGUMENT;
    if (np) *np = c->npairs;
    if (nc) *nc = c->ncons;
    if (npts) *npts = c->npoints;
    return KZG_OK;
}

int kzg_cache_info(kzg_ctx* c, uint32_t* cap, uint32_t* used) {
    if (!c) return KZG_ERR_ARGUMENT;
    hipSetDevice(c->device);
    uint32_t u = 0;
    if (c->cache_cap) {
        int rc = download(c, &u, c->cache_used, 1);
        if (rc) return rc;
    }
    if (cap) *cap = c->cache_cap;
    if (used) *used = u;
    return KZG_OK;
}

int kzg_download_cache(kzg_ctx* c, void* out, uint64_t bytes) {
    if (!c || !out) return KZG_ERR_ARGUMENT;
    const size_t cap = c->cache_cap;
    if (bytes != cap * KZG_CACHE_SLOT_BYTES) return fail(c, KZG_ERR_ARGUMENT, "cache image size mismatch");
    hipSetDevice(c->device);
    char* o = static_cast<char*>(out);
    int rc;
    if ((rc = download(c, reinterpret_cast<unsigned long long*>(o), c->cache_key, cap)) ||
        (rc = download(c, reinterpret_cast<uint32_t*>(o + 8 * cap), c->cache_count, cap)) ||
        (rc = download(c, reinterpret_cast<float4*>(o + 12 * cap), c->cache_p, cap * kPairPoints)) ||
        (rc = download(c, reinterpret_cast<float2*>(o + 76 * cap), c->cache_t, cap * kPairPoints)))
        return rc;
    return KZG_OK;
}

int kzg_upload_cache(kzg_ctx* c, const void* in, uint64_t bytes, uint32_t cap, uint32_t used) {
    if (!c || !in || cap < 2 || (cap & (cap - 1)) || used > cap / 2 || bytes != uint64_t(cap) * KZG_CACHE_SLOT_BYTES)
        return fail(c, KZG_ERR_ARGUMENT, "bad cache image (capacity a power of two, load at most 1/2)");
    hipSetDevice(c->device);
    const char* p = static_cast<const char*>(in);
    int rc;
    if ((rc = upload(c, c->cache_key, reinterpret_cast<const unsigned long long*>(p), cap)) ||
        (rc = upload(c, c->cache_count, reinterpret_cast<const uint32_t*>(p + 8 * size_t(cap)), cap)) ||
        (rc = upload(c, c->cache_p, reinterpret_cast<const float4*>(p + 12 * size_t(cap)), size_t(cap) * kPairPoints)) ||
        (rc = upload(c, c->cache_t, reinterpret_cast<const float2*>(p + 76 * size_t(cap)), size_t(cap) * kPairPoints)) ||
        (rc = upload(c, c->cache_used, &used, 1)))
        return rc;
    c->cache_cap = cap;
    return KZG_OK;
}

int kzg_download_states(kzg_ctx* c, kz_body_state* out, uint32_t n) {
    if (!c || !out || n != c->nbodies) return fail(c, KZG_ERR_ARGUMENT, "state buffer does not match the body table");
    hipSetDevice(c->device);
    std::vector<float4> x(n), q(n), v(n), w(n);
    int rc;
    if ((rc = download(c, x.data(), c->b_x, n)) || (rc = download(c, q.data(), c->b_q, n)) ||
        (rc = download(c, v.data(), c->b_v, n)) || (rc = download(c, w.data(), c->b_w, n)))
        return rc;
    if ((rc = check_sweep(c))) return rc;
    for (uint32_t i = 0; i < n; ++i) {
        kz_body_state& s = out[i];
        s.position[0] = x[i].x; s.position[1] = x[i].y; s.position[2] = x[i].z;
        s.rotation[0] = q[i].x; s.rotation[1] = q[i].y; s.rotation[2] = q[i].z; s.rotation[3] = q[i].w;
        s.linear_velocity[0] = v[i].x; s.linear_velocity[1] = v[i].y; s.linear_velocity[2] = v[i].z;
        s.angular_velocity[0] = w[i].x; s.angular_velocity[1] = w[i].y; s.angular_velocity[2] = w[i].z;
    }
    return KZG_OK;
}

int kzg_download_manifolds(kzg_ctx* c, uint32_t* counts, kz_contact_point* pts) {
    if (!c || !counts || !pts) return fail(c, KZG_ERR_ARGUMENT, "null output");
    hipSetDevice(c->device);
    const uint32_t n = c->npairs;
    std::vector<float4> pa(size_t(n) * kPairPoints), pb(pa.size()), nn(pa.size());
    int rc;
    if ((rc = download(c, counts, c->m_count, n)) || (rc = download(c, pa.data(), c->m_pa, pa.size())) ||
        (rc = download(c, pb.data(), c->m_pb, pb.size())) || (rc = download(c, nn.data(), c->m_n, nn.size())))
        return rc;
    for (size_t i = 0; i < pa.size(); ++i) {
        kz_contact_point& o = pts[i];
        const bool live = (i % kPairPoints) < counts[i / kPairPoints];
        o.point_on_a[0] = live ? pa[i].x : 0.0f; o.point_on_a[1] = live ? pa[i].y : 0.0f; o.point_on_a[2] = live ? pa[i].z : 0.0f;
        o.point_on_b[0] = live ? pb[i].x : 0.0f; o.point_on_b[1] = live ? pb[i].y : 0.0f; o.point_on_b[2] = live ? pb[i].z : 0.0f;
        o.normal[0] = live ? nn[i].x : 0.0f; o.normal[1] = live ? nn[i].y : 1.0f; o.normal[2] = live ? nn[i].z : 0.0f;
        o.separation = live ? pa[i].w : 0.0f;
        uint32_t fid;
        std::memcpy(&fid, &pb[i].w, 4);
        o.feature_id = live ? fid : 0u;
    }
    return KZG_OK;
}

int kzg_download_constraints(kzg_ctx* c, uint32_t* hdr, float* pts) {
    if (!c || !hdr || !pts) return fail(c, KZG_ERR_ARGUMENT, "null output");
    hipSetDevice(c->device);
    const uint32_t np = c->npoints;
    std::vector<uint4> body(np);
    std::vector<float4> f[10];
    DevVec<float4>* src[10] = {&c->p_pa, &c->p_pb, &c->p_n, &c->p_t1, &c->p_t2, &c->p_ra, &c->p_rb, &c->p_la, &c->p_lb, &c->p_imp};
    std::vector<float> nvel(np);
    int rc = download(c, body.data(), c->p_body, np);
    if (rc) return rc;
    for (int k = 0; k < 10; ++k) {
        f[k].resize(np);
        if ((rc = download(c, f[k].data(), *src[k], np))) return rc;
    }
    if ((rc = download(c, nvel.data(), c->p_nvel, np))) return rc;
    for (uint32_t ci = 0; ci < c->ncons; ++ci) {
        const U4& info = c->h_info[ci];
        const uint32_t k = info.w & 0xFFu, first = info.z;
        uint32_t flags = (info.w >> 8) & 1u;
        if (k) flags = body[first].z;
        hdr[4 * ci] = info.x; hdr[4 * ci + 1] = info.y; hdr[4 * ci + 2] = k; hdr[4 * ci + 3] = flags;
        for (uint32_t j = 0; j < 8; ++j) {
            float* o = pts + (8 * size_t(ci) + j) * kPointFloats;
            std::memset(o, 0, sizeof(float) * kPointFloats);
            if (j >= k) continue;
            const uint32_t p = first + j;
            auto v3 = [&](int at, const float4& v) { o[at] = v.x; o[at + 1] = v.y; o[at + 2] = v.z; };
            const float4 &pa = f[0][p], &pb = f[1][p], &nn = f[2][p], &t1 = f[3][p], &t2 = f[4][p], &ra = f[5][p],
                         &rb = f[6][p], &la = f[7][p], &lb = f[8][p], &imp = f[9][p];
            v3(0, pa); v3(3, pb); v3(6, ra); v3(9, rb); v3(12, la); v3(15, lb); v3(18, nn); v3(21, t1); v3(24, t2);
            o[27] = nn.w; o[28] = t1.w; o[29] = t2.w;
            o[30] = imp.x; o[31] = imp.y; o[32] = imp.z;
            o[33] = pa.w; o[34] = ra.w; o[35] = rb.w; o[36] = imp.w; o[37] = la.w; o[38] = lb.w; o[39] = nvel[p];
            o[40] = pb.w;
        }
    }
    return KZG_OK;
}

int kzg_get_residuals(kzg_ctx* c, double* out, uint32_t cap) {
    if (!c || !out) return KZG_ERR_ARGUMENT;
    hipSetDevice(c->device);
    const uint32_t n = c->world.velocity_iterations;
    std::vector<float> r(std::max<uint32_t>(n, 1));
    int rc = download(c, r.data(), c->resid, n);
    if (rc) return rc;
    for (uint32_t i = 0; i < n && i < cap; ++i)
        out[i] = c->resid_points ? std::sqrt(double(r[i]) / double(c->resid_points)) : 0.0;
    return int(std::min(n, cap));
}

int kzg_get_natural_residual(kzg_ctx* c, double* out) {
    if (!c || !out) return KZG_ERR_ARGUMENT;
    if (c->phase != 3)
        return fail(c, KZG_ERR_STATE, "kzg_get_natural_residual needs a kzg_solve and no kzg_integrate_positions since");
    hipSetDevice(c->device);
    KZG_HIP(c, c->natural.ensure(4));
    launch_fill_zero(c->stream, c->natural.p, 4);
    launch_natural_residual(c->stream, c->contacts(), c->bodies(), c->npoints, c->world.use_velocity_bias ? 1 : 0,
                            c->natural.p);
    KZG_HIP(c, hipGetLastError());
    double acc[4];
    int rc = download(c, acc, c->natural, 4);
    if (rc) return rc;
    const double n = acc[3] > 0.0 ? acc[3] : 1.0;
    out[0] = std::sqrt((acc[0] + acc[1]) / n);
    out[1] = std::sqrt(acc[0] / n);
    out[2] = std::sqrt(acc[1] / n);
    out[3] = std::sqrt(acc[2] / n);
    return KZG_OK;
}

int kzg_get_timings(kzg_ctx* c, kzg_timings* out) {
    if (!c || !out) return KZG_ERR_ARGUMENT;
    join_schedule(c);
    hipSetDevice(c->device);
    KZG_HIP(c, hipStreamSynchronize(c->stream));
    int rc0;
    if ((rc0 = check_sweep(c))) return rc0;
    kzg_timings t{};
    // Only pairs of events recorded since the last narrowphase are read, so
    // no query fails and leaves a sticky runtime error behind.
    auto span = [&](int a, int b, float& out) {
        float ms = 0.0f;
        if ((c->recorded >> a & 1u) && (c->recorded >> b & 1u) && hipEventElapsedTime(&ms, c->ev[a], c->ev[b]) == hipSuccess)
            out = ms;
    };
    span(0, 1, t.narrowphase_ms);
    span(1, 13, t.build_device_ms);
    span(2, 3, t.build_kernel_ms);
    t.build_ms = float(c->build_host_ms);
    t.schedule_ms = float(c->schedule_host_ms);
    t.schedule_ws_ms = float(c->sched_builder.ws_ms);
    t.schedule_wait_ms = float(c->sched_builder.wait_ms);
    span(8, 4, t.integrate_velocities_ms);
    if (c->early_ws) {
        // The solve's device work without the wait for the velocity schedule
        // and its upload: the warm-start part, then the velocity part.
        float part = 0.0f, sweep = 0.0f;
        span(4, 15, t.early_warm_start_ms);
        span(16, 9, part);
        t.solve_ms = t.early_warm_start_ms + part;
        span(14, 15, sweep);
        span(5, 6, t.solve_kernel_ms);
        t.solve_kernel_ms += sweep;
        t.schedule_velocity_after_ms = float(c->vel_after_ms);
        t.overlap_ms = std::min(t.schedule_velocity_after_ms, t.integrate_velocities_ms + t.early_warm_start_ms);
    } else {
        span(4, 9, t.solve_ms);
        span(5, 6, t.solve_kernel_ms);
    }
    span(9, 7, t.integrate_positions_ms);
    (void)hipGetLastError();
    t.velocity_levels = c->sched.vel.max_depth;
    t.bias_levels = c->sched.ws.max_depth;
    t.groups = c->sched.groups;
    t.cooperative_blocks = c->coop_used;
    if (c->coop_used > 1) {
        t.cooperative_ws_points = c->coop_ws_points;
        t.cooperative_vel_points = c->coop_vel_points;
        span(10, 11, t.cooperative_kernel_ms);
        float early = 0.0f;
        span(17, 18, early);
        t.cooperative_kernel_ms += early;
    }
    t.islands = c->sched.islands;
    *out = t;
    return KZG_OK;
}

int kzg_schedule(uint32_t nb, const uint8_t* dyn, uint32_t nc, const uint32_t* info, uint32_t flags,
                 uint32_t max_groups, uint32_t* vel_level, uint32_t* ws_level, uint32_t* group, uint32_t* vel_unit,
                 uint32_t* ws_unit, uint32_t* stats) {
    if ((nb && !dyn) || (nc && !info) || !stats) return KZG_ERR_ARGUMENT;
    std::vector<U4> in(nc);
    uint32_t np = 0;
    for (uint32_t i = 0; i < nc; ++i) {
        in[i] = U4{info[4 * i], info[4 * i + 1], info[4 * i + 2], info[4 * i + 3]};
        if (in[i].x >= nb || in[i].y >= nb || in[i].z != np) return KZG_ERR_ARGUMENT;
        np += in[i].w & 0xFFu;
    }
    ScheduleBuilder b;
    ScheduleOutput out;
    ScheduleInput si;
    si.num_bodies = nb;
    si.dynamic = dyn;
    si.num_constraints = nc;
    si.info = in.data();
    si.use_simd = (flags & KZG_SCHEDULE_SIMD) != 0;
    si.max_unit_vel = (flags & KZG_SCHEDULE_UNITS) ? kDefaultUnitVel : 1u;
    si.max_unit_ws = (flags & KZG_SCHEDULE_UNITS) ? kDefaultUnitWs : 1u;
    si.max_groups = std::max<uint32_t>(max_groups, 1);
    if (flags & KZG_SCHEDULE_DEFER) {
        b.start(si, out, ScheduleBuilder::kDeferAlways);
        b.finish();
    } else {
        b.build(si, out);
    }
    auto fill = [&](const LevelSchedule& s, uint32_t* lvl, uint32_t* unit, uint32_t* grp) {
        if (lvl) std::fill(lvl, lvl + np, 0u);
        if (unit) std::fill(unit, unit + np, 0xFFFFFFFFu);
        for (uint32_t g = 0; g < s.grp.size(); ++g)
            for (uint32_t l = s.grp[g].x; l < s.grp[g].y; ++l)
                for (uint32_t u = s.off[l]; u < s.off[l + 1]; ++u)
                    for (uint32_t e = s.unit[u]; e < s.unit[u + 1]; ++e) {
                        if (lvl) lvl[s.idx[e]] = l - s.grp[g].x + 1;
                        if (unit) unit[s.idx[e]] = u;
                        if (grp) grp[s.idx[e]] = g;
                    }
    };
    fill(out.vel, vel_level, vel_unit, nullptr);
    fill(out.ws, ws_level, ws_unit, group);
    stats[0] = out.islands;
    stats[1] = out.groups;
    stats[2] = out.vel.max_depth;
    stats[3] = out.ws.max_depth;
    stats[4] = out.vel.units;
    stats[5] = out.ws.units;
    return KZG_OK;
}

int kzg_collide_batch(kzg_ctx* c, const uint32_t* sa, const uint32_t* sb, const float* xa, const float* xb,
                      uint32_t n, float max_sep, uint32_t* counts, kz_contact_point* pts, float* ms) {
    if (!c || (n && (!sa || !sb || !xa || !xb || !counts || !pts))) return fail(c, KZG_ERR_ARGUMENT, "null argument");
    if (n > kWorkPairMask) return fail(c, KZG_ERR_UNSUPPORTED, "more than 2^27 - 1 pairs");
    hipSetDevice(c->device);
    std::vector<float4> x(2 * size_t(n)), q(2 * size_t(n));
    std::vector<uint32_t> meta(2 * size_t(n));
    std::vector<uint2> pr(n);
    for (uint32_t i = 0; i < n; ++i) {
        if (sa[i] >= c->nshapes || sb[i] >= c->nshapes) return fail(c, KZG_ERR_ARGUMENT, "bad shape index");
        const float* a = xa + 7 * size_t(i);
        const float* b = xb + 7 * size_t(i);
        q[i] = f4(a[0], a[1], a[2], a[3]);
        x[i] = f4(a[4], a[5], a[6], 0.0f);
        q[n + i] = f4(b[0], b[1], b[2], b[3]);
        x[n + i] = f4(b[4], b[5], b[6], 0.0f);
        meta[i] = sa[i] | (2u << 24);
        meta[n + i] = sb[i] | (2u << 24);
        pr[i] = make_uint2(i, n + i);
    }
    int rc;
    if ((rc = upload(c, c->x_x, x.data(), x.size())) || (rc = upload(c, c->x_q, q.data(), q.size())) ||
        (rc = upload(c, c->x_meta, meta.data(), meta.size())) || (rc = upload(c, c->x_pairs, pr.data(), n)))
        return rc;
    KZG_HIP(c, c->m_count.ensure(std::max<uint32_t>(n, 1)));
    KZG_HIP(c, c->m_pa.ensure(size_t(std::max<uint32_t>(n, 1)) * kPairPoints));
    KZG_HIP(c, c->m_pb.ensure(size_t(std::max<uint32_t>(n, 1)) * kPairPoints));
    KZG_HIP(c, c->m_n.ensure(size_t(std::max<uint32_t>(n, 1)) * kPairPoints));
    KZG_HIP(c, c->work.ensure(work_entries(std::max<uint32_t>(n, 1))));
    KZG_HIP(c, c->gjk.ensure(size_t(std::max<uint32_t>(n, 1)) * kGjkRecBytes));
    BodyBuf bb{c->x_x.p, c->x_q.p, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, c->x_meta.p, nullptr};
    c->recorded = 0;
    KZG_HIP(c, record(c, 0));
    launch_narrowphase(c->stream, c->x_pairs.p, n, bb, c->shapes(), c->manifolds(), c->work.p, c->work_count.p,
                       c->gjk.p, max_sep, 0,
                       c->np_waves ? c->np_waves : std::min<uint32_t>(std::max<uint32_t>(n, 1), c->cu_count * 16u));
    KZG_HIP(c, hipGetLastError());
    KZG_HIP(c, record(c, 1));
    KZG_HIP(c, hipStreamSynchronize(c->stream));
    if (ms) KZG_HIP(c, hipEventElapsedTime(ms, c->ev[0], c->ev[1]));
    c->npairs = n;
    c->phase = 0;
    return kzg_download_manifolds(c, counts, pts);
}


int kzg_collide_batch_device(kzg_ctx* c, const uint32_t* shape_a, const uint32_t* shape_b, const float* xf_a,
                             const float* xf_b, uint32_t n, float max_sep, uint32_t* counts, kz_contact_point* points) {
    if (!c) return KZG_ERR_ARGUMENT;
    if (n && (!shape_a || !shape_b || !xf_a || !xf_b || !counts || !points))
        return fail(c, KZG_ERR_ARGUMENT, "kzg_collide_batch_device: null device pointer");
    if (n > kScratchPairs)
        return fail(c, KZG_ERR_ARGUMENT, "kzg_collide_batch_device: " + std::to_string(n) + " pairs exceed the "
                                             "allocation-free limit of " + std::to_string(kScratchPairs));
    if (!(max_sep >= 0.0f) || !std::isfinite(max_sep))
        return fail(c, KZG_ERR_ARGUMENT, "kzg_collide_batch_device: bad max separation");
    if (!c->scratch) return fail(c, KZG_ERR_STATE, "kzg_collide_batch_device: scratch not resolved");
    if (!n) return KZG_OK;
    if (!c->nshapes) return fail(c, KZG_ERR_STATE, "kzg_collide_batch_device: no shape table uploaded");
    join_schedule(c);
    hipSetDevice(c->device);
    unsigned char* base = c->scratch;
    auto at = [&](size_t off) { return static_cast<void*>(base + off); };
    BodyBuf bb{static_cast<float4*>(at(kScratch.x)), static_cast<float4*>(at(kScratch.q)), nullptr, nullptr, nullptr,
               nullptr, nullptr, nullptr, nullptr, nullptr, static_cast<uint32_t*>(at(kScratch.meta)),
               static_cast<uint32_t*>(at(kScratch.packed))};
    ManifoldBuf mb{static_cast<uint32_t*>(at(kScratch.mcount)), static_cast<float4*>(at(kScratch.mpa)),
                   static_cast<float4*>(at(kScratch.mpb)), static_cast<float4*>(at(kScratch.mn))};
    uint2* pairs = static_cast<uint2*>(at(kScratch.pairs));
    uint32_t* counters = static_cast<uint32_t*>(at(kScratch.counters));
    k_scratch_reset<<<1, 64, 0, c->stream>>>(counters, n);
    k_scratch_setup<<<(n + 255) / 256, 256, 0, c->stream>>>(shape_a, shape_b, xf_a, xf_b, n, c->nshapes, bb, pairs,
                                                             counters);
    launch_narrowphase_counted(c->stream, pairs, n, counters + 3, bb, c->shapes(), mb,
                               static_cast<uint32_t*>(at(kScratch.work)), counters, at(kScratch.gjk), max_sep,
                               c->np_waves);
    k_scratch_output<<<(n * kPairPoints + 255) / 256, 256, 0, c->stream>>>(mb, n, counts, points);
    KZG_HIP(c, hipGetLastError());
    uint32_t* flag = c->pin_flag.as<uint32_t>();
    KZG_HIP(c, hipMemcpyAsync(flag, counters + 4, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    KZG_HIP(c, hipStreamSynchronize(c->stream));
    if (*flag) return fail(c, KZG_ERR_ARGUMENT, "kzg_collide_batch_device: shape index outside the shape table "
                                                 "(those pairs report no contact)");
    return KZG_OK;
}

// ---------------------------------------------------------------------------
// Warm-start mode, iteration counts and the learned warm start.
// ---------------------------------------------------------------------------
int kzg_set_warm_start(kzg_ctx* c, int mode) {
    if (!c) return KZG_ERR_ARGUMENT;
    if ((mode & ~(0xFF | KZG_WARM_NO_STORE)) || (mode & 0xFF) > KZG_WARM_LEARNED_MISS)
        return fail(c, KZG_ERR_ARGUMENT, "unknown warm-start mode " + std::to_string(mode));
    c->warm_mode = mode;
    return KZG_OK;
}

int kzg_set_iterations(kzg_ctx* c, uint32_t velocity, uint32_t position) {
    if (!c) return KZG_ERR_ARGUMENT;
    if (velocity > 1024 || position > 1024) return fail(c, KZG_ERR_ARGUMENT, "iteration count above 1024");
    c->world.velocity_iterations = velocity;
    c->world.position_iterations = position;
    return KZG_OK;
}

int kzg_set_cooperative(kzg_ctx* c, uint32_t blocks) {
    if (!c) return KZG_ERR_ARGUMENT;
    if (blocks > kMaxCoopBlocks) return fail(c, KZG_ERR_ARGUMENT, "more than 64 cooperative workgroups");
    c->coop_blocks = blocks;
    return KZG_OK;
}

int kzg_set_narrowphase_waves(kzg_ctx* c, uint32_t waves) {
    if (!c) return KZG_ERR_ARGUMENT;
    if (waves > (1u << 20)) return fail(c, KZG_ERR_ARGUMENT, "more than 2^20 narrowphase waves");
    c->np_waves = waves;
    return KZG_OK;
}

int kzg_set_cooperative_placement(kzg_ctx* c, uint32_t placement) {
    if (!c) return KZG_ERR_ARGUMENT;
    if (placement != KZG_COOP_ONE_XCD && placement != KZG_COOP_SPREAD)
        return fail(c, KZG_ERR_ARGUMENT, "placement is KZG_COOP_ONE_XCD or KZG_COOP_SPREAD");
    c->coop_placement = placement;
    return KZG_OK;
}

int kzg_set_device_bucketing(kzg_ctx* c, uint32_t mode) {
    if (!c) return KZG_ERR_ARGUMENT;
    if (mode > 1) return fail(c, KZG_ERR_ARGUMENT, "device bucketing mode is 0 (auto) or 1 (off)");
    c->device_bucketing = mode;
    return KZG_OK;
}

int kzg_set_spin_limit(kzg_ctx* c, uint32_t polls) {
    if (!c) return KZG_ERR_ARGUMENT;
    c->spin_limit = polls;
    return KZG_OK;
}

int kzg_set_island_lds_limit(kzg_ctx* c, uint32_t max_bodies) {
    if (!c) return KZG_ERR_ARGUMENT;
    c->island_lds_limit = max_bodies;
    return KZG_OK;
}

int kzg_set_early_warm_start(kzg_ctx* c, uint32_t mode) {
    if (!c) return KZG_ERR_ARGUMENT;
    if (mode > 2) return fail(c, KZG_ERR_ARGUMENT, "early warm start mode is 0 (auto), 1 (off) or 2 (always)");
    c->early_ws_mode = mode;
    return KZG_OK;
}

int kzg_set_sweep_units(kzg_ctx* c, uint32_t velocity, uint32_t warm_start) {
    if (!c) return KZG_ERR_ARGUMENT;
    if (velocity > kMaxUnit || warm_start > kMaxUnit)
        return fail(c, KZG_ERR_ARGUMENT, "a sweep unit holds at most 4 points");
    c->unit_cap_vel = velocity;
    c->unit_cap_ws = warm_start;
    return KZG_OK;
}

}  // extern "C"

namespace {

int mlp_alloc(kzg_ctx* c) {
    if (c->mlp_master.p) return KZG_OK;
    hipSetDevice(c->device);
    KZG_HIP(c, c->mlp_master.ensure(kMlpParams));
    KZG_HIP(c, c->mlp_m.ensure(kMlpParams));
    KZG_HIP(c, c->mlp_v.ensure(kMlpParams));
    KZG_HIP(c, c->mlp_norm.ensure(32));
    KZG_HIP(c, c->mlp_out.ensure(kMlpParams + 1));
    KZG_HIP(c, c->mlp_w1b.ensure(64 * 16));
    KZG_HIP(c, c->mlp_w2b.ensure(64 * 64));
    KZG_HIP(c, c->mlp_w3b.ensure(16 * 64));
    // Zero padding of the BF16 operands (rows 3-15 of W3).
    KZG_HIP(c, hipMemsetAsync(c->mlp_w1b.p, 0, 64 * 16 * sizeof(uint16_t), c->stream));
    KZG_HIP(c, hipMemsetAsync(c->mlp_w3b.p, 0, 16 * 64 * sizeof(uint16_t), c->stream));
    std::vector<float> norm(32, 0.0f);
    for (int f = 0; f < kMlpIn; ++f) norm[16 + f] = 1.0f;
    return upload(c, c->mlp_norm, norm.data(), 32);
}

// Upload master weights, reset Adam and refresh the BF16 copies.
int mlp_load(kzg_ctx* c, const float* params) {
    int rc;
    if ((rc = mlp_alloc(c)) || (rc = upload(c, c->mlp_master, params, kMlpParams))) return rc;
    KZG_HIP(c, hipMemsetAsync(c->mlp_m.p, 0, kMlpParams * sizeof(float), c->stream));
    KZG_HIP(c, hipMemsetAsync(c->mlp_v.p, 0, kMlpParams * sizeof(float), c->stream));
    launch_mlp_sync(c->stream, c->mlp());
    KZG_HIP(c, hipGetLastError());
    KZG_HIP(c, hipStreamSynchronize(c->stream));
    c->adam_t = 0;
    c->mlp_ready = true;
    return KZG_OK;
}

int mlp_check_rows(kzg_ctx* c, const float* rows, uint32_t n) {
    if (!rows && n) return fail(c, KZG_ERR_ARGUMENT, "null row array");
    if (!c->mlp_ready) return fail(c, KZG_ERR_STATE, "no network weights (kzg_mlp_init or kzg_mlp_set_params)");
    return KZG_OK;
}

}  // namespace

extern "C" {

int kzg_mlp_init(kzg_ctx* c, uint64_t seed) {
    if (!c) return KZG_ERR_ARGUMENT;
    // He-uniform weights, zero biases; SplitMix64 stream.
    uint64_t st = seed;
    auto uni = [&]() {
        st += 0x9E3779B97F4A7C15ull;
        uint64_t z = st;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return double(z >> 11) * (1.0 / 9007199254740992.0);
    };
    std::vector<float> p(kMlpParams, 0.0f);
    auto fill = [&](int off, int count, int fan_in) {
        const double bound = std::sqrt(6.0 / fan_in);
        for (int k = 0; k < count; ++k) p[off + k] = float((2.0 * uni() - 1.0) * bound);
    };
    fill(kMlpW1, kMlpHid * kMlpIn, kMlpIn);
    fill(kMlpW2, kMlpHid * kMlpHid, kMlpHid);
    fill(kMlpW3, kMlpOut * kMlpHid, kMlpHid);
    return mlp_load(c, p.data());
}

int kzg_mlp_set_params(kzg_ctx* c, const float* params, uint32_t n) {
    if (!c || !params) return KZG_ERR_ARGUMENT;
    if (n != uint32_t(kMlpParams)) return fail(c, KZG_ERR_ARGUMENT, "expected " + std::to_string(kMlpParams) + " parameters");
    return mlp_load(c, params);
}

int kzg_mlp_get_params(kzg_ctx* c, float* params, uint32_t n) {
    if (!c || !params) return KZG_ERR_ARGUMENT;
    if (n != uint32_t(kMlpParams)) return fail(c, KZG_ERR_ARGUMENT, "expected " + std::to_string(kMlpParams) + " parameters");
    if (!c->mlp_ready) return fail(c, KZG_ERR_STATE, "no network weights");
    hipSetDevice(c->device);
    return download(c, params, c->mlp_master, kMlpParams);
}

int kzg_mlp_set_normalization(kzg_ctx* c, const float* mean, const float* inv_std) {
    if (!c || !mean || !inv_std) return KZG_ERR_ARGUMENT;
    int rc;
    if ((rc = mlp_alloc(c))) return rc;
    std::vector<float> norm(32, 0.0f);
    for (int f = 0; f < kMlpIn; ++f) {
        norm[f] = mean[f];
        norm[16 + f] = inv_std[f];
    }
    return upload(c, c->mlp_norm, norm.data(), 32);
}

int kzg_mlp_get_normalization(kzg_ctx* c, float* mean, float* inv_std) {
    if (!c || !mean || !inv_std) return KZG_ERR_ARGUMENT;
    int rc;
    if ((rc = mlp_alloc(c))) return rc;
    std::vector<float> norm(32);
    if ((rc = download(c, norm.data(), c->mlp_norm, 32))) return rc;
    for (int f = 0; f < kMlpIn; ++f) {
        mean[f] = norm[f];
        inv_std[f] = norm[16 + f];
    }
    return KZG_OK;
}

int kzg_mlp_collect(kzg_ctx* c, uint32_t capacity) {
    if (!c) return KZG_ERR_ARGUMENT;
    hipSetDevice(c->device);
    if (capacity) KZG_HIP(c, c->samples.ensure(5 * size_t(capacity)));
    c->sample_cap = capacity;
    c->sample_count = 0;
    return KZG_OK;
}

int kzg_mlp_num_samples(kzg_ctx* c, uint64_t* written, uint32_t* held) {
    if (!c) return KZG_ERR_ARGUMENT;
    if (written) *written = c->sample_count;
    if (held) *held = uint32_t(std::min<uint64_t>(c->sample_count, c->sample_cap));
    return KZG_OK;
}

int kzg_mlp_get_samples(kzg_ctx* c, uint32_t first, uint32_t count, float* out) {
    if (!c || (count && !out)) return KZG_ERR_ARGUMENT;
    const uint64_t held = std::min<uint64_t>(c->sample_count, c->sample_cap);
    if (uint64_t(first) + count > held) return fail(c, KZG_ERR_ARGUMENT, "sample range beyond the collected rows");
    if (!count) return KZG_OK;
    hipSetDevice(c->device);
    KZG_HIP(c, c->stage2.ensure(size_t(count) * 5 * sizeof(float4)));
    KZG_HIP(c, hipMemcpyAsync(c->stage2.p, c->samples.p + 5 * size_t(first), size_t(count) * 5 * sizeof(float4),
                              hipMemcpyDeviceToHost, c->stream));
    KZG_HIP(c, hipStreamSynchronize(c->stream));
    std::memcpy(out, c->stage2.p, size_t(count) * 5 * sizeof(float4));
    return KZG_OK;
}

int kzg_mlp_fit_normalization(kzg_ctx* c) {
    if (!c) return KZG_ERR_ARGUMENT;
    const uint32_t held = uint32_t(std::min<uint64_t>(c->sample_count, c->sample_cap));
    if (!held) return fail(c, KZG_ERR_STATE, "no collected samples");
    // Mean and standard deviation of every feature over at most 65536 rows
    // spread evenly over the ring (double accumulation on the host).
    const uint32_t take = std::min<uint32_t>(held, 65536), stride = held / take;
    const size_t row_bytes = kSampleRow * sizeof(float);
    hipSetDevice(c->device);
    KZG_HIP(c, c->stage2.ensure(size_t(take) * row_bytes));
    KZG_HIP(c, hipMemcpy2DAsync(c->stage2.p, row_bytes, c->samples.p, row_bytes * stride, row_bytes, take,
                                hipMemcpyDeviceToHost, c->stream));
    KZG_HIP(c, hipStreamSynchronize(c->stream));
    const float* rows = c->stage2.as<float>();
    double sum[kMlpIn] = {}, sq[kMlpIn] = {};
    for (uint32_t k = 0; k < take; ++k) {
        const float* r = rows + size_t(k) * kSampleRow;
        for (int f = 0; f < kMlpIn; ++f) {
            sum[f] += r[f];
            sq[f] += double(r[f]) * r[f];
        }
    }
    float mean[kMlpIn], istd[kMlpIn];
    for (int f = 0; f < kMlpIn; ++f) {
        const double m = sum[f] / take, var = std::max(0.0, sq[f] / take - m * m), sd = std::sqrt(var);
        mean[f] = float(m);
        istd[f] = sd > 1e-6 ? float(1.0 / sd) : 0.0f;  // a constant feature is switched off
    }
    return kzg_mlp_set_normalization(c, mean, istd);
}

int kzg_set_learned_features(kzg_ctx* c, uint32_t flags) {
    if (!c) return KZG_ERR_ARGUMENT;
    if (flags & ~uint32_t(KZG_FEATURES_LOAD)) return fail(c, KZG_ERR_ARGUMENT, "unknown feature flags");
    c->mlp_features = flags;
    return KZG_OK;
}

int kzg_mlp_set_target_clip(kzg_ctx* c, float clip) {
    if (!c) return KZG_ERR_ARGUMENT;
    if (!(clip >= 0.0f) || !std::isfinite(clip)) return fail(c, KZG_ERR_ARGUMENT, "target clip must be finite and >= 0");
    c->mlp_clip = clip;
    return KZG_OK;
}

int kzg_mlp_train(kzg_ctx* c, uint32_t steps, uint32_t batch, float lr, uint32_t seed, float* losses) {
    if (!c) return KZG_ERR_ARGUMENT;
    if (!c->mlp_ready) return fail(c, KZG_ERR_STATE, "no network weights");
    const uint32_t held = uint32_t(std::min<uint64_t>(c->sample_count, c->sample_cap));
    if (!held) return fail(c, KZG_ERR_STATE, "no collected samples");
    if (batch < 16 || batch > (1u << 20)) return fail(c, KZG_ERR_ARGUMENT, "batch outside [16, 2^20]");
    if (!(lr > 0.0f)) return fail(c, KZG_ERR_ARGUMENT, "learning rate must be positive");
    hipSetDevice(c->device);
    const uint32_t rows = mlp_grad_blocks(batch);
    KZG_HIP(c, c->mlp_partial.ensure(size_t(rows) * (kMlpParams + 1)));
    KZG_HIP(c, c->mlp_loss.ensure(std::max<uint32_t>(steps, 1)));
    const float b1 = 0.9f, b2 = 0.999f;
    for (uint32_t s = 0; s < steps; ++s) {
        ++c->adam_t;
        AdamParams ap{lr, b1, b2, 1e-8f, float(1.0 - std::pow(double(b1), double(c->adam_t))),
                      float(1.0 - std::pow(double(b2), double(c->adam_t))), 1.0f / (3.0f * float(batch)), rows};
        launch_mlp_grad(c->stream, c->mlp(), c->samples.p, held, batch, 0, seed + 0x9E3779B9u * s, c->mlp_partial.p,
                        c->mlp_clip);
        launch_mlp_adam(c->stream, c->mlp(), c->mlp_partial.p, ap, c->mlp_loss.p + s);
    }
    KZG_HIP(c, hipGetLastError());
    KZG_HIP(c, hipStreamSynchronize(c->stream));
    if (losses && steps) return download(c, losses, c->mlp_loss, steps);
    return KZG_OK;
}

int kzg_mlp_forward(kzg_ctx* c, const float* features, uint32_t n, float* out) {
    if (!c || (n && !out)) return KZG_ERR_ARGUMENT;
    int rc;
    if ((rc = mlp_check_rows(c, features, n))) return rc;
    if (!n) return KZG_OK;
    hipSetDevice(c->device);
    if ((rc = upload(c, c->x_samples, reinterpret_cast<const float4*>(features), 4 * size_t(n)))) return rc;
    KZG_HIP(c, c->x_out.ensure(size_t(n)));
    launch_mlp_infer(c->stream, c->mlp(), c->x_samples.p, n, c->x_out.p);
    KZG_HIP(c, hipGetLastError());
    return download(c, reinterpret_cast<float4*>(out), c->x_out, n);
}

int kzg_mlp_gradients(kzg_ctx* c, const float* samples, uint32_t n, float* grad, float* loss) {
    if (!c || !grad) return KZG_ERR_ARGUMENT;
    int rc;
    if ((rc = mlp_check_rows(c, samples, n))) return rc;
    if (n < 1 || n > (1u << 20)) return fail(c, KZG_ERR_ARGUMENT, "sample count outside [1, 2^20]");
    hipSetDevice(c->device);
    if ((rc = upload(c, c->x_samples, reinterpret_cast<const float4*>(samples), 5 * size_t(n)))) return rc;
    const uint32_t rows = mlp_grad_blocks(n);
    KZG_HIP(c, c->mlp_partial.ensure(size_t(rows) * (kMlpParams + 1)));
    launch_mlp_grad(c->stream, c->mlp(), c->x_samples.p, n, n, 1, 0, c->mlp_partial.p, c->mlp_clip);
    launch_mlp_reduce(c->stream, c->mlp_partial.p, rows, c->mlp_out.p);
    KZG_HIP(c, hipGetLastError());
    std::vector<float> g(kMlpParams + 1);
    if ((rc = download(c, g.data(), c->mlp_out, kMlpParams + 1))) return rc;
    std::memcpy(grad, g.data(), kMlpParams * sizeof(float));
    if (loss) *loss = g[kMlpParams];
    return KZG_OK;
}

}  // extern "C"
