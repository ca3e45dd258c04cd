// ctx_internal.h - the kzg_ctx object shared by the single-world C-ABI
// (context.hip) and the batched-worlds C-ABI (batch.hip): device buffers,
// staging, events and the upload / download helpers.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <unordered_set>
#include <vector>

#include "kz_device.h"
#include "kz_math.h"
#include "kzg.h"
#include "mlp.h"
#include "schedule.h"

namespace kz {
// work: 2 n entries; work_count: 2 counters; gjk: n records of kGjkRecBytes.
void launch_narrowphase(hipStream_t s, const uint2* pairs, uint32_t n, const BodyBuf& b, const ShapeBuf& sb,
                        const ManifoldBuf& m, uint32_t* work, uint32_t* work_count, void* gjk, float max_sep,
                        int world_mode, uint32_t wave_blocks);
// The same kernels with the pair count read on the device and no memset:
// the two list counters must already be zero.  world_mode 1 (skip pairs
// with no dynamic body, packed-id order).  work: work_entries(cap).
void launch_narrowphase_counted(hipStream_t s, const uint2* pairs, uint32_t cap, const uint32_t* n_dev,
                                const BodyBuf& b, const ShapeBuf& sb, const ManifoldBuf& m, uint32_t* work,
                                uint32_t* work_count, void* gjk, float max_sep, uint32_t wave_blocks);
void launch_integrate_velocities(hipStream_t s, const BodyBuf& b, uint32_t n, float dt, const float* g);
void launch_build(hipStream_t s, const ContactBuf& c, const ManifoldBuf& m, const BodyBuf& b, const CacheBuf& cache,
                  uint32_t nc, int warm);
void launch_prepare(hipStream_t s, const ContactBuf& c, const BodyBuf& b, uint32_t np, float dt, float rthr,
                    float baum, float slop, float maxb, int use_simd, int use_vbias, const VelRecBuf& vr,
                    const WsRecBuf& wr, const uint32_t* vel_pos, const uint32_t* ws_pos, float gdt, float4* feat,
                    float4* base, unsigned long long* load = nullptr, uint32_t nbodies = 0);
void launch_writeback(hipStream_t s, const VelRecBuf& vr, const WsRecBuf& wr, const ContactBuf& c,
                      const uint32_t* vel_pos, const uint32_t* ws_pos, uint32_t np);
// `bodies`: kPackBodies / kUnpackBodies (kz_device.h).
void launch_solve_groups(hipStream_t s, const VelRecBuf& vr, const WsRecBuf& wr, const BodyBuf& b, uint32_t nbodies,
                         float4* packed, const Schedule& ws, const Schedule& vel, const SolveParams& sp,
                         uint32_t groups, uint32_t threads, float* resid, UnitKernel uk,
                         uint32_t bodies = kPackBodies | kUnpackBodies);
void launch_velocity_records(hipStream_t s, const ContactBuf& c, const WsRecBuf& wr, const VelRecBuf& vr,
                             const uint32_t* vel_pos, const uint32_t* ws_pos, uint32_t np, int use_simd,
                             int use_vbias);
void launch_unit_desc(hipStream_t s, const uint32_t* unit, const uint4* bd, uint32_t nu, uint4* desc);
hipError_t launch_solve_split(hipStream_t s, const VelRecBuf& vr, const WsRecBuf& wr, const BodyBuf& b,
                              uint32_t nbodies, float4* packed, const Schedule& ws, const Schedule& vel,
                              const SolveParams& sp, uint32_t groups, uint32_t threads, float* resid,
                              UnitKernel uk, uint32_t coop_blocks, uint32_t coop_threads, uint32_t* sync,
                              uint32_t limit, bool one_xcd, hipEvent_t coop_begin, hipEvent_t coop_end,
                              uint32_t bodies = kPackBodies | kUnpackBodies);
// Clears a sync block of kSyncWords words, its give-up flag included.
void launch_sync_reset(hipStream_t s, uint32_t* sync);
// The velocity sequence's bucketing on the device (bucket.hip): inputs from
// the host's level walk (ScheduleBuilder::velocity_sequence), scratch, and
// the outputs the sweeps read (level offsets, unit bounds, point records).
struct VelBucketArgs {
    const uint32_t* pts;
    const uint32_t* uid;
    const uint32_t* lvl;
    uint32_t n;
    const uint2* runs;
    uint32_t num_runs;
    const uint32_t* group_base;
    const uint32_t* rank;  // [n] place in its level of a group-0 unit's first point, else 0xFFFFFFFF
    uint32_t g0_levels;    // count[0 .. g0_levels) holds group 0's counts already
    uint32_t levels, units, num_points;
    uint32_t* gl;      // [n]
    uint32_t* count;   // [levels]
    uint32_t* fill;    // [levels]
    uint32_t* uslot;   // [units]
    uint32_t* usize;   // [units]
    uint32_t* off;     // [levels + 1] out
    uint32_t* unit;    // [units + 1] out
    uint32_t* pos;     // [num_points] out
    uint32_t* sums;    // [vel_bucket_sums_words(levels, units)] scan scratch
};
uint32_t vel_bucket_sums_words(uint32_t levels, uint32_t units);
void launch_vel_bucket(hipStream_t s, const VelBucketArgs& a);
void launch_compact(hipStream_t s, const uint32_t* m_count, uint32_t n, const uint2* pairs, const BodyBuf& b,
                    const ContactBuf& c, uint2* tile, uint32_t* totals);
// lds_limit: most bodies for the one-workgroup LDS union-find (0: the
// kernel's own limit); larger worlds take the global-memory union-find.
void launch_islands(hipStream_t s, const ContactBuf& c, uint32_t cap, const BodyBuf& b, uint32_t nb,
                    uint32_t* parent, uint32_t* island, uint2* tile, uint32_t* totals, uint32_t* con_island,
                    uint32_t lds_limit = 0);
void launch_island_index(hipStream_t s, const uint32_t* act, uint32_t* sleep, const BodyBuf& b, uint32_t* isl);
void launch_islands_indexed(hipStream_t s, const ContactBuf& c, uint32_t cap, const BodyBuf& b, uint32_t nb,
                            const uint32_t* isl, const uint32_t* sleep, uint32_t* parent, uint32_t* island,
                            uint2* tile, uint32_t* totals, uint32_t* con_island);
void launch_wake(hipStream_t s, const ContactBuf& c, const uint32_t* totals, const BodyBuf& b, uint32_t nb,
                 uint32_t* t, uint32_t* act, uint32_t* sleep, float* timer, uint32_t limit);
void launch_sleep_decide(hipStream_t s, const ContactBuf& c, uint32_t cap, const uint32_t* totals, const BodyBuf& b,
                         const uint32_t* act, uint32_t n, float* timer, uint32_t* isl, uint32_t* parent,
                         uint32_t* tmin, uint8_t* q, float dt, float lin2, float ang2, float thr);
void launch_set_activity(hipStream_t s, const BodyBuf& b, uint32_t nb, const uint8_t* active, float* timer,
                         uint32_t* isl);
void launch_cache_store(hipStream_t s, const ContactBuf& c, const BodyBuf& b, const CacheBuf& cache, uint32_t nc);
void launch_cache_rehash(hipStream_t s, const CacheBuf& src, const CacheBuf& dst);
// acc[4] (zeroed by the caller): squared sums of the normal, friction and
// cone terms of the natural-map residual, and the point count.
void launch_natural_residual(hipStream_t s, const ContactBuf& c, const BodyBuf& b, uint32_t np, int use_vbias,
                             double* acc);
void launch_finish(hipStream_t s, const BodyBuf& b, uint32_t n, float dt, int split);
void launch_world_inertia(hipStream_t s, const BodyBuf& b, uint32_t n);
}  // namespace kz

using namespace kz;

namespace kzctx {

// Device array that grows geometrically and never copies on growth.
template <class T> struct DevVec {
    T* p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t n) {
        if (n <= cap && p) return hipSuccess;
        size_t nc = std::max<size_t>(n, cap + cap / 2);
        nc = std::max<size_t>(nc, 64);
        if (p) hipFree(p);
        p = nullptr;
        cap = 0;
        hipError_t e = hipMalloc(&p, nc * sizeof(T));
        if (e == hipSuccess) cap = nc;
        return e;
    }
    void release() {
        if (p) hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

// Pinned host staging buffer.
struct Pinned {
    void* p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes) {
        if (bytes <= cap && p) return hipSuccess;
        size_t nc = std::max<size_t>(bytes, cap + cap / 2);
        nc = std::max<size_t>(nc, 4096);
        if (p) hipHostFree(p);
        p = nullptr;
        cap = 0;
        hipError_t e = hipHostMalloc(&p, nc, hipHostMallocDefault);
        if (e == hipSuccess) cap = nc;
        return e;
    }
    template <class T> T* as() { return static_cast<T*>(p); }
    void release() {
        if (p) hipHostFree(p);
        p = nullptr;
        cap = 0;
    }
};

constexpr int kPointFloats = 44;

}  // namespace kzctx

using namespace kzctx;

struct kzg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    kz_world_desc world{};
    uint32_t cu_count = 256;

    // shapes
    uint32_t nshapes = 0;
    std::vector<uint32_t> h_shape_type;
    DevVec<uint32_t> s_type;
    DevVec<float4> s_params;
    DevVec<uint4> s_hull;
    DevVec<float4> s_hverts;
    DevVec<uint32_t> s_htris;
    DevVec<float4> s_hfn;

    // bodies
    uint32_t nbodies = 0;
    std::vector<uint32_t> h_packed;
    std::vector<uint8_t> h_dynamic, h_sensor, h_motion;
    DevVec<float4> b_x, b_q, b_v, b_w, b_pv, b_pw, b_iw, b_il, b_mat, b_lim;
    DevVec<float4> b_packed_solver;  // 8 float4 per body during the solve
    DevVec<uint32_t> coop_sync;      // sync block of the cooperative solve (kSyncWords words)
    uint32_t coop_blocks = 0;        // kzg_set_cooperative: 0 auto, 1 off, n workgroups
    uint32_t coop_placement = 0;     // kzg_set_cooperative_placement: KZG_COOP_ONE_XCD or KZG_COOP_SPREAD
    uint32_t device_bucketing = 0;   // kzg_set_device_bucketing: 0 auto (on), 1 off
    // Device bucketing of the velocity sequence: the walk's per-point arrays,
    // island runs and group bases (uploaded), and the bucketing's scratch.
    DevVec<uint32_t> vb_pts, vb_uid, vb_lvl, vb_rank, vb_gl, vb_base, vb_count, vb_fill, vb_uslot, vb_usize, vb_sums;
    DevVec<uint32_t> vb_seq;  // the four per-point arrays in one block (from stage_seq)
    DevVec<uint2> vb_runs;
    uint32_t spin_limit = 0;         // kzg_set_spin_limit: 0 default
    bool coop_launched = false;      // this step's solve made a cooperative launch
    uint32_t spin_limit_or_default() const { return spin_limit ? spin_limit : kDefaultSpinLimit; }
    uint32_t unit_cap_vel = 0, unit_cap_ws = 0;  // kzg_set_sweep_units: 0 default, 1..kMaxUnit
    uint32_t coop_used = 1;          // workgroups the last solve gave group 0
    uint32_t coop_ws_points = 0, coop_vel_points = 0;
    DevVec<uint32_t> b_meta, b_packed;

    // Activation (BodyManager's active list, src/Body/BodyManager.cpp:138-160)
    // and sleeping (PhysicsWorld::UpdateSleeping, PhysicsWorld.cpp:866-951).
    // The host keeps the list: entries past h_act_n are the stale values
    // swap-removes leave there.  d_act is the device's copy; d_sleep holds
    // its length, the count of compact island indices, the bodies woken by
    // the last build and the wake pass's give-up flag.
    bool sleep_model = false;        // the active list may differ from every body in slot order
    float sleep_time = 0.5f, sleep_lin = 0.05f, sleep_ang = 0.1f;
    std::vector<uint32_t> h_act, h_act_pos;  // list, and each body's position in it (kNone: asleep)
    uint32_t h_act_n = 0;
    std::vector<uint8_t> h_may_sleep;
    DevVec<uint32_t> d_act, d_island, d_wake, d_sleep, sl_parent, sl_min;
    DevVec<float> d_timer;
    DevVec<uint8_t> sl_q;
    Pinned pin_sleep;                // d_sleep after a build; UpdateSleeping's decisions
    bool wake_launched = false;      // this step's build ran the wake-up pass
    std::vector<uint32_t> h_owner, h_orig;  // host scratch: island of each body, the list before UpdateSleeping
    // Islands that share a dynamic body (a sleeping body reached through
    // stale indices) are solved in generations, one solve each, in order:
    // per generation the offsets of its level, group and unit arrays in the
    // uploaded schedule arrays, its unit count, groups and launch shape.
    struct GenPlan {
        uint32_t ws_loff, ws_goff, ws_uoff, ws_udoff, ws_units;
        uint32_t vel_loff, vel_goff, vel_uoff, vel_udoff, vel_units;
        uint32_t groups, width, largest_ws, largest_vel;
    };
    std::vector<GenPlan> gens;  // empty: one solve over all islands
    std::vector<uint32_t> h_gen_mask;

    // pairs and manifolds
    uint32_t npairs = 0;
    DevVec<uint2> d_pairs;
    bool pairs_resident = false;  // d_pairs holds the last device broadphase query's pairs
    // device broadphase query (broadphase.hip)
    DevVec<float4> bp_nodes;
    DevVec<uint32_t> bp_layer, bp_mask, bp_counts, bp_sums, bp_counters;
    DevVec<uint2> bp_items_a, bp_items_b;
    Pinned pin_bp, stage_bp;
    uint32_t bp_items_cap = 0, bp_pairs_cap = 0;
    // the tree image kept on the device between queries, and the proxy
    // moves on the device (kzg_broadphase_move)
    uint32_t bp_num_nodes = 0, bp_num_layers = 0;  // of the image on the device (0: none)
    DevVec<float4> bp_local;     // per body: local bounds min, max
    DevVec<uint32_t> bp_proxy;   // per body: proxy (leaf node) id
    DevVec<uint32_t> bp_moved;   // [0] count, then (active position, 6 floats) records
    DevVec<uint32_t> bp_upd_ids; // records of kzg_broadphase_pairs_update
    DevVec<float4> bp_upd_rec;
    uint32_t bp_proxies = 0;     // bodies with proxies registered
    uint32_t bp_root = 0xFFFFFFFFu;  // root and its height at the last query
    int32_t bp_root_height = 0;
    DevVec<uint32_t> m_count;
    DevVec<float4> m_pa, m_pb, m_n;
    DevVec<uint32_t> work, work_count;
    DevVec<uint8_t> gjk;  // GJK results handed from the lane kernel to the wave kernel

    // constraints and points
    uint32_t ncons = 0, npoints = 0;
    const U4* h_info = nullptr;           // in pin_info, valid after kzg_build_constraints
    const uint32_t* h_con_island = nullptr;  // in pin_info, after the table
    DevVec<uint4> c_info;
    DevVec<uint32_t> c_pair;
    DevVec<uint4> p_body;
    DevVec<float4> p_pa, p_pb, p_n, p_t1, p_t2, p_ra, p_rb, p_la, p_lb, p_imp;
    DevVec<float> p_nvel;

    // schedule
    ScheduleBuilder sched_builder;
    ScheduleOutput sched;
    DevVec<uint32_t> ws_off, vel_off, ws_pos, vel_pos, ws_unit, vel_unit;
    DevVec<uint4> ws_udesc, vel_udesc;
    DevVec<uint2> ws_grp, vel_grp;
    // sweep records
    DevVec<uint4> vr_bd, wr_bd;
    DevVec<float4> vr_ra, vr_rb, vr_n, vr_t1, vr_t2, vr_imp;
    DevVec<float2> vr_fb;
    DevVec<float4> wr_ra, wr_rb, wr_qa, wr_qb, wr_n, wr_t1, wr_t2, wr_imp;

    // warm-start cache
    DevVec<unsigned long long> cache_key;
    DevVec<uint32_t> cache_used;
    DevVec<uint32_t> cache_count;
    DevVec<float4> cache_p;
    DevVec<float2> cache_t;
    uint32_t cache_cap = 0;

    // residuals
    DevVec<float> resid;
    DevVec<double> natural;  // natural-map residual sums
    uint32_t resid_points = 0;

    // staging and timing
    Pinned stage, stage2, stage_a, stage_b, stage_v;  // stage_v: a deferred velocity schedule
    Pinned stage_seq;  // the level walk's per-point arrays of a velocity schedule bucketed on the device
    Pinned pin_totals;  // constraints, points and cache occupancy of the step
    Pinned pin_info;    // the step's constraint table (the schedule's input)
    DevVec<uint2> compact_tile;
    DevVec<uint32_t> compact_tot;  // constraints, points, islands, 0
    DevVec<uint32_t> uf_parent, uf_island, con_island;
    hipEvent_t ev[19] = {};  // 12: the constraint table has reached the host; 13: device build end;
    // 14/15: early warm-start sweep begin/end, 16: velocity part begin, 17/18: early cooperative kernel
    uint32_t recorded = 0;  // bit i: ev[i] recorded since the last narrowphase
    kzg_timings timings{};
    double build_host_ms = 0.0, schedule_host_ms = 0.0;
    // This step's warm-start sweep was launched before the velocity schedule
    // was ready (the builder finishes it in kzg_solve); the host time the
    // velocity schedule took after the warm-start one was ready.
    bool early_ws = false;
    uint32_t np_waves = 0;       // kzg_set_narrowphase_waves: 0 auto (16 per CU, at most one per pair)
    uint32_t island_lds_limit = 0;  // kzg_set_island_lds_limit: 0 default
    uint32_t early_ws_mode = 0;  // kzg_set_early_warm_start: 0 auto, 1 off, 2 always
    double vel_after_ms = 0.0;
    int phase = 0;  // 0 idle, 1 narrowphase done, 2 built, 3 solved

    // learned warm start
    bool mlp_ready = false;
    DevVec<float> mlp_master, mlp_m, mlp_v, mlp_norm, mlp_partial, mlp_out, mlp_loss;
    DevVec<uint16_t> mlp_w1b, mlp_w2b, mlp_w3b;
    uint32_t adam_t = 0;
    float mlp_clip = 0.0f;     // kzg_mlp_set_target_clip: 0 none
    uint32_t mlp_features = 0; // kzg_set_learned_features: KZG_FEATURES_LOAD fills features 14-15
    DevVec<unsigned long long> body_load;  // per-body load sums of the load features
    DevVec<float4> feat;       // feature rows of the current step, 4 float4 per point
    DevVec<float4> base;       // the cache's clamped impulses of the current step, 1 float4 per point
    DevVec<float4> samples;    // training ring, 5 float4 per sample
    DevVec<float4> x_samples;  // caller-supplied rows (forward / gradient entry points)
    DevVec<float4> x_out;      // forward outputs
    uint32_t sample_cap = 0;
    uint64_t sample_count = 0;
    int warm_mode = KZG_WARM_CACHE;  // kzg_step's mode
    int cur_mode = KZG_WARM_CACHE;   // mode of the step being built and solved
    bool no_store = false;
    bool no_collect = false;  // this step's solve appends no training rows

    // allocation-free standalone narrowphase: this device's copy of the
    // library's scratch array (resolved in kzg_create) and a pinned word for
    // its error flag
    unsigned char* scratch = nullptr;
    Pinned pin_flag;

    // standalone batch scratch
    DevVec<float4> x_x, x_q;
    DevVec<uint32_t> x_meta, x_packed;
    DevVec<uint2> x_pairs;

    BodyBuf bodies() {
        return BodyBuf{b_x.p, b_q.p, b_v.p, b_w.p, b_pv.p, b_pw.p, b_iw.p, b_il.p, b_mat.p, b_lim.p, b_meta.p, b_packed.p};
    }
    ShapeBuf shapes() { return ShapeBuf{s_type.p, s_params.p, s_hull.p, s_hverts.p, s_htris.p, s_hfn.p}; }
    ManifoldBuf manifolds() { return ManifoldBuf{m_count.p, m_pa.p, m_pb.p, m_n.p}; }
    ContactBuf contacts() {
        return ContactBuf{c_info.p, c_pair.p, p_body.p, p_pa.p, p_pb.p, p_n.p, p_t1.p, p_t2.p,
                          p_ra.p, p_rb.p, p_la.p, p_lb.p, p_imp.p, p_nvel.p};
    }
    CacheBuf cache() { return CacheBuf{cache_key.p, cache_used.p, cache_count.p, cache_p.p, cache_t.p, cache_cap}; }
    MlpBuf mlp() {
        return MlpBuf{mlp_master.p, mlp_w1b.p, mlp_w2b.p, mlp_w3b.p, mlp_norm.p, mlp_m.p, mlp_v.p};
    }
    VelRecBuf vel_records() { return VelRecBuf{vr_bd.p, vr_ra.p, vr_rb.p, vr_n.p, vr_t1.p, vr_t2.p, vr_fb.p, vr_imp.p}; }
    WsRecBuf ws_records() {
        return WsRecBuf{wr_bd.p, wr_ra.p, wr_rb.p, wr_qa.p, wr_qb.p, wr_n.p, wr_t1.p, wr_t2.p, wr_imp.p};
    }
};


namespace kzctx {

inline int fail(kzg_ctx* c, int code, const std::string& msg);

inline hipError_t record(kzg_ctx* c, int k) {
    c->recorded |= 1u << k;
    return hipEventRecord(c->ev[k], c->stream);
}

inline int fail(kzg_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

#define KZG_HIP(ctx, call)                                                                            \
    do {                                                                                              \
        hipError_t e_ = (call);                                                                       \
        if (e_ != hipSuccess)                                                                         \
            return fail(ctx, KZG_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_));      \
    } while (0)

template <class T> inline int upload(kzg_ctx* c, DevVec<T>& d, const T* src, size_t n) {
    KZG_HIP(c, d.ensure(std::max<size_t>(n, 1)));
    if (n) {
        KZG_HIP(c, c->stage.ensure(n * sizeof(T)));
        std::memcpy(c->stage.p, src, n * sizeof(T));
        KZG_HIP(c, hipMemcpyAsync(d.p, c->stage.p, n * sizeof(T), hipMemcpyHostToDevice, c->stream));
        KZG_HIP(c, hipStreamSynchronize(c->stream));
    }
    return KZG_OK;
}

template <class T> inline int download(kzg_ctx* c, T* dst, const DevVec<T>& d, size_t n) {
    if (!n) return KZG_OK;
    KZG_HIP(c, c->stage2.ensure(n * sizeof(T)));
    KZG_HIP(c, hipMemcpyAsync(c->stage2.p, d.p, n * sizeof(T), hipMemcpyDeviceToHost, c->stream));
    KZG_HIP(c, hipStreamSynchronize(c->stream));
    std::memcpy(dst, c->stage2.p, n * sizeof(T));
    return KZG_OK;
}

inline float4 f4(float a, float b, float cc, float d) { return make_float4(a, b, cc, d); }

// A deferred velocity schedule must be complete before anything reads the
// host arrays it is built from or the schedule itself.
inline void join_schedule(kzg_ctx* c) {
    if (c->sched_builder.pending()) c->sched_builder.finish();
}

inline int ensure_cache(kzg_ctx* c, size_t needed_keys) {
    // Load factor at most 1/2 so that probing always ends quickly.
    uint32_t want = c->cache_cap ? c->cache_cap : 1024u;
    while (size_t(want) < 2 * needed_keys) want *= 2u;
    if (want == c->cache_cap) return KZG_OK;
    DevVec<unsigned long long> nk;
    DevVec<uint32_t> nused, ncnt;
    DevVec<float4> np;
    DevVec<float2> nt;
    KZG_HIP(c, nk.ensure(want));
    KZG_HIP(c, nused.ensure(1));
    KZG_HIP(c, ncnt.ensure(want));
    KZG_HIP(c, np.ensure(size_t(want) * kPairPoints));
    KZG_HIP(c, nt.ensure(size_t(want) * kPairPoints));
    KZG_HIP(c, hipMemsetAsync(nk.p, 0, sizeof(unsigned long long) * want, c->stream));
    KZG_HIP(c, hipMemsetAsync(nused.p, 0, sizeof(uint32_t), c->stream));
    CacheBuf dst{nk.p, nused.p, ncnt.p, np.p, nt.p, want};
    if (c->cache_cap) launch_cache_rehash(c->stream, c->cache(), dst);
    KZG_HIP(c, hipGetLastError());
    KZG_HIP(c, hipStreamSynchronize(c->stream));
    c->cache_key.release();
    c->cache_used.release();
    c->cache_count.release();
    c->cache_p.release();
    c->cache_t.release();
    c->cache_key = nk;
    c->cache_used = nused;
    c->cache_count = ncnt;
    c->cache_p = np;
    c->cache_t = nt;
    c->cache_cap = want;
    return KZG_OK;
}

// Asynchronous upload through a pinned staging buffer: the caller guarantees
// that the buffer is not reused before the stream has drained.
template <class T> inline int upload_async(kzg_ctx* c, DevVec<T>& d, Pinned& st, size_t& off, const T* src, size_t n) {
    KZG_HIP(c, d.ensure(std::max<size_t>(n, 1)));
    if (!n) return KZG_OK;
    off = (off + 255) & ~size_t(255);
    std::memcpy(static_cast<char*>(st.p) + off, src, n * sizeof(T));
    KZG_HIP(c, hipMemcpyAsync(d.p, static_cast<char*>(st.p) + off, n * sizeof(T), hipMemcpyHostToDevice, c->stream));
    off += n * sizeof(T);
    return KZG_OK;
}

}  // namespace kzctx

using namespace kzctx;
