// mlp.h - learned warm start: a per-contact-point MLP 16 -> 64 -> 64 -> 3
// (ReLU hidden layers) with BF16 weights and FP32 accumulation on the
// gfx950 matrix cores.
//
// The network predicts the accumulated normal and friction impulses of a
// contact point from its contact-frame features; it stands in for (or
// complements) the warm-start match of PhysicsWorld::BuildContactConstraints
// (src/World/PhysicsWorld.cpp:514-530), and is trained on the impulses the
// solver stores at the end of a step (PhysicsWorld.cpp:568-581).
//
// Feature row (16 floats per point, written by the Prepare kernel; all 16
// are network inputs):
//    0  normal relative velocity (vA + wA x rA - vB - wB x rB) . n
//    1  relative velocity . t1        2  relative velocity . t2
//    3-5 normal n (world)             6  penetration depth (-separation)
//    7  inverse mass A                8  inverse mass B
//    9  (rA x n) . IA^-1 (rA x n)    10  (rB x n) . IB^-1 (rB x n)
//   11  mean over t1, t2 of (rA x t) . IA^-1 (rA x t)
//   12  the same for B               13  1 if the manifold cache supplied the point, else 0
//   14  load of body A, 15  load of body B: with the load features on, the
//       body's summed cached normal impulses (clamped at 0) over all its
//       contact points, in units of its weight impulse m |gravity| dt (0 for
//       a static or kinematic body); else 0, which leaves the 14 features
//       above as the network's only information.
// The impulse scale s = normal mass * |gravity| * dt of the point rides in
// the w slot of its base row (the cache's clamped impulses).  Network
// outputs y are impulse corrections in units of s on top of the cache's
// clamped impulses c (zero where the cache has no match, which feature 13
// tells the network): lambda_n = max(c_n + y0 s, 0),
// lambda_t1,2 = clamp(c_t1,2 + y1,2 s, -mu lambda_n, mu lambda_n).  Training
// targets are (final impulses - c) / s.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace kz {

constexpr int kMlpIn = 16, kMlpHid = 64, kMlpOut = 3;
constexpr int kMlpW1 = 0;                                // [64][16] row-major (unit, feature)
constexpr int kMlpB1 = kMlpW1 + kMlpHid * kMlpIn;        // [64]
constexpr int kMlpW2 = kMlpB1 + kMlpHid;                 // [64][64] (unit, input unit)
constexpr int kMlpB2 = kMlpW2 + kMlpHid * kMlpHid;       // [64]
constexpr int kMlpW3 = kMlpB2 + kMlpHid;                 // [3][64]
constexpr int kMlpB3 = kMlpW3 + kMlpOut * kMlpHid;       // [3]
constexpr int kMlpParams = kMlpB3 + kMlpOut;             // 5443
constexpr int kFeatRow = 16;                             // floats per feature row
constexpr int kSampleRow = 20;                           // feature row + (3 target corrections / s, 0)

struct MlpBuf {
    float* master;     // [kMlpParams] FP32 master weights
    uint16_t* w1b;     // [64][16] BF16
    uint16_t* w2b;     // [64][64] BF16
    uint16_t* w3b;     // [16][64] BF16, rows 3-15 zero
    float* norm;       // [32] feature mean[16], inverse std[16]
    float* adam_m;     // [kMlpParams]
    float* adam_v;     // [kMlpParams]
};

struct AdamParams {
    float lr, beta1, beta2, eps;
    float corr1, corr2;   // 1 - beta1^t, 1 - beta2^t
    float loss_scale;     // 1 / (3 * batch): turns the summed squared error into the mean
    uint32_t rows;        // partial-gradient rows to sum
};

void launch_mlp_sync(hipStream_t s, const MlpBuf& m);
void launch_mlp_infer(hipStream_t s, const MlpBuf& m, const float4* feat, uint32_t n, float4* out);
void launch_mlp_warm(hipStream_t s, const MlpBuf& m, const float4* feat, const float4* base, uint32_t np,
                     int only_miss, float4* c_imp, const float4* c_la, float4* vr_imp, float4* wr_imp,
                     const uint32_t* vel_pos, const uint32_t* ws_pos);
void launch_mlp_collect(hipStream_t s, const float4* feat, const float4* base, const float4* c_imp, uint32_t np,
                        float4* samples, uint32_t cap, uint32_t start);
// One minibatch: forward + backward over `batch` samples, per-block partial
// gradients (and summed squared error) into partial[blocks][kMlpParams + 1].
// identity != 0 takes samples 0..batch-1 in order, else hashed indices in [0, n).
uint32_t mlp_grad_blocks(uint32_t batch);
// clip > 0 clamps the targets to [-clip, clip] (0: as stored).
void launch_mlp_grad(hipStream_t s, const MlpBuf& m, const float4* samples, uint32_t n, uint32_t batch,
                     int identity, uint32_t seed, float* partial, float clip = 0.0f);
void launch_mlp_reduce(hipStream_t s, const float* partial, uint32_t rows, float* out);
void launch_mlp_adam(hipStream_t s, const MlpBuf& m, const float* partial, const AdamParams& ap, float* loss_out);

}  // namespace kz
