// mlp.hip - learned warm start on the gfx950 matrix cores: inference,
// training (forward + backward + Adam) of the 16 -> 64 -> 64 -> 3 network.
//
// Matrix layout.  A wave works on a tile of 16 contact points with
// v_mfma_f32_16x16x16_bf16: D[16x16] += A[16x16] B[16x16].  Lane l holds
// A[l%16][4(l/16)+e], B[4(l/16)+e][l%16] and D[4(l/16)+e][l%16], e = 0..3.
// Every layer is computed as out[units x points] = W[units x in] * act[in x points],
// so a layer's D tile (units 4g+e of point l%16 in lane l) is exactly the B
// operand of the next layer: activations stay in registers, the weights
// come from LDS.  The backward pass runs the transposed products the same
// way; the weight gradients (sums over points) need the activations with
// points along K, which a per-wave LDS transpose provides.
//
// Rounding points (mirrored by the CPU oracle): normalised inputs, weights,
// hidden activations and backpropagated errors are rounded to BF16 (round to
// nearest even) where they enter a matrix product; every product
// accumulates in FP32; biases, bias gradients, losses and the Adam update
// stay FP32.
#include "mlp.h"

#include <algorithm>

namespace kz {

namespace {

typedef short bf4 __attribute__((ext_vector_type(4)));
typedef float fx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}
__device__ __forceinline__ bf4 pack4(const float* v) {
    bf4 r;
    r[0] = short(to_bf16(v[0]));
    r[1] = short(to_bf16(v[1]));
    r[2] = short(to_bf16(v[2]));
    r[3] = short(to_bf16(v[3]));
    return r;
}
__device__ __forceinline__ bf4 ld_bf4(const uint16_t* p) { return *reinterpret_cast<const bf4*>(p); }
__device__ __forceinline__ bf4 gather_bf4(const uint16_t* p, int stride) {
    bf4 r;
    r[0] = short(p[0]);
    r[1] = short(p[stride]);
    r[2] = short(p[2 * stride]);
    r[3] = short(p[3 * stride]);
    return r;
}
__device__ __forceinline__ fx4 mfma(bf4 a, bf4 b, fx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ fx4 zero4() { return fx4{0.0f, 0.0f, 0.0f, 0.0f}; }

// Weights of one network in LDS.
struct alignas(16) MlpLds {
    uint16_t w1[64 * 16];
    uint16_t w2[64 * 64];
    uint16_t w3[16 * 64];
    float b1[64], b2[64], b3[4];
    float mean[16], istd[16];
};

__device__ void load_weights(MlpLds& L, const MlpBuf& m) {
    const uint32_t t = threadIdx.x, nt = blockDim.x;
    const uint32_t* s1 = reinterpret_cast<const uint32_t*>(m.w1b);
    const uint32_t* s2 = reinterpret_cast<const uint32_t*>(m.w2b);
    const uint32_t* s3 = reinterpret_cast<const uint32_t*>(m.w3b);
    uint32_t* d1 = reinterpret_cast<uint32_t*>(L.w1);
    uint32_t* d2 = reinterpret_cast<uint32_t*>(L.w2);
    uint32_t* d3 = reinterpret_cast<uint32_t*>(L.w3);
    for (uint32_t k = t; k < 64 * 16 / 2; k += nt) d1[k] = s1[k];
    for (uint32_t k = t; k < 64 * 64 / 2; k += nt) d2[k] = s2[k];
    for (uint32_t k = t; k < 16 * 64 / 2; k += nt) d3[k] = s3[k];
    for (uint32_t k = t; k < 64; k += nt) {
        L.b1[k] = m.master[kMlpB1 + k];
        L.b2[k] = m.master[kMlpB2 + k];
    }
    for (uint32_t k = t; k < 4; k += nt) L.b3[k] = k < 3 ? m.master[kMlpB3 + k] : 0.0f;
    for (uint32_t k = t; k < 16; k += nt) {
        L.mean[k] = m.norm[k];
        L.istd[k] = m.norm[16 + k];
    }
    __syncthreads();
}

// Activations of one 16-point tile, forward pass.  Lane (j = l%16, g = l/16)
// holds features / units 4g+e (+16m for tile m) of point j.
struct Act {
    bf4 x;            // normalised inputs, features 4g+e
    bf4 h1[4], h2[4]; // hidden layers, units 16m+4g+e
    uint32_t mask1, mask2;  // bit 4m+e: pre-activation > 0
    fx4 y;            // outputs 4g+e (valid in lane group 0, e < 3)
};

__device__ __forceinline__ void mlp_forward(const MlpLds& L, float4 xin, int lane, Act& a) {
    const int i = lane & 15, g = lane >> 4;
    const float xs[4] = {xin.x, xin.y, xin.z, xin.w};
    float v[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int f = 4 * g + e;
        v[e] = f < kMlpIn ? (xs[e] - L.mean[f]) * L.istd[f] : 0.0f;
    }
    a.x = pack4(v);
    a.mask1 = a.mask2 = 0u;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const fx4 acc = mfma(ld_bf4(&L.w1[(16 * m + i) * 16 + 4 * g]), a.x, zero4());
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const float z = acc[e] + L.b1[16 * m + 4 * g + e];
            const bool on = z > 0.0f;
            a.mask1 |= uint32_t(on) << (4 * m + e);
            v[e] = on ? z : 0.0f;
        }
        a.h1[m] = pack4(v);
    }
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        fx4 acc = zero4();
#pragma unroll
        for (int t = 0; t < 4; ++t) acc = mfma(ld_bf4(&L.w2[(16 * m + i) * 64 + 16 * t + 4 * g]), a.h1[t], acc);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const float z = acc[e] + L.b2[16 * m + 4 * g + e];
            const bool on = z > 0.0f;
            a.mask2 |= uint32_t(on) << (4 * m + e);
            v[e] = on ? z : 0.0f;
        }
        a.h2[m] = pack4(v);
    }
    fx4 acc = zero4();
#pragma unroll
    for (int t = 0; t < 4; ++t) acc = mfma(ld_bf4(&L.w3[i * 64 + 16 * t + 4 * g]), a.h2[t], acc);
#pragma unroll
    for (int e = 0; e < 4; ++e) a.y[e] = acc[e] + (g == 0 ? L.b3[e] : 0.0f);
}

__device__ __forceinline__ uint32_t sample_hash(uint32_t seed, uint32_t q) {
    uint32_t h = seed ^ (q * 0x9E3779B1u);
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    h ^= h >> 16;
    return h;
}

// Sum over the 16 lanes of a lane group (the points of a tile).
__device__ __forceinline__ float sum16(float v) {
    v += __shfl_xor(v, 1);
    v += __shfl_xor(v, 2);
    v += __shfl_xor(v, 4);
    v += __shfl_xor(v, 8);
    return v;
}

// ---------------------------------------------------------------------------
// Refresh the BF16 operand copies from the FP32 master weights.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void store_copy(const MlpBuf& m, int p, float w) {
    const uint16_t b = to_bf16(w);
    if (p < kMlpB1) {
        const int u = (p - kMlpW1) / kMlpIn, f = (p - kMlpW1) % kMlpIn;
        m.w1b[u * 16 + f] = b;
    } else if (p >= kMlpW2 && p < kMlpB2) {
        m.w2b[p - kMlpW2] = b;
    } else if (p >= kMlpW3 && p < kMlpB3) {
        m.w3b[p - kMlpW3] = b;  // rows 0-2 of [16][64]
    }
}

__global__ void k_mlp_sync(MlpBuf m) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < kMlpParams) store_copy(m, p, m.master[p]);
}

// ---------------------------------------------------------------------------
// Inference over feature rows; out[p] = normalised outputs (y0, y1, y2, 0).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_mlp_infer(MlpBuf m, const float4* feat, uint32_t n, float4* out) {
    __shared__ MlpLds L;
    load_weights(L, m);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, i = lane & 15, g = lane >> 4;
    const uint32_t tiles = (n + 15u) / 16u;
    for (uint32_t tile = blockIdx.x * 4u + wave; tile < tiles; tile += gridDim.x * 4u) {
        const uint32_t p = tile * 16u + i;
        const bool valid = p < n;
        const float4 xin = valid ? feat[size_t(p) * 4 + g] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        Act a;
        mlp_forward(L, xin, lane, a);
        if (valid && g == 0) out[p] = make_float4(a.y[0], a.y[1], a.y[2], 0.0f);
    }
}

// ---------------------------------------------------------------------------
// Warm start from the network: the cache's impulses (zero where the cache has
// no match) plus the predicted correction, clamped as Prepare clamps cached
// impulses (ContactSolver.cpp:107-110), into the constraint points and both
// sweep-record sets.  only_miss keeps the cache's impulses where the cache
// supplied the point.
// ---------------------------------------------------------------------------
struct WarmArgs {
    const float4* feat;
    const float4* base;
    uint32_t np;
    int only_miss;
    float4* c_imp;
    const float4* c_la;
    float4* vr_imp;
    float4* wr_imp;
    const uint32_t* vel_pos;
    const uint32_t* ws_pos;
};

__global__ void __launch_bounds__(256) k_mlp_warm(MlpBuf m, WarmArgs w) {
    __shared__ MlpLds L;
    load_weights(L, m);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, i = lane & 15, g = lane >> 4;
    const uint32_t tiles = (w.np + 15u) / 16u;
    for (uint32_t tile = blockIdx.x * 4u + wave; tile < tiles; tile += gridDim.x * 4u) {
        const uint32_t p = tile * 16u + i;
        const bool valid = p < w.np;
        const float4 xin = valid ? w.feat[size_t(p) * 4 + g] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        Act a;
        mlp_forward(L, xin, lane, a);
        if (valid && g == 0) {
            const float4 tail = w.feat[size_t(p) * 4 + 3];  // feature 12, cache hit, loads A and B
            if (!(w.only_miss && tail.y != 0.0f)) {
                const float4 b0 = w.base[p];  // clamped cache impulses, impulse scale s
                const float s = b0.w;
                const float ln = fmaxf(b0.x + a.y[0] * s, 0.0f);
                const float mf = w.c_la[p].w * ln;
                const float t1 = fminf(fmaxf(b0.y + a.y[1] * s, -mf), mf);
                const float t2 = fminf(fmaxf(b0.z + a.y[2] * s, -mf), mf);
                const float4 imp = make_float4(ln, t1, t2, 0.0f);
                w.c_imp[p] = imp;
                const uint32_t e = w.vel_pos[p];
                if (e != 0xFFFFFFFFu) w.vr_imp[e] = imp;
                const uint32_t q = w.ws_pos[p];
                if (q != 0xFFFFFFFFu) w.wr_imp[q] = imp;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Training samples: feature row + (final impulses - the cache's impulses) in
// units of the scale.
// ---------------------------------------------------------------------------
__global__ void k_mlp_collect(const float4* feat, const float4* base, const float4* c_imp, uint32_t np,
                              float4* samples, uint32_t cap, uint32_t start) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const size_t row = size_t((uint64_t(start) + p) % cap) * 5;
    float4 f[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        f[k] = feat[size_t(p) * 4 + k];
        samples[row + k] = f[k];
    }
    const float4 imp = c_imp[p], b0 = base[p];
    const float s = b0.w;
    samples[row + 4] = s > 0.0f ? make_float4((imp.x - b0.x) / s, (imp.y - b0.y) / s, (imp.z - b0.z) / s, 0.0f)
                                : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// ---------------------------------------------------------------------------
// Forward + backward of one minibatch.  Loss = mean over the batch and the
// three outputs of (y - target)^2.  Every block handles `tpb` tiles, its four
// waves take every fourth; all waves run the same number of rounds (tiles
// past the block's range are masked), so __syncthreads() can order the LDS
// transposes.
// ---------------------------------------------------------------------------
struct alignas(16) WaveScratch {
    uint16_t x[16 * 16];    // [feature][point]
    uint16_t h1[64 * 16];   // [unit][point]
    uint16_t h2[64 * 16];
    uint16_t dy[16 * 16];   // [output][point]
    uint16_t dz2[64 * 16];
    uint16_t dz1[64 * 16];
};
struct TrainLds {
    MlpLds w;
    union {
        WaveScratch s[4];
        float red[kMlpParams + 1];
    } u;
};

struct GradArgs {
    const float4* samples;  // 5 float4 per sample
    uint32_t n, batch, tpb, seed;
    int identity;
    float* partial;         // [gridDim.x][kMlpParams + 1]
    float clip;             // > 0: targets clamped to [-clip, clip]
};

__global__ void __launch_bounds__(256) k_mlp_grad(MlpBuf m, GradArgs ga) {
    __shared__ TrainLds L;
    load_weights(L.w, m);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, i = lane & 15, g = lane >> 4;
    WaveScratch& S = L.u.s[wave];
    fx4 gw2[4][4], gw1[4], gw3[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        gw1[a] = gw3[a] = zero4();
#pragma unroll
        for (int b = 0; b < 4; ++b) gw2[a][b] = zero4();
    }
    float gb1[16], gb2[16], gb3[4];
#pragma unroll
    for (int k = 0; k < 16; ++k) gb1[k] = gb2[k] = 0.0f;
#pragma unroll
    for (int k = 0; k < 4; ++k) gb3[k] = 0.0f;
    float loss = 0.0f;
    const float cscale = 2.0f / (3.0f * float(ga.batch));
    const uint32_t rounds = (ga.tpb + 3u) / 4u;
    for (uint32_t r = 0; r < rounds; ++r) {
        const uint32_t local = r * 4u + uint32_t(wave);
        const uint32_t q = (blockIdx.x * ga.tpb + local) * 16u + uint32_t(i);
        const bool valid = local < ga.tpb && q < ga.batch;
        const uint32_t idx = ga.identity ? q : sample_hash(ga.seed, q) % ga.n;
        const float4 z4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        const float4 xin = valid ? ga.samples[size_t(idx) * 5 + g] : z4;
        const float4 tg = valid ? ga.samples[size_t(idx) * 5 + 4] : z4;
        Act a;
        mlp_forward(L.w, xin, lane, a);
        auto cl = [&](float t) { return ga.clip > 0.0f ? fminf(fmaxf(t, -ga.clip), ga.clip) : t; };
        const float tv[4] = {cl(tg.x), cl(tg.y), cl(tg.z), 0.0f};
        float dy[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const bool on = valid && g == 0 && e < 3;
            const float d = a.y[e] - tv[e];
            dy[e] = on ? cscale * d : 0.0f;
            loss += on ? d * d : 0.0f;
            gb3[e] += dy[e];
        }
        const bf4 dyb = pack4(dy);
        // dh2 = W3^T dy, masked by the ReLU of layer 2.
        bf4 dz2b[4];
        float v[4];
#pragma unroll
        for (int mm = 0; mm < 4; ++mm) {
            const fx4 acc = mfma(gather_bf4(&L.w.w3[(4 * g) * 64 + 16 * mm + i], 64), dyb, zero4());
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                v[e] = (a.mask2 >> (4 * mm + e) & 1u) ? acc[e] : 0.0f;
                gb2[4 * mm + e] += v[e];
            }
            dz2b[mm] = pack4(v);
        }
        // dh1 = W2^T dz2, masked by the ReLU of layer 1.
        bf4 dz1b[4];
#pragma unroll
        for (int mm = 0; mm < 4; ++mm) {
            fx4 acc = zero4();
#pragma unroll
            for (int t = 0; t < 4; ++t)
                acc = mfma(gather_bf4(&L.w.w2[(16 * t + 4 * g) * 64 + 16 * mm + i], 64), dz2b[t], acc);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                v[e] = (a.mask1 >> (4 * mm + e) & 1u) ? acc[e] : 0.0f;
                gb1[4 * mm + e] += v[e];
            }
            dz1b[mm] = pack4(v);
        }
        // Transpose through LDS: rows (features / units / outputs) by points.
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            S.x[(4 * g + e) * 16 + i] = uint16_t(a.x[e]);
            S.dy[(4 * g + e) * 16 + i] = uint16_t(dyb[e]);
#pragma unroll
            for (int mm = 0; mm < 4; ++mm) {
                const int row = (16 * mm + 4 * g + e) * 16 + i;
                S.h1[row] = uint16_t(a.h1[mm][e]);
                S.h2[row] = uint16_t(a.h2[mm][e]);
                S.dz2[row] = uint16_t(dz2b[mm][e]);
                S.dz1[row] = uint16_t(dz1b[mm][e]);
            }
        }
        __syncthreads();
        // Weight gradients, K = the 16 points of the tile.
        const bf4 bx = ld_bf4(&S.x[i * 16 + 4 * g]);
        const bf4 ady = ld_bf4(&S.dy[i * 16 + 4 * g]);
        bf4 bh1[4], bh2[4];
#pragma unroll
        for (int nn = 0; nn < 4; ++nn) {
            bh1[nn] = ld_bf4(&S.h1[(16 * nn + i) * 16 + 4 * g]);
            bh2[nn] = ld_bf4(&S.h2[(16 * nn + i) * 16 + 4 * g]);
        }
#pragma unroll
        for (int mm = 0; mm < 4; ++mm) {
            const bf4 adz2 = ld_bf4(&S.dz2[(16 * mm + i) * 16 + 4 * g]);
            const bf4 adz1 = ld_bf4(&S.dz1[(16 * mm + i) * 16 + 4 * g]);
#pragma unroll
            for (int nn = 0; nn < 4; ++nn) gw2[mm][nn] = mfma(adz2, bh1[nn], gw2[mm][nn]);
            gw1[mm] = mfma(adz1, bx, gw1[mm]);
            gw3[mm] = mfma(ady, bh2[mm], gw3[mm]);
        }
        __syncthreads();
    }
    // Bias gradients and loss: sums over the points of the lane group.
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        gb1[k] = sum16(gb1[k]);
        gb2[k] = sum16(gb2[k]);
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) gb3[k] = sum16(gb3[k]);
    loss = sum16(loss);
    // The four waves add into LDS in wave order (deterministic).
    float* red = L.u.red;
    for (int wv = 0; wv < 4; ++wv) {
        if (wave == wv) {
            const bool first = wv == 0;
            auto put = [&](int k, float x) { red[k] = first ? x : red[k] + x; };
#pragma unroll
            for (int mm = 0; mm < 4; ++mm)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int u = 16 * mm + 4 * g + e;
#pragma unroll
                    for (int nn = 0; nn < 4; ++nn) put(kMlpW2 + u * 64 + 16 * nn + i, gw2[mm][nn][e]);
                    if (i < kMlpIn) put(kMlpW1 + u * kMlpIn + i, gw1[mm][e]);
                    const int o = 4 * g + e;
                    if (o < kMlpOut) put(kMlpW3 + o * 64 + 16 * mm + i, gw3[mm][e]);
                    if (i == 0) {
                        put(kMlpB1 + u, gb1[4 * mm + e]);
                        put(kMlpB2 + u, gb2[4 * mm + e]);
                    }
                }
            if (i == 0 && g == 0) {
#pragma unroll
                for (int e = 0; e < kMlpOut; ++e) put(kMlpB3 + e, gb3[e]);
                put(kMlpParams, loss);
            }
        }
        __syncthreads();
    }
    float* dst = ga.partial + size_t(blockIdx.x) * (kMlpParams + 1);
    for (int k = threadIdx.x; k < kMlpParams + 1; k += blockDim.x) dst[k] = red[k];
}

__global__ void k_mlp_reduce(const float* partial, uint32_t rows, float* out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > kMlpParams) return;
    float s = 0.0f;
    for (uint32_t r = 0; r < rows; ++r) s += partial[size_t(r) * (kMlpParams + 1) + p];
    out[p] = s;
}

__global__ void k_mlp_adam(MlpBuf m, const float* partial, AdamParams ap, float* loss_out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > kMlpParams) return;
    float gsum = 0.0f;
    for (uint32_t r = 0; r < ap.rows; ++r) gsum += partial[size_t(r) * (kMlpParams + 1) + p];
    if (p == kMlpParams) {
        if (loss_out) *loss_out = gsum * ap.loss_scale;
        return;
    }
    const float mo = ap.beta1 * m.adam_m[p] + (1.0f - ap.beta1) * gsum;
    const float ve = ap.beta2 * m.adam_v[p] + (1.0f - ap.beta2) * gsum * gsum;
    m.adam_m[p] = mo;
    m.adam_v[p] = ve;
    const float mh = mo / ap.corr1, vh = ve / ap.corr2;
    const float w = m.master[p] - ap.lr * mh / (sqrtf(vh) + ap.eps);
    m.master[p] = w;
    store_copy(m, p, w);
}

inline uint32_t cdiv(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

}  // namespace

void launch_mlp_sync(hipStream_t s, const MlpBuf& m) { k_mlp_sync<<<cdiv(kMlpParams, 256), 256, 0, s>>>(m); }

void launch_mlp_infer(hipStream_t s, const MlpBuf& m, const float4* feat, uint32_t n, float4* out) {
    if (!n) return;
    const uint32_t grid = std::min<uint32_t>(cdiv(cdiv(n, 16), 4), 512);
    k_mlp_infer<<<grid, 256, 0, s>>>(m, feat, n, out);
}

void launch_mlp_warm(hipStream_t s, const MlpBuf& m, const float4* feat, const float4* base, uint32_t np,
                     int only_miss, float4* c_imp, const float4* c_la, float4* vr_imp, float4* wr_imp,
                     const uint32_t* vel_pos, const uint32_t* ws_pos) {
    if (!np) return;
    const uint32_t grid = std::min<uint32_t>(cdiv(cdiv(np, 16), 4), 512);
    WarmArgs w{feat, base, np, only_miss, c_imp, c_la, vr_imp, wr_imp, vel_pos, ws_pos};
    k_mlp_warm<<<grid, 256, 0, s>>>(m, w);
}

void launch_mlp_collect(hipStream_t s, const float4* feat, const float4* base, const float4* c_imp, uint32_t np,
                        float4* samples, uint32_t cap, uint32_t start) {
    if (np && cap) k_mlp_collect<<<cdiv(np, 256), 256, 0, s>>>(feat, base, c_imp, np, samples, cap, start);
}

// Tiles per block chosen so that about 256 blocks share the batch.
static uint32_t grad_tpb(uint32_t batch) { return std::max<uint32_t>(4, cdiv(cdiv(batch, 16), 256)); }
uint32_t mlp_grad_blocks(uint32_t batch) { return cdiv(cdiv(batch, 16), grad_tpb(batch)); }

void launch_mlp_grad(hipStream_t s, const MlpBuf& m, const float4* samples, uint32_t n, uint32_t batch,
                     int identity, uint32_t seed, float* partial, float clip) {
    const uint32_t tpb = grad_tpb(batch);
    GradArgs ga{samples, n, batch, tpb, seed, identity, partial, clip};
    k_mlp_grad<<<mlp_grad_blocks(batch), 256, 0, s>>>(m, ga);
}

void launch_mlp_reduce(hipStream_t s, const float* partial, uint32_t rows, float* out) {
    k_mlp_reduce<<<cdiv(kMlpParams + 1, 256), 256, 0, s>>>(partial, rows, out);
}

void launch_mlp_adam(hipStream_t s, const MlpBuf& m, const float* partial, const AdamParams& ap, float* loss_out) {
    k_mlp_adam<<<cdiv(kMlpParams + 1, 256), 256, 0, s>>>(m, partial, ap, loss_out);
}

}  // namespace kz
