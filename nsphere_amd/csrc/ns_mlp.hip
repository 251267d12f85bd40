// ns_mlp.hip -- learned refinement binning: a 5 -> 64 -> 64 -> 8 MLP on
// MI355X matrix cores (BF16 operands, FP32 accumulation), trained and run by
// this library's own kernels.
//
// No reference counterpart: the reference runs every particle's adaptive
// level search from level 0 (doAdaptiveFullLeap, nsphere.c:1648-1735).  Here
// the network predicts a leapfrog particle's converged level from
//   (log r, v_r, log ell, log M_enc, log dt),
// normalised by the training set's mean and spread, and the integrator uses
// the prediction only to schedule: the bulk kernel's lanes take the ranks in
// stable order of predicted level, so that a wave holds particles of like
// work.  The level search itself still starts at level 0 for every particle,
// so a misprediction costs time, never accuracy.
//
// Matrix layout (v_mfma_f32_16x16x16_bf16, one wave): A 16x16 with lane l
// holding A[l % 16][4 (l / 16) + j], B 16x16 with lane l holding
// B[4 (l / 16) + j][l % 16], D 16x16 with lane l holding D[4 (l / 16) + r][l % 16]
// (j, r = 0..3).  Activations are kept transposed -- features x 16 samples --
// so that one layer's D tile is, converted to BF16, the next layer's B
// operand as it stands; the weight gradients (K = samples) need their
// operands sample-major and go through a 16x16 LDS transpose per tile.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "ns_device.h"
#include "ns_internal.h"
#include "ns_kernels.h"
#include "ns_physics.h"

namespace ns {

namespace {

constexpr int kIn = 5, kInPad = 16, kHid = 64, kOut = 8, kOutPad = 16;
// parameter block (FP32 master copy and gradients share this layout)
constexpr int kW1 = 0;                       // [kHid][kInPad]
constexpr int kB1 = kW1 + kHid * kInPad;     // [kHid]
constexpr int kW2 = kB1 + kHid;              // [kHid][kHid]
constexpr int kB2 = kW2 + kHid * kHid;       // [kHid]
constexpr int kW3 = kB2 + kHid;              // [kOutPad][kHid] (rows >= kOut zero)
constexpr int kB3 = kW3 + kOutPad * kHid;    // [kOutPad]
constexpr int kParams = kB3 + kOutPad;       // 6288
constexpr int kTrainThreads = 256;           // 4 waves
constexpr int kTrainBlocks = 256;            // 1024 waves

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// float -> bfloat16 bits, round to nearest even (inputs here are finite)
__host__ __device__ inline uint16_t to_bf16(float x)
{
    const uint32_t u = __builtin_bit_cast(uint32_t, x);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

__device__ inline v4f mfma(v4s a, v4s b, v4f c) { return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0); }

__device__ inline v4s pack4(float a, float b, float c, float d)
{
    v4s r;
    r[0] = (short)to_bf16(a);
    r[1] = (short)to_bf16(b);
    r[2] = (short)to_bf16(c);
    r[3] = (short)to_bf16(d);
    return r;
}

__device__ inline v4s pack(v4f x) { return pack4(x[0], x[1], x[2], x[3]); }

// BF16 weight operands of one wave, fetched from the FP32 master copy.
struct Weights {
    v4s w1[4];      // A of layer 1: W1[16 t + l%16][4 (l/16) + j]
    v4s w2[4][4];   // A of layer 2: W2[16 t + l%16][16 ks + 4 (l/16) + j]
    v4s w3[4];      // A of layer 3: W3[l%16][16 ks + 4 (l/16) + j]
    float b1[4][4], b2[4][4], b3[4];  // bias of D row 4 (l/16) + r of each tile
};

__device__ inline void load_weights(const float *__restrict__ P, Weights &W)
{
    const int lane = threadIdx.x & 63, lr = lane & 15, lq = lane >> 4;
    for (int t = 0; t < 4; t++) {
        const float *w1 = P + kW1 + (16 * t + lr) * kInPad + 4 * lq;
        W.w1[t] = pack4(w1[0], w1[1], w1[2], w1[3]);
        for (int ks = 0; ks < 4; ks++) {
            const float *w2 = P + kW2 + (16 * t + lr) * kHid + 16 * ks + 4 * lq;
            W.w2[t][ks] = pack4(w2[0], w2[1], w2[2], w2[3]);
        }
        const float *w3 = P + kW3 + lr * kHid + 16 * t + 4 * lq;
        W.w3[t] = pack4(w3[0], w3[1], w3[2], w3[3]);
        for (int r = 0; r < 4; r++) {
            W.b1[t][r] = P[kB1 + 16 * t + 4 * lq + r];
            W.b2[t][r] = P[kB2 + 16 * t + 4 * lq + r];
        }
    }
    for (int r = 0; r < 4; r++)
        W.b3[r] = P[kB3 + 4 * lq + r];
}

// Forward pass of 16 samples.  xb: this lane's B operand of layer 1 (features
// 4 (l/16) + j of sample l % 16, normalised).  h1, h2: ReLU activations
// (D layout per 16-feature tile); z: logits (D rows 0..7 valid).
__device__ inline void forward(const Weights &W, v4s xb, v4f h1[4], v4f h2[4], v4f &z)
{
    const v4f zero = {0.f, 0.f, 0.f, 0.f};
    for (int t = 0; t < 4; t++) {
        v4f a = mfma(W.w1[t], xb, zero);
        for (int r = 0; r < 4; r++)
            a[r] = fmaxf(a[r] + W.b1[t][r], 0.f);
        h1[t] = a;
    }
    v4s hb[4];
    for (int ks = 0; ks < 4; ks++)
        hb[ks] = pack(h1[ks]);
    for (int t = 0; t < 4; t++) {
        v4f a = zero;
        for (int ks = 0; ks < 4; ks++)
            a = mfma(W.w2[t][ks], hb[ks], a);
        for (int r = 0; r < 4; r++)
            a[r] = fmaxf(a[r] + W.b2[t][r], 0.f);
        h2[t] = a;
    }
    for (int ks = 0; ks < 4; ks++)
        hb[ks] = pack(h2[ks]);
    v4f a = zero;
    for (int ks = 0; ks < 4; ks++)
        a = mfma(W.w3[ks], hb[ks], a);
    for (int r = 0; r < 4; r++)
        a[r] += W.b3[r];
    z = a;
}

// The five raw features of the particle at rank i (before normalisation):
// single-precision logarithms of the values rounded to float (the network's
// inputs, not physics: training and inference both come through here).
__device__ inline void raw_features(const StepArgs &a, int64_t i, double r, double v, double ell, float f[kIn])
{
    const RankInfo k = rank_info(a, i);
    const double menc = a.menc ? a.menc[i] : ((double)k.i / (double)k.n) * k.M;
    f[0] = logf(fmaxf((float)r, 1e-30f));
    f[1] = (float)v;
    f[2] = logf(fmaxf(fabsf((float)ell), 1e-30f));
    f[3] = logf(fmaxf((float)menc, 1e-30f));
    f[4] = logf((float)a.dt);
}

// Layer-1 B operand of this lane from the 16 samples' normalised features
// held by lanes 0..15 (feature k of sample s in lane s): lanes 16..31 take
// feature 4, lanes 32..63 the zero padding.
__device__ inline v4s layer1_operand(const float f[kIn])
{
    const int lane = threadIdx.x & 63, lq = lane >> 4;
    const float x4 = __shfl(f[4], lane & 15, 64);
    if (lq == 0)
        return pack4(f[0], f[1], f[2], f[3]);
    if (lq == 1)
        return pack4(x4, 0.f, 0.f, 0.f);
    return pack4(0.f, 0.f, 0.f, 0.f);
}

// ---- inference ------------------------------------------------------------
//
// Layers 2 and 3 contract over 64 hidden features with the K = 32 BF16
// instruction (v_mfma_f32_16x16x32_bf16: A lane l holds A[l % 16][8 (l/16) + j],
// B lane l holds B[8 (l/16) + j][l % 16], j = 0..7).  The order of the
// contraction index is free as long as A and B agree, so chunk kk's index
// 8 lq + j is taken to be the hidden feature that lane group lq already holds
// in D tiles 2 kk (j < 4) and 2 kk + 1 (j >= 4): the B operand is the two
// tiles converted to BF16 as they stand, and the weights are gathered to
// match once per wave.
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef __bf16 v2bf __attribute__((ext_vector_type(2)));
typedef short v8s __attribute__((ext_vector_type(8)));
typedef short v2s __attribute__((ext_vector_type(2)));
typedef float v2f __attribute__((ext_vector_type(2)));

// Two floats to packed BF16 in one instruction (v_cvt_pk_bf16_f32, round to
// nearest even, as to_bf16); with relu, negative values (sign bit set, a
// negative 16-bit integer) become +0 by a packed 16-bit integer max -- the
// same as ReLU before rounding, since rounding keeps the sign.
template <bool RELU>
__device__ inline v2s cvt2(float a, float b)
{
    const v2f x = {a, b};
    v2s s = __builtin_bit_cast(v2s, __builtin_convertvector(x, v2bf));
    if (RELU) {
        const v2s zero = {0, 0};
        s = __builtin_elementwise_max(s, zero);
    }
    return s;
}

__device__ inline v4f mfma32(v8bf a, v8bf b, v4f c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }

__device__ inline int chunk_feature(int kk, int lq, int j) { return 32 * kk + 16 * (j >> 2) + 4 * lq + (j & 3); }

// ReLU of two D tiles (pre-activations) as one K = 32 B operand
__device__ inline v8bf relu_cat_bf16(v4f lo, v4f hi)
{
    const v2s p0 = cvt2<true>(lo[0], lo[1]), p1 = cvt2<true>(lo[2], lo[3]), p2 = cvt2<true>(hi[0], hi[1]),
              p3 = cvt2<true>(hi[2], hi[3]);
    const v8s s = {p0[0], p0[1], p1[0], p1[1], p2[0], p2[1], p3[0], p3[1]};
    return __builtin_bit_cast(v8bf, s);
}

struct InferWeights {
    v4s w1[4];       // A of layer 1 (K = 16): W1[16 t + lr][4 lq + j]
    v8bf w2[4][2];   // A of layer 2, tile t, chunk kk: W2[16 t + lr][chunk_feature(kk, lq, j)]
    v8bf w3[2];      // A of layer 3: W3[lr][chunk_feature(kk, lq, j)]
    v4f b1[4], b2[4], b3;  // biases of D rows 4 lq + r (the accumulators' initial values)
};

__device__ inline void load_infer_weights(const float *__restrict__ P, InferWeights &W)
{
    const int lane = threadIdx.x & 63, lr = lane & 15, lq = lane >> 4;
    for (int t = 0; t < 4; t++) {
        const float *w1 = P + kW1 + (16 * t + lr) * kInPad + 4 * lq;
        W.w1[t] = pack4(w1[0], w1[1], w1[2], w1[3]);
        for (int kk = 0; kk < 2; kk++) {
            v8s s;
            for (int j = 0; j < 8; j++)
                s[j] = (short)to_bf16(P[kW2 + (16 * t + lr) * kHid + chunk_feature(kk, lq, j)]);
            W.w2[t][kk] = __builtin_bit_cast(v8bf, s);
        }
        for (int r = 0; r < 4; r++) {
            W.b1[t][r] = P[kB1 + 16 * t + 4 * lq + r];
            W.b2[t][r] = P[kB2 + 16 * t + 4 * lq + r];
        }
    }
    for (int kk = 0; kk < 2; kk++) {
        v8s s;
        for (int j = 0; j < 8; j++)
            s[j] = (short)to_bf16(P[kW3 + lr * kHid + chunk_feature(kk, lq, j)]);
        W.w3[kk] = __builtin_bit_cast(v8bf, s);
    }
    for (int r = 0; r < 4; r++)
        W.b3[r] = P[kB3 + 4 * lq + r];
}

// Logits of the 16 samples of one sub-group (D layout, rows 0..7 valid).
// The pre-activations stay in the accumulators; ReLU is applied as they are
// converted to the next layer's BF16 operand.
__device__ inline v4f forward_infer(const InferWeights &W, v4s xb)
{
    v4f p1[4], p2[4];
    for (int t = 0; t < 4; t++)
        p1[t] = mfma(W.w1[t], xb, W.b1[t]);
    const v8bf a0 = relu_cat_bf16(p1[0], p1[1]), a1 = relu_cat_bf16(p1[2], p1[3]);
    for (int t = 0; t < 4; t++)
        p2[t] = mfma32(W.w2[t][1], a1, mfma32(W.w2[t][0], a0, W.b2[t]));
    const v8bf c0 = relu_cat_bf16(p2[0], p2[1]), c1 = relu_cat_bf16(p2[2], p2[3]);
    return mfma32(W.w3[1], c1, mfma32(W.w3[0], c0, W.b3));
}

// argmax over the 8 logits of sample l % 16, valid in lanes 0..15: each of
// lanes l and l + 16 takes the best of its four classes, then lane l keeps
// the upper half's only if strictly larger -- the first of equal maxima, as
// numpy's argmax.
__device__ inline int argmax8(v4f z)
{
    const int lq = (threadIdx.x & 63) >> 4;
    float best_v = z[0];
    int best = 0;
    for (int r = 1; r < 4; r++)
        if (z[r] > best_v) {
            best_v = z[r];
            best = r;
        }
    best += 4 * lq;
    const float other_v = __shfl_xor(best_v, 16, 64);
    const int other = __shfl_xor(best, 16, 64);
    return other_v > best_v ? other : best;
}

// Predicted level of every rank of a step (gathering through perm when set),
// or of every row of a raw feature array X (when X is set).  Each wave takes
// 64 ranks at a time, one per lane for the loads and the features, then runs
// them through the network as four sub-groups of 16; the next 64 ranks' loads
// are issued before the current ones are evaluated.
template <bool GATHER, typename S>
__global__ __launch_bounds__(256) void k_mlp_predict(StepArgs a, const float *__restrict__ X,
                                                      const float *__restrict__ P, const float *__restrict__ norm,
                                                      uint32_t *__restrict__ pred)
{
    const int lane = threadIdx.x & 63, lr = lane & 15, lq = lane >> 4;
    InferWeights W;
    load_infer_weights(P, W);
    float mean[kIn], inv[kIn];
    for (int k = 0; k < kIn; k++) {
        mean[k] = norm[k];
        inv[k] = norm[kIn + k];
    }
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t waves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    const S *r0 = static_cast<const S *>(a.r0), *v0 = static_cast<const S *>(a.v0),
            *l0 = static_cast<const S *>(a.l0);
    // the next 64 ranks' inputs: the state at its own precision, or the five
    // raw features
    int64_t base = wave * 64;
    S nr = 0, nv = 0, nl = 0;
    float nx[kIn] = {0.f, 0.f, 0.f, 0.f, 0.f};
    auto prefetch = [&](int64_t b) {
        const int64_t i = b + lane;
        if (i >= a.n)
            return;
        if (X) {
            for (int k = 0; k < kIn; k++)
                nx[k] = X[i * kIn + k];
        } else {
            const int64_t src = GATHER ? (int64_t)a.perm[i] : i;
            nr = r0[src];
            nv = v0[src];
            nl = l0[src];
        }
    };
    if (base < a.n)
        prefetch(base);
    for (; base < a.n; base += waves * 64) {
        const int64_t i = base + lane;
        const S cr = nr, cv = nv, cl = nl;
        float f[kIn];
        for (int k = 0; k < kIn; k++)
            f[k] = nx[k];
        if (base + waves * 64 < a.n)
            prefetch(base + waves * 64);
        if (i < a.n) {
            if (!X)
                raw_features(a, i, (double)cr, (double)cv, (double)cl, f);
            for (int k = 0; k < kIn; k++)
                f[k] = (f[k] - mean[k]) * inv[k];
        } else {
            for (int k = 0; k < kIn; k++)
                f[k] = 0.f;
        }
        uint32_t mine = 0;
        for (int q = 0; q < 4; q++) {
            if (base + 16 * q >= a.n)  // wave-uniform
                break;
            const int src = 16 * q + lr;
            float g[kIn];
            for (int k = 0; k < kIn; k++)
                g[k] = __shfl(f[k], src, 64);
            const float x0 = lq == 0 ? g[0] : (lq == 1 ? g[4] : 0.f), x1 = lq == 0 ? g[1] : 0.f,
                        x2 = lq == 0 ? g[2] : 0.f, x3 = lq == 0 ? g[3] : 0.f;
            const v2s s01 = cvt2<false>(x0, x1), s23 = cvt2<false>(x2, x3);
            const v4s xb = {s01[0], s01[1], s23[0], s23[1]};
            const int c = argmax8(forward_infer(W, xb));
            const int cq = __shfl(c, lr, 64);
            if (lq == q)
                mine = (uint32_t)cq;
        }
        if (i < a.n)
            pred[i] = mine;
    }
}

// Raw features of the current state in rank order (training data).
template <typename S>
__global__ void k_mlp_features(StepArgs a, float *__restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n)
        return;
    float f[kIn];
    raw_features(a, i, (double)static_cast<const S *>(a.r0)[i], (double)static_cast<const S *>(a.v0)[i],
                 (double)static_cast<const S *>(a.l0)[i], f);
    for (int k = 0; k < kIn; k++)
        out[i * kIn + k] = f[k];
}

// 16x16 tile transpose through LDS: from D layout (lane l holds rows
// 4 (l/16) + r of column l % 16) to an A operand (lane l: row l % 16,
// columns 4 (l/16) + j), in BF16.  One wave, its own scratch.
__device__ inline v4s transpose_to_a(v4f d, float *tile)
{
    const int lane = threadIdx.x & 63, lr = lane & 15, lq = lane >> 4;
    for (int r = 0; r < 4; r++)
        tile[(4 * lq + r) * 17 + lr] = d[r];
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    const v4s out = pack4(tile[lr * 17 + 4 * lq], tile[lr * 17 + 4 * lq + 1], tile[lr * 17 + 4 * lq + 2],
                          tile[lr * 17 + 4 * lq + 3]);
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    return out;
}

// The same tile as a B operand with the roles of rows and columns swapped:
// B[k][n] = tile[n][k], i.e. lane l gets D rows l % 16, columns 4 (l/16) + j
// -- which is transpose_to_a again.
__device__ inline v4s transpose_to_b(v4f d, float *tile) { return transpose_to_a(d, tile); }

// Gradients of the mean (class-weighted) cross-entropy over `count` samples
// starting at `first`: every wave takes 16 samples at a time and accumulates
// its weight gradients in MFMA accumulators, then writes them to its own row
// of `partial` (kParams floats); loss and weight sums go to partial's tail.
__global__ __launch_bounds__(kTrainThreads) void k_mlp_grad(const float *__restrict__ P,
                                                            const float *__restrict__ norm,
                                                            const float *__restrict__ X, const uint8_t *__restrict__ y,
                                                            const float *__restrict__ cw, int64_t first,
                                                            int64_t count, float scale, float *__restrict__ partial)
{
    __shared__ float s_tile[kTrainThreads / 64][16 * 17];
    const int lane = threadIdx.x & 63, lr = lane & 15, lq = lane >> 4, wv = threadIdx.x >> 6;
    float *tile = s_tile[wv];
    Weights W;
    load_weights(P, W);
    // transposed weight operands: W3^T (layer-3 backward) and W2^T (layer 2)
    v4s w3t[4], w2t[4][4];
    for (int t = 0; t < 4; t++) {
        w3t[t] = pack4(P[kW3 + (4 * lq + 0) * kHid + 16 * t + lr], P[kW3 + (4 * lq + 1) * kHid + 16 * t + lr],
                       P[kW3 + (4 * lq + 2) * kHid + 16 * t + lr], P[kW3 + (4 * lq + 3) * kHid + 16 * t + lr]);
        for (int ks = 0; ks < 4; ks++) {
            const float *c = P + kW2 + (16 * ks + 4 * lq) * kHid + 16 * t + lr;
            w2t[t][ks] = pack4(c[0], c[kHid], c[2 * kHid], c[3 * kHid]);
        }
    }
    float mean[kIn], inv[kIn];
    for (int k = 0; k < kIn; k++) {
        mean[k] = norm[k];
        inv[k] = norm[kIn + k];
    }
    const v4f zero = {0.f, 0.f, 0.f, 0.f};
    v4f g1[4], g2[4][4], g3[4];  // dW1 (f1 tile x inputs), dW2 (f2 x f1 tiles), dW3 (classes x f2 tile)
    float gb1[4][4] = {}, gb2[4][4] = {}, gb3[4] = {};
    for (int t = 0; t < 4; t++) {
        g1[t] = zero;
        g3[t] = zero;
        for (int u = 0; u < 4; u++)
            g2[t][u] = zero;
    }
    float loss = 0.f, wsum = 0.f;
    const int64_t groups = (count + 15) / 16;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t waves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t g = wave; g < groups; g += waves) {
        const int64_t s_idx = g * 16 + lr;  // this lane's sample column
        const bool valid = s_idx < count;
        float f[kIn] = {0.f, 0.f, 0.f, 0.f, 0.f};
        int label = 0;
        float wgt = 0.f;
        if (valid) {
            for (int k = 0; k < kIn; k++)
                f[k] = (X[(first + s_idx) * kIn + k] - mean[k]) * inv[k];
            label = y[first + s_idx];
            wgt = cw[label];
        }
        const v4s xb = layer1_operand(f);
        v4f h1[4], h2[4], z;
        forward(W, xb, h1, h2, z);
        // softmax over classes 0..7 of sample lr: lanes lr (0-3) and lr + 16 (4-7)
        float m = -1e30f;
        for (int r = 0; r < 4; r++)
            m = fmaxf(m, z[r]);
        const float mo = __shfl_xor(m, 16, 64);
        m = lq < 2 ? fmaxf(m, mo) : 0.f;
        float e[4], sum = 0.f;
        for (int r = 0; r < 4; r++) {
            e[r] = lq < 2 ? expf(z[r] - m) : 0.f;
            sum += e[r];
        }
        sum += __shfl_xor(sum, 16, 64);
        v4f dz;
        for (int r = 0; r < 4; r++) {
            const int c = 4 * lq + r;
            const float pr = lq < 2 ? e[r] / sum : 0.f;
            dz[r] = lq < 2 && valid ? wgt * scale * (pr - (c == label ? 1.f : 0.f)) : 0.f;
            if (lq < 2 && valid && c == label) {
                loss -= wgt * logf(fmaxf(pr, 1e-30f));
                wsum += wgt;
            }
        }
        // dW3 += dZ . H2^T ; db3 += sum_s dZ
        const v4s dz_a = transpose_to_a(dz, tile);
        for (int t = 0; t < 4; t++)
            g3[t] = mfma(dz_a, transpose_to_b(h2[t], tile), g3[t]);
        for (int r = 0; r < 4; r++)
            gb3[r] += dz[r];
        // dH2 = W3^T . dZ (masked by ReLU)
        const v4s dz_b = pack(dz);
        v4f dh2[4];
        for (int t = 0; t < 4; t++) {
            dh2[t] = mfma(w3t[t], dz_b, zero);
            for (int r = 0; r < 4; r++) {
                dh2[t][r] = h2[t][r] > 0.f ? dh2[t][r] : 0.f;
                gb2[t][r] += dh2[t][r];
            }
        }
        // dW2 += dH2 . H1^T
        v4s h1b[4];
        for (int u = 0; u < 4; u++)
            h1b[u] = transpose_to_b(h1[u], tile);
        for (int t = 0; t < 4; t++) {
            const v4s a2 = transpose_to_a(dh2[t], tile);
            for (int u = 0; u < 4; u++)
                g2[t][u] = mfma(a2, h1b[u], g2[t][u]);
        }
        // dH1 = W2^T . dH2 (masked by ReLU)
        v4s dh2b[4];
        for (int ks = 0; ks < 4; ks++)
            dh2b[ks] = pack(dh2[ks]);
        v4f dh1[4];
        for (int t = 0; t < 4; t++) {
            v4f acc = zero;
            for (int ks = 0; ks < 4; ks++)
                acc = mfma(w2t[t][ks], dh2b[ks], acc);
            for (int r = 0; r < 4; r++) {
                acc[r] = h1[t][r] > 0.f ? acc[r] : 0.f;
                gb1[t][r] += acc[r];
            }
            dh1[t] = acc;
        }
        // dW1 += dH1 . X: B[k = sample][n = input] from this lane's inputs
        // (lane l: samples 4 (l/16) + j, input l % 16)
        v4f xin;
        for (int j = 0; j < 4; j++) {
            const int src = 4 * lq + j;  // sample column held by lane src
            float x = 0.f;
            for (int k = 0; k < kIn; k++) {
                const float fk = __shfl(f[k], src, 64);
                if (lr == k)
                    x = fk;
            }
            xin[j] = x;
        }
        const v4s xbk = pack(xin);
        for (int t = 0; t < 4; t++)
            g1[t] = mfma(transpose_to_a(dh1[t], tile), xbk, g1[t]);
    }
    // bias gradients: D rows are features, columns samples -> sum the 16
    // sample lanes of each row group
    for (int off = 1; off < 16; off <<= 1) {
        for (int t = 0; t < 4; t++)
            for (int r = 0; r < 4; r++) {
                gb1[t][r] += __shfl_xor(gb1[t][r], off, 64);
                gb2[t][r] += __shfl_xor(gb2[t][r], off, 64);
            }
        for (int r = 0; r < 4; r++)
            gb3[r] += __shfl_xor(gb3[r], off, 64);
    }
    for (int off = 1; off < 64; off <<= 1) {
        loss += __shfl_xor(loss, off, 64);
        wsum += __shfl_xor(wsum, off, 64);
    }
    float *out = partial + wave * (kParams + 2);
    for (int t = 0; t < 4; t++)
        for (int r = 0; r < 4; r++) {
            const int row = 4 * lq + r;
            out[kW1 + (16 * t + row) * kInPad + lr] = g1[t][r];
            for (int u = 0; u < 4; u++)
                out[kW2 + (16 * t + row) * kHid + 16 * u + lr] = g2[t][u][r];
            out[kW3 + row * kHid + 16 * t + lr] = g3[t][r];
            if (lr == 0) {
                out[kB1 + 16 * t + row] = gb1[t][r];
                out[kB2 + 16 * t + row] = gb2[t][r];
            }
        }
    if (lr == 0)
        for (int r = 0; r < 4; r++)
            out[kB3 + 4 * lq + r] = gb3[r];
    if (lane == 0) {
        out[kParams] = loss;
        out[kParams + 1] = wsum;
    }
}

// Sum of the waves' partial gradients (one thread per parameter).
__global__ void k_mlp_reduce(const float *__restrict__ partial, int nwaves, float *__restrict__ grad)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kParams + 2)
        return;
    double s = 0.0;
    for (int w = 0; w < nwaves; w++)
        s += partial[(int64_t)w * (kParams + 2) + p];
    grad[p] = (float)s;
}

// Adam on the FP32 master copy; padded weights stay zero.
__global__ void k_mlp_adam(float *__restrict__ P, const float *__restrict__ grad, float *__restrict__ m1,
                           float *__restrict__ m2, float lr, float b1, float b2, float eps, float c1, float c2)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kParams)
        return;
    const bool pad = (p < kB1 && (p % kInPad) >= kIn) || (p >= kW3 && p < kB3 && (p - kW3) / kHid >= kOut) ||
                     (p >= kB3 + kOut);
    if (pad)
        return;
    const float g = grad[p];
    m1[p] = b1 * m1[p] + (1.f - b1) * g;
    m2[p] = b2 * m2[p] + (1.f - b2) * g * g;
    P[p] -= lr * (m1[p] / c1) / (sqrtf(m2[p] / c2) + eps);
}

} // namespace

constexpr int kPredictBlocks = 2048;

hipError_t launch_mlp_predict(const StepArgs &a, bool gather, bool f32, const float *P, const float *norm,
                              uint32_t *pred, hipStream_t stream)
{
    if (a.n <= 0)
        return hipSuccess;
    if (f32) {
        if (gather)
            hipLaunchKernelGGL((k_mlp_predict<true, float>), dim3(kPredictBlocks), dim3(256), 0, stream, a, nullptr, P,
                               norm, pred);
        else
            hipLaunchKernelGGL((k_mlp_predict<false, float>), dim3(kPredictBlocks), dim3(256), 0, stream, a, nullptr,
                               P, norm, pred);
    } else {
        if (gather)
            hipLaunchKernelGGL((k_mlp_predict<true, double>), dim3(kPredictBlocks), dim3(256), 0, stream, a, nullptr,
                               P, norm, pred);
        else
            hipLaunchKernelGGL((k_mlp_predict<false, double>), dim3(kPredictBlocks), dim3(256), 0, stream, a,
                               nullptr, P, norm, pred);
    }
    return hipGetLastError();
}

hipError_t launch_mlp_features(const StepArgs &a, bool f32, float *out, hipStream_t stream)
{
    if (a.n <= 0)
        return hipSuccess;
    const dim3 grid((unsigned)((a.n + 255) / 256));
    if (f32)
        hipLaunchKernelGGL(k_mlp_features<float>, grid, dim3(256), 0, stream, a, out);
    else
        hipLaunchKernelGGL(k_mlp_features<double>, grid, dim3(256), 0, stream, a, out);
    return hipGetLastError();
}

} // namespace ns

// ------------------------------------------------------------------ C-ABI

struct ns_mlp {
    int device = 0;
    float *P = nullptr, *m1 = nullptr, *m2 = nullptr, *grad = nullptr, *partial = nullptr;
    float *norm = nullptr;  // mean[5], 1/std[5]
    float *cw = nullptr;    // class weights [8]
    int64_t steps = 0;      // Adam steps taken
    float *X = nullptr;
    uint8_t *y = nullptr;
    int64_t cap = 0;        // samples the device data buffers hold
    hipStream_t stream = nullptr;
};

namespace {

#define MLP_TRY(call)                                                                                       \
    do {                                                                                                    \
        hipError_t e_ = (call);                                                                             \
        if (e_ != hipSuccess)                                                                               \
            return ns::set_error(NS_ERR_HIP, "%s: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__,     \
                                 __LINE__);                                                                 \
    } while (0)

int mlp_upload(ns_mlp *m, const float *X, const uint8_t *y, int64_t n)
{
    if (n > m->cap) {
        if (m->X)
            (void)hipFree(m->X);
        if (m->y)
            (void)hipFree(m->y);
        m->X = nullptr;
        m->y = nullptr;
        m->cap = 0;
        MLP_TRY(hipMalloc((void **)&m->X, (size_t)n * ns::kIn * sizeof(float)));
        MLP_TRY(hipMalloc((void **)&m->y, (size_t)n));
        m->cap = n;
    }
    MLP_TRY(hipMemcpy(m->X, X, (size_t)n * ns::kIn * sizeof(float), hipMemcpyHostToDevice));
    MLP_TRY(hipMemcpy(m->y, y, (size_t)n, hipMemcpyHostToDevice));
    return NS_OK;
}

// gradient of the batch [first, first + count) into m->grad (+ loss, weight)
int mlp_grad(ns_mlp *m, int64_t first, int64_t count, float scale)
{
    // at least kGroupsPerWave groups of 16 samples per wave, so that the
    // partial gradients (one row per wave) stay small against the batch
    constexpr int64_t kGroupsPerWave = 8, kWavesPerBlock = ns::kTrainThreads / 64;
    const int64_t groups = (count + 15) / 16;
    int64_t blocks = (groups + kGroupsPerWave * kWavesPerBlock - 1) / (kGroupsPerWave * kWavesPerBlock);
    blocks = blocks < 1 ? 1 : (blocks > ns::kTrainBlocks ? ns::kTrainBlocks : blocks);
    const int waves = (int)(blocks * kWavesPerBlock);
    hipLaunchKernelGGL(ns::k_mlp_grad, dim3((unsigned)blocks), dim3(ns::kTrainThreads), 0, m->stream, m->P, m->norm,
                       m->X, m->y, m->cw, first, count, scale, m->partial);
    MLP_TRY(hipGetLastError());
    hipLaunchKernelGGL(ns::k_mlp_reduce, dim3((ns::kParams + 2 + 255) / 256), dim3(256), 0, m->stream, m->partial,
                       waves, m->grad);
    MLP_TRY(hipGetLastError());
    return NS_OK;
}

} // namespace

const float *ns::mlp_params(const ns_mlp *m) { return m->P; }
const float *ns::mlp_norm(const ns_mlp *m) { return m->norm; }
int ns::mlp_device(const ns_mlp *m) { return m->device; }

extern "C" {

int ns_mlp_create(int device, uint64_t seed, ns_mlp **out)
{
    if (!out)
        return ns::set_error(NS_ERR_ARG, "ns_mlp_create: null out");
    *out = nullptr;
    ns_mlp *m = new (std::nothrow) ns_mlp;
    if (!m)
        return ns::set_error(NS_ERR_NOMEM, "ns_mlp_create");
    m->device = device;
    if (hipSetDevice(device) != hipSuccess) {
        delete m;
        return ns::set_error(NS_ERR_NODEV, "ns_mlp_create: device %d", device);
    }
    const int waves = ns::kTrainBlocks * (ns::kTrainThreads / 64);
    hipError_t e = hipMalloc((void **)&m->P, ns::kParams * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void **)&m->m1, ns::kParams * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void **)&m->m2, ns::kParams * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void **)&m->grad, (ns::kParams + 2) * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void **)&m->partial, (size_t)waves * (ns::kParams + 2) * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void **)&m->norm, 2 * ns::kIn * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void **)&m->cw, ns::kOut * sizeof(float));
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        ns_mlp_destroy(m);
        return ns::set_error(NS_ERR_NOMEM, "ns_mlp_create: %s", hipGetErrorString(e));
    }
    // He-uniform initial weights from a splitmix64 stream, zero biases and
    // padding; identity normalisation and unit class weights
    std::vector<float> P(ns::kParams, 0.f);
    uint64_t s = seed;
    auto uni = [&]() {
        s += 0x9E3779B97F4A7C15ull;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
    };
    for (int o = 0; o < ns::kHid; o++)
        for (int i = 0; i < ns::kIn; i++)
            P[ns::kW1 + o * ns::kInPad + i] = (float)(uni() * std::sqrt(6.0 / ns::kIn));
    for (int o = 0; o < ns::kHid; o++)
        for (int i = 0; i < ns::kHid; i++)
            P[ns::kW2 + o * ns::kHid + i] = (float)(uni() * std::sqrt(6.0 / ns::kHid));
    for (int o = 0; o < ns::kOut; o++)
        for (int i = 0; i < ns::kHid; i++)
            P[ns::kW3 + o * ns::kHid + i] = (float)(uni() * std::sqrt(6.0 / ns::kHid));
    float norm[2 * ns::kIn], cw[ns::kOut];
    for (int k = 0; k < ns::kIn; k++) {
        norm[k] = 0.f;
        norm[ns::kIn + k] = 1.f;
    }
    for (int c = 0; c < ns::kOut; c++)
        cw[c] = 1.f;
    e = hipMemcpy(m->P, P.data(), ns::kParams * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(m->m1, 0, ns::kParams * sizeof(float));
    if (e == hipSuccess) e = hipMemset(m->m2, 0, ns::kParams * sizeof(float));
    if (e == hipSuccess) e = hipMemcpy(m->norm, norm, sizeof norm, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(m->cw, cw, sizeof cw, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        ns_mlp_destroy(m);
        return ns::set_error(NS_ERR_HIP, "ns_mlp_create: %s", hipGetErrorString(e));
    }
    *out = m;
    return NS_OK;
}

int ns_mlp_destroy(ns_mlp *m)
{
    if (!m)
        return NS_OK;
    (void)hipSetDevice(m->device);
    if (m->stream)
        (void)hipStreamSynchronize(m->stream);
    for (void *p : {(void *)m->P, (void *)m->m1, (void *)m->m2, (void *)m->grad, (void *)m->partial,
                    (void *)m->norm, (void *)m->cw, (void *)m->X, (void *)m->y})
        if (p)
            (void)hipFree(p);
    if (m->stream)
        (void)hipStreamDestroy(m->stream);
    delete m;
    return NS_OK;
}

int ns_mlp_get_params(ns_mlp *m, float *params, float *norm, float *class_weight)
{
    if (!m)
        return ns::set_error(NS_ERR_ARG, "ns_mlp_get_params: null network");
    (void)hipSetDevice(m->device);
    MLP_TRY(hipStreamSynchronize(m->stream));
    if (params)
        MLP_TRY(hipMemcpy(params, m->P, ns::kParams * sizeof(float), hipMemcpyDeviceToHost));
    if (norm)
        MLP_TRY(hipMemcpy(norm, m->norm, 2 * ns::kIn * sizeof(float), hipMemcpyDeviceToHost));
    if (class_weight)
        MLP_TRY(hipMemcpy(class_weight, m->cw, ns::kOut * sizeof(float), hipMemcpyDeviceToHost));
    return NS_OK;
}

int ns_mlp_set_params(ns_mlp *m, const float *params, const float *norm, const float *class_weight)
{
    if (!m)
        return ns::set_error(NS_ERR_ARG, "ns_mlp_set_params: null network");
    (void)hipSetDevice(m->device);
    MLP_TRY(hipStreamSynchronize(m->stream));
    if (params)
        MLP_TRY(hipMemcpy(m->P, params, ns::kParams * sizeof(float), hipMemcpyHostToDevice));
    if (norm)
        MLP_TRY(hipMemcpy(m->norm, norm, 2 * ns::kIn * sizeof(float), hipMemcpyHostToDevice));
    if (class_weight)
        MLP_TRY(hipMemcpy(m->cw, class_weight, ns::kOut * sizeof(float), hipMemcpyHostToDevice));
    return NS_OK;
}

int ns_mlp_gradient(ns_mlp *m, const float *X, const uint8_t *y, int64_t n, float *grad, float *loss)
{
    if (!m || !X || !y || !grad || n <= 0)
        return ns::set_error(NS_ERR_ARG, "ns_mlp_gradient: bad argument");
    for (int64_t i = 0; i < n; i++)
        if (y[i] >= ns::kOut)
            return ns::set_error(NS_ERR_ARG, "ns_mlp_gradient: label %d at %lld outside 0..7", (int)y[i],
                                 (long long)i);
    (void)hipSetDevice(m->device);
    if (int rc = mlp_upload(m, X, y, n))
        return rc;
    // class-weighted mean: scale by 1 / sum of the batch's weights
    std::vector<float> cw(ns::kOut);
    MLP_TRY(hipMemcpy(cw.data(), m->cw, ns::kOut * sizeof(float), hipMemcpyDeviceToHost));
    double ws = 0.0;
    for (int64_t i = 0; i < n; i++)
        ws += y[i] < ns::kOut ? cw[y[i]] : 0.0;
    if (int rc = mlp_grad(m, 0, n, (float)(1.0 / ws)))
        return rc;
    std::vector<float> g(ns::kParams + 2);
    MLP_TRY(hipMemcpyAsync(g.data(), m->grad, g.size() * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    MLP_TRY(hipStreamSynchronize(m->stream));
    std::memcpy(grad, g.data(), ns::kParams * sizeof(float));
    if (loss)
        *loss = (float)(g[ns::kParams] / g[ns::kParams + 1]);
    return NS_OK;
}

int ns_mlp_train(ns_mlp *m, const float *X, const uint8_t *y, int64_t n, int epochs, int64_t batch, float lr,
                 float *loss_out)
{
    if (!m || !X || !y || n <= 0 || epochs < 1 || batch < 16 || !(lr > 0.f))
        return ns::set_error(NS_ERR_ARG, "ns_mlp_train: bad argument");
    for (int64_t i = 0; i < n; i++)
        if (y[i] >= ns::kOut)
            return ns::set_error(NS_ERR_ARG, "ns_mlp_train: label %d at %lld outside 0..7", (int)y[i], (long long)i);
    (void)hipSetDevice(m->device);
    // normalisation from the training set; class weights (1 / frequency)^(1/2),
    // normalised to mean 1 over the classes present
    double mean[ns::kIn] = {}, sq[ns::kIn] = {};
    int64_t cnt[ns::kOut] = {};
    for (int64_t i = 0; i < n; i++) {
        for (int k = 0; k < ns::kIn; k++) {
            mean[k] += X[i * ns::kIn + k];
            sq[k] += (double)X[i * ns::kIn + k] * X[i * ns::kIn + k];
        }
        cnt[y[i]]++;
    }
    float norm[2 * ns::kIn], cw[ns::kOut];
    for (int k = 0; k < ns::kIn; k++) {
        const double mu = mean[k] / n, var = sq[k] / n - mu * mu;
        norm[k] = (float)mu;
        norm[ns::kIn + k] = (float)(var > 1e-20 ? 1.0 / std::sqrt(var) : 1.0);
    }
    double wsum = 0.0;
    int present = 0;
    for (int c = 0; c < ns::kOut; c++)
        if (cnt[c]) {
            cw[c] = (float)std::sqrt((double)n / cnt[c]);
            wsum += cw[c];
            present++;
        } else {
            cw[c] = 0.f;
        }
    for (int c = 0; c < ns::kOut; c++)
        cw[c] = (float)(cw[c] * present / wsum);
    MLP_TRY(hipMemcpy(m->norm, norm, sizeof norm, hipMemcpyHostToDevice));
    MLP_TRY(hipMemcpy(m->cw, cw, sizeof cw, hipMemcpyHostToDevice));
    if (int rc = mlp_upload(m, X, y, n))
        return rc;
    // per-batch weight sums (the loss is the class-weighted mean per batch)
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    double last = 0.0;
    for (int ep = 0; ep < epochs; ep++) {
        double lsum = 0.0, wtot = 0.0;
        for (int64_t first = 0; first < n; first += batch) {
            const int64_t count = first + batch <= n ? batch : n - first;
            double ws = 0.0;
            for (int64_t i = first; i < first + count; i++)
                ws += cw[y[i]];
            if (int rc = mlp_grad(m, first, count, (float)(1.0 / ws)))
                return rc;
            m->steps++;
            const float c1 = 1.f - std::pow(b1, (float)m->steps), c2 = 1.f - std::pow(b2, (float)m->steps);
            hipLaunchKernelGGL(ns::k_mlp_adam, dim3((ns::kParams + 255) / 256), dim3(256), 0, m->stream, m->P,
                               m->grad, m->m1, m->m2, lr, b1, b2, eps, c1, c2);
            MLP_TRY(hipGetLastError());
            if (first + batch >= n || ((first / batch) % 64) == 0) {
                float tail[2];
                MLP_TRY(hipMemcpyAsync(tail, m->grad + ns::kParams, sizeof tail, hipMemcpyDeviceToHost,
                                       m->stream));
                MLP_TRY(hipStreamSynchronize(m->stream));
                lsum += tail[0];
                wtot += tail[1];
            }
        }
        last = wtot > 0 ? lsum / wtot : 0.0;
    }
    MLP_TRY(hipStreamSynchronize(m->stream));
    if (loss_out)
        *loss_out = (float)last;
    return NS_OK;
}

int ns_mlp_predict(ns_mlp *m, const float *X, int64_t n, uint8_t *level)
{
    if (!m || !X || !level || n <= 0)
        return ns::set_error(NS_ERR_ARG, "ns_mlp_predict: bad argument");
    (void)hipSetDevice(m->device);
    std::vector<uint8_t> zeros((size_t)n, 0);
    if (int rc = mlp_upload(m, X, zeros.data(), n))
        return rc;
    uint32_t *d_pred = nullptr;
    MLP_TRY(hipMalloc((void **)&d_pred, (size_t)n * sizeof(uint32_t)));
    ns::StepArgs a{};
    a.n = n;
    hipLaunchKernelGGL((ns::k_mlp_predict<false, double>), dim3(ns::kPredictBlocks), dim3(256), 0, m->stream, a, m->X,
                       m->P, m->norm, d_pred);
    hipError_t e = hipGetLastError();
    std::vector<uint32_t> h((size_t)n);
    if (e == hipSuccess)
        e = hipMemcpyAsync(h.data(), d_pred, (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToHost, m->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(m->stream);
    (void)hipFree(d_pred);
    if (e != hipSuccess)
        return ns::set_error(NS_ERR_HIP, "ns_mlp_predict: %s", hipGetErrorString(e));
    for (int64_t i = 0; i < n; i++)
        level[i] = (uint8_t)h[(size_t)i];
    return NS_OK;
}

} // extern "C"
