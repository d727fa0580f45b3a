// Learned radiance cache: fused MLP inference and training on BF16 MFMA (gfx950
// v_mfma_f32_16x16x32_bf16, FP32 accumulation) and an FP32 Adam update.
//
// A 32-sample tile goes through all six layers with its activations in LDS as bf16 (sample-major
// rows): one wave per tile for queries, the four waves of a block per tile for training. Every
// product is a 16x16x32 MFMA tile:
//   forward       Y[s][o]  = sum_k X[s][k] W[o][k]      A = X rows, B = W rows
//   input grad    G'[s][k] = sum_o G[s][o] W[o][k]      A = G rows, B = W^T rows
//   weight grad   dW[o][k] = sum_s G[s][o] X[s][k]      A = G columns, B = X columns
// Operand lane layout of the 16x16x32 bf16 MFMA (lane l, r = l % 16, g = l / 16): A supplies
// row r, k = 8g..8g+7; B supplies column r, k = 8g..8g+7; D holds column r, rows 4g..4g+3.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "cache.h"
#include "common.h"
#include "shading.h"

namespace ptk {

namespace {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

constexpr int WS = 32;       // samples per wave
// Rows of per-block partial gradients: a training launch has at most this many blocks, each
// sums its tiles' weight and bias gradients into its own row (no atomics), and the Adam step
// adds the rows in a fixed order (the gradient of a batch is the same on every run).
constexpr int PARTIAL_ROWS = 512;
constexpr int ST = 104;      // LDS row stride (bf16 elements) of every activation buffer
constexpr MlpLayout LAYOUT{};
constexpr int PSTRIDE = (LAYOUT.count + 63) / 64 * 64;  // floats per partial row

__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7F800000u) == 0x7F800000u)
        return uint16_t(u >> 16);  // inf / nan pass through
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
    return uint16_t(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }

__device__ __forceinline__ bf16x8_t load8(const uint16_t* p)  // 16-byte aligned
{
    return __builtin_bit_cast(bf16x8_t, *reinterpret_cast<const uint4*>(p));
}

__device__ __forceinline__ bf16x8_t gather8(const uint16_t* p, int stride)
{
    uint16_t v[8];
#pragma unroll
    for (int j = 0; j < 8; j++)
        v[j] = p[j * stride];
    uint4 u;
    u.x = uint32_t(v[0]) | (uint32_t(v[1]) << 16);
    u.y = uint32_t(v[2]) | (uint32_t(v[3]) << 16);
    u.z = uint32_t(v[4]) | (uint32_t(v[5]) << 16);
    u.w = uint32_t(v[6]) | (uint32_t(v[7]) << 16);
    return __builtin_bit_cast(bf16x8_t, u);
}

__device__ __forceinline__ f32x4_t mfma(bf16x8_t a, bf16x8_t b, f32x4_t c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Input encoding (96 features): position normalised to the scene bounds and its sines and
// cosines at 12 octaves (72), the position itself (3), direction (3), normal (3),
// roughness (1), albedo (3); the last 11 are zero.
__device__ __forceinline__ void encode_row(const float* rec, const float* lo, const float* inv_ext, uint16_t* row)
{
    float p[3];
    for (int c = 0; c < 3; c++)
        p[c] = fminf(fmaxf((rec[c] - lo[c]) * inv_ext[c], 0.0f), 1.0f);
    int f = 0;
    for (int b = 0; b < 12; b++) {
        const float s = float(1 << b) * 3.14159265f;
        for (int c = 0; c < 3; c++) {
            float sn, cs;
            sincosf(s * p[c], &sn, &cs);
            row[f++] = f2bf(sn);
            row[f++] = f2bf(cs);
        }
    }
    for (int c = 0; c < 3; c++)
        row[f++] = f2bf(p[c]);
    for (int c = 3; c < 13; c++)
        row[f++] = f2bf(rec[c]);
    while (f < MLP_IN)
        row[f++] = 0;
}

// Part w (0..3) of the encoded features of one record, as encode_row computes them: the bands
// 3w..3w+2 (features 18w..18w+17), and for w = 0 also features 72..95.
__device__ __forceinline__ void encode_part(const float* rec, const float* lo, const float* inv_ext, uint16_t* row,
                                            int w)
{
    float p[3];
    for (int c = 0; c < 3; c++)
        p[c] = fminf(fmaxf((rec[c] - lo[c]) * inv_ext[c], 0.0f), 1.0f);
    for (int b = 3 * w; b < 3 * w + 3; b++) {
        const float s = float(1 << b) * 3.14159265f;
        for (int c = 0; c < 3; c++) {
            float sn, cs;
            sincosf(s * p[c], &sn, &cs);
            row[6 * b + 2 * c] = f2bf(sn);
            row[6 * b + 2 * c + 1] = f2bf(cs);
        }
    }
    if (w == 0) {
        int f = 72;
        for (int c = 0; c < 3; c++)
            row[f++] = f2bf(p[c]);
        for (int c = 3; c < 13; c++)
            row[f++] = f2bf(rec[c]);
        while (f < MLP_IN)
            row[f++] = 0;
    }
}

// LDS buffer holding the input of layer l: one buffer per layer (training keeps them all for
// the backward pass) or two alternating buffers (inference).
template <bool PING>
__device__ __forceinline__ uint16_t* act_buf(uint16_t* base, int l)
{
    return base + (PING ? (l & 1) : l) * WS * ST;
}

// One layer of the forward pass over a 32-sample tile: Y = relu(X W^T + b) into LDS, or, for
// the last layer (LAST), X W^T + b into y (lane holds rows 4g..4g+3 of tile mt for output r).
// The wave's weight fragments (global, L2-resident) and activation fragments (LDS) are all
// loaded before the first product, so the layer waits once for their latency.
template <int K, int NO, bool LAST>
__device__ __forceinline__ void forward_layer(const uint16_t* X, uint16_t* Y, const uint16_t* W, const float* bias,
                                              int nv, f32x4_t y[2])
{
    const int lane = __lane_id(), r = lane & 15, g = lane >> 4;
    constexpr int KC = K / 32, NT = NO / 16;
    bf16x8_t bw[NT][KC], ax[2][KC];
    float bo[NT];
#pragma unroll
    for (int nt = 0; nt < NT; nt++) {
#pragma unroll
        for (int kc = 0; kc < KC; kc++)
            bw[nt][kc] = load8(W + (nt * 16 + r) * K + kc * 32 + 8 * g);
        const int o = nt * 16 + r;
        bo[nt] = o < nv ? bias[o] : 0.0f;
    }
#pragma unroll
    for (int mt = 0; mt < 2; mt++)
#pragma unroll
        for (int kc = 0; kc < KC; kc++)
            ax[mt][kc] = load8(X + (mt * 16 + r) * ST + kc * 32 + 8 * g);
#pragma unroll
    for (int nt = 0; nt < NT; nt++)
#pragma unroll
        for (int mt = 0; mt < 2; mt++) {
            f32x4_t acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int kc = 0; kc < KC; kc++)
                acc = mfma(ax[mt][kc], bw[nt][kc], acc);
            if (LAST) {
#pragma unroll
                for (int i = 0; i < 4; i++)
                    acc[i] += bo[nt];
                y[mt] = acc;
            } else {
#pragma unroll
                for (int i = 0; i < 4; i++)
                    Y[(mt * 16 + 4 * g + i) * ST + nt * 16 + r] = f2bf(fmaxf(acc[i] + bo[nt], 0.0f));
            }
        }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
}

// Forward of one 32-sample tile whose encoded inputs are in act_buf(base, 0); the output of
// the last layer is returned in y (lane holds rows 4g..4g+3 of tile mt for output r < 3).
template <bool PING>
__device__ __forceinline__ void forward_tile(uint16_t* base, const uint16_t* wb, const float* params, f32x4_t y[2])
{
    forward_layer<MLP_IN, MLP_H, false>(act_buf<PING>(base, 0), act_buf<PING>(base, 1), wb + LAYOUT.wb[0],
                                        params + LAYOUT.b[0], MLP_H, y);
#pragma unroll 1
    for (int l = 1; l < MLP_LAYERS - 1; l++)
        forward_layer<MLP_H, MLP_H, false>(act_buf<PING>(base, l), act_buf<PING>(base, l + 1), wb + LAYOUT.wb[l],
                                           params + LAYOUT.b[l], MLP_H, y);
    constexpr int L = MLP_LAYERS - 1;
    forward_layer<MLP_H, mlp_outp(L), true>(act_buf<PING>(base, L), nullptr, wb + LAYOUT.wb[L], params + LAYOUT.b[L],
                                            MLP_OUT, y);
}

struct CacheArgs {
    const float* params;
    const uint16_t* wb;
    const uint16_t* wt;
    float* partials;  // PARTIAL_ROWS rows of PSTRIDE floats
    float* loss_acc;
    float lo[3], inv_ext[3];
};

// Where the training samples of a batch come from: records (sample i is record
// start + i * stride mod n_rec, or, with n_rec_dev, a hashed draw among the min(*n_rec_dev,
// cap) records written so far, of which only the completed ones count) or, when features
// != nullptr, explicit feature rows and targets (n x 96 and n x 3 floats).
struct TrainSrc {
    const float* records;
    uint32_t n_rec;
    const uint32_t* n_rec_dev;
    uint32_t cap;
    uint32_t start, stride;
    const float* features;
    const float* targets;
};

// The training kernel's layers, split over the four waves of a block that share one 32-sample
// tile in LDS (the caller synchronises the block between layers).
//
// Forward: a 64-wide layer gives wave w output tile nt = w for both sample tiles; the last
// layer (LAST, 16 padded outputs) gives waves 0 and 1 sample tile mt = w, into y.
template <int K, int NO, bool LAST>
__device__ __forceinline__ void forward_layer4(const uint16_t* X, uint16_t* Y, const uint16_t* W, const float* bias,
                                               int nv, int w, f32x4_t& y)
{
    const int lane = __lane_id(), r = lane & 15, g = lane >> 4;
    constexpr int KC = K / 32;
    static_assert(LAST ? NO == 16 : NO == 64, "layer widths");
    if (LAST && w >= 2)
        return;
    const int nt = LAST ? 0 : w;
    bf16x8_t bw[KC];
#pragma unroll
    for (int kc = 0; kc < KC; kc++)
        bw[kc] = load8(W + (nt * 16 + r) * K + kc * 32 + 8 * g);
    const int o = nt * 16 + r;
    const float bo = o < nv ? bias[o] : 0.0f;
    constexpr int MT = LAST ? 1 : 2;
    bf16x8_t ax[MT][KC];
#pragma unroll
    for (int j = 0; j < MT; j++)
#pragma unroll
        for (int kc = 0; kc < KC; kc++)
            ax[j][kc] = load8(X + ((LAST ? w : j) * 16 + r) * ST + kc * 32 + 8 * g);
#pragma unroll
    for (int j = 0; j < MT; j++) {
        const int mt = LAST ? w : j;
        f32x4_t acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kc = 0; kc < KC; kc++)
            acc = mfma(ax[j][kc], bw[kc], acc);
        if (LAST) {
#pragma unroll
            for (int i = 0; i < 4; i++)
                acc[i] += bo;
            y = acc;
        } else {
#pragma unroll
            for (int i = 0; i < 4; i++)
                Y[(mt * 16 + 4 * g + i) * ST + o] = f2bf(fmaxf(acc[i] + bo, 0.0f));
        }
    }
}

// Backward: the weight-gradient tiles of column blocks kt = w, w + 4, ... (every row block) go
// to the block's partial row (stored by its first tile, first = true, added by later ones; each
// element has one writer lane), wave 0 adds the bias gradient, and with INPUT_GRAD wave w
// writes input-gradient column block kt = w (masked by the ReLU that made X) to Gp.
template <int K, int NO, int NTT, bool INPUT_GRAD>
__device__ __forceinline__ void backward_layer4(const uint16_t* G, const uint16_t* X, uint16_t* Gp,
                                                const uint16_t* WT, float* gW, float* gb, int nv, bool first, int w)
{
    const int lane = __lane_id(), r = lane & 15, g = lane >> 4;
    constexpr int OT = NO / 16, KT = K / 16;
    auto put = [first](float* p, float v) {
        if (first)
            *p = v;
        else
            *p += v;
    };
    bf16x8_t ga[OT];
#pragma unroll
    for (int ot = 0; ot < OT; ot++)
        ga[ot] = gather8(G + (8 * g) * ST + ot * 16 + r, ST);
    for (int kt = w; kt < KT; kt += 4) {
        const bf16x8_t xb = gather8(X + (8 * g) * ST + kt * 16 + r, ST);
#pragma unroll
        for (int ot = 0; ot < OT; ot++) {
            const f32x4_t d = mfma(ga[ot], xb, f32x4_t{0.0f, 0.0f, 0.0f, 0.0f});
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int o = ot * 16 + 4 * g + i;
                if (o < nv)
                    put(&gW[o * K + kt * 16 + r], d[i]);
            }
        }
    }
    if (w == 0 && lane < nv) {
        float sb = 0.0f;
        for (int sr = 0; sr < WS; sr++)
            sb += bf2f(G[sr * ST + lane]);
        put(&gb[lane], sb);
    }
    if (!INPUT_GRAD)
        return;
    static_assert(!INPUT_GRAD || KT == 4, "one input-gradient column block per wave");
    constexpr int OC = NTT / 32;
    const int kt = w;
    bf16x8_t bw[OC], ag[2][OC];
#pragma unroll
    for (int oc = 0; oc < OC; oc++)
        bw[oc] = load8(WT + (kt * 16 + r) * NTT + oc * 32 + 8 * g);
#pragma unroll
    for (int mt = 0; mt < 2; mt++)
#pragma unroll
        for (int oc = 0; oc < OC; oc++)
            ag[mt][oc] = load8(G + (mt * 16 + r) * ST + oc * 32 + 8 * g);
#pragma unroll
    for (int mt = 0; mt < 2; mt++) {
        f32x4_t acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int oc = 0; oc < OC; oc++)
            acc = mfma(ag[mt][oc], bw[oc], acc);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int srow = mt * 16 + 4 * g + i, k = kt * 16 + r;
            const bool on = X[srow * ST + k] != 0;  // post-ReLU activation > 0
            Gp[srow * ST + k] = on ? f2bf(acc[i]) : uint16_t(0);
        }
    }
}

// Training: forward + backward over n samples, one 32-sample tile at a time per block of four
// waves; the gradients (summed over the batch, already scaled by the loss normalisation) of the
// tiles block b takes (b, b + grid, ...) go to row b of the partial gradients, which
// k_apply_gradients sums. At most PARTIAL_ROWS blocks.
__global__ __launch_bounds__(256) void k_mlp_train(CacheArgs a, TrainSrc src, uint32_t n)
{
    float* const prow = a.partials + size_t(blockIdx.x) * PSTRIDE;
    __shared__ __attribute__((aligned(16))) uint16_t lds[(MLP_LAYERS + 2) * WS * ST];
    __shared__ float4 tl[WS];  // targets and use (1: the sample counts) of the tile's samples
    uint16_t* const gbase = lds + MLP_LAYERS * WS * ST;  // two gradient buffers
    const int w = threadIdx.x >> 6, lane = __lane_id(), r = lane & 15, g = lane >> 4;
    const float inv_n3 = 1.0f / (3.0f * float(n));
    float loss = 0.0f;
    const uint32_t n_rec = src.n_rec_dev ? min(*src.n_rec_dev, src.cap) : src.n_rec;
    if (!src.features && n_rec == 0) {  // nothing recorded yet: a zero gradient
        for (int i = threadIdx.x; i < PSTRIDE; i += 256)
            prow[i] = 0.0f;
        return;
    }
    constexpr int L = MLP_LAYERS - 1;
    for (uint32_t base = blockIdx.x * WS; base < n; base += gridDim.x * WS) {
        const bool first = base == blockIdx.x * WS;
        const uint32_t s = base + uint32_t(lane);
        if (lane < WS) {  // the tile's inputs, a quarter of the features per wave
            uint16_t* row = lds + lane * ST;
            float4 t = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (s < n) {
                if (src.features) {
                    for (int f = 24 * w; f < 24 * w + 24; f++)
                        row[f] = f2bf(src.features[size_t(s) * MLP_IN + f]);
                    t = make_float4(src.targets[size_t(s) * 3], src.targets[size_t(s) * 3 + 1],
                                    src.targets[size_t(s) * 3 + 2], 1.0f);
                } else {
                    uint32_t j;
                    if (src.n_rec_dev)
                        j = hash_u32(src.start + s * 0x9E3779B9u) % n_rec;
                    else
                        j = uint32_t((uint64_t(src.start) + uint64_t(s) * src.stride) % n_rec);
                    const float* rec = src.records + size_t(j) * MLP_REC_FLOATS;
                    encode_part(rec, a.lo, a.inv_ext, row, w);
                    // the target is complete: the sample counts
                    t = make_float4(rec[13], rec[14], rec[15], rec[19] != 0.0f ? 1.0f : 0.0f);
                }
            } else {
                for (int f = 24 * w; f < 24 * w + 24; f++)
                    row[f] = 0;
            }
            if (w == 0)
                tl[lane] = t;
        }
        __syncthreads();
        f32x4_t y = {0.0f, 0.0f, 0.0f, 0.0f};
        forward_layer4<MLP_IN, MLP_H, false>(act_buf<false>(lds, 0), act_buf<false>(lds, 1), a.wb + LAYOUT.wb[0],
                                             a.params + LAYOUT.b[0], MLP_H, w, y);
        __syncthreads();
#pragma unroll 1
        for (int l = 1; l < L; l++) {
            forward_layer4<MLP_H, MLP_H, false>(act_buf<false>(lds, l), act_buf<false>(lds, l + 1),
                                                a.wb + LAYOUT.wb[l], a.params + LAYOUT.b[l], MLP_H, w, y);
            __syncthreads();
        }
        forward_layer4<MLP_H, mlp_outp(L), true>(act_buf<false>(lds, L), nullptr, a.wb + LAYOUT.wb[L],
                                                 a.params + LAYOUT.b[L], MLP_OUT, w, y);
        // dL/dy of the relative L2 loss mean((y - t)^2 / (sg(y)^2 + 0.01)) into G5 (32 wide): waves
        // 0 and 1, sample tile mt = w, whose rows 4g..4g+3 the lane holds for output r
        uint16_t* G = gbase + (L & 1) * WS * ST;
        if (w < 2) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int srow = w * 16 + 4 * g + i;
                const float4 t = tl[srow];
                const float t_c = r == 0 ? t.x : (r == 1 ? t.y : t.z);
                float gr = 0.0f;
                if (r < 3 && t.w != 0.0f) {
                    const float yv = y[i];
                    const float den = yv * yv + 0.01f;
                    const float d = yv - t_c;
                    loss += d * d / den * inv_n3;
                    gr = 2.0f * d / den * inv_n3;
                }
                G[srow * ST + r] = f2bf(gr);
                G[srow * ST + 16 + r] = 0;
            }
        }
        __syncthreads();
        backward_layer4<MLP_H, mlp_outp(L), mlp_outt(L), true>(G, act_buf<false>(lds, L),
                                                               gbase + ((L - 1) & 1) * WS * ST, a.wt + LAYOUT.wt[L],
                                                               prow + LAYOUT.w[L], prow + LAYOUT.b[L], MLP_OUT, first, w);
        __syncthreads();
#pragma unroll 1
        for (int l = L - 1; l >= 1; l--) {
            backward_layer4<MLP_H, MLP_H, MLP_H, true>(gbase + (l & 1) * WS * ST, act_buf<false>(lds, l),
                                                       gbase + ((l - 1) & 1) * WS * ST, a.wt + LAYOUT.wt[l],
                                                       prow + LAYOUT.w[l], prow + LAYOUT.b[l], MLP_H, first, w);
            __syncthreads();
        }
        backward_layer4<MLP_IN, MLP_H, MLP_H, false>(gbase, act_buf<false>(lds, 0), nullptr, nullptr,
                                                     prow + LAYOUT.w[0], prow + LAYOUT.b[0], MLP_H, first, w);
        __syncthreads();  // the next tile's inputs overwrite what this one read
    }
    // wave-reduce the loss
    for (int off = 32; off >= 1; off >>= 1)
        loss += __shfl_xor(loss, off);
    if (lane == 0 && w < 2)
        atomicAdd(a.loss_acc, loss);
}

// Inference. Query mode (features == nullptr): each query record's prediction, clamped at
// zero and weighted by the record's throughput and weight, is added to the radiance of its
// path: next_rad[target] for a path that goes on, rad_out[path] for one that has ended
// (target = TARGET_ENDED | path). Test mode: predictions of explicit feature rows are
// written to out (n x 3).
__global__ __launch_bounds__(256) void k_mlp_infer(CacheArgs a, const float* queries, const uint32_t* query_slot,
                                                   const uint32_t* count_ptr, float4* next_rad, float4* rad_out,
                                                   const float* features, uint32_t n_explicit, float* out)
{
    __shared__ __attribute__((aligned(16))) uint16_t lds[4 * 2 * WS * ST];
    const int w = threadIdx.x >> 6;
    const int lane = __lane_id(), r = lane & 15, g = lane >> 4;
    uint16_t* const wbuf = lds + w * 2 * WS * ST;
    const uint32_t n = features ? n_explicit : *count_ptr;
    const uint32_t waves = gridDim.x * 4;
    for (uint32_t base = (blockIdx.x * 4 + w) * WS; base < n; base += waves * WS) {
        const uint32_t s = base + uint32_t(lane);
        if (lane < WS) {
            uint16_t* row = wbuf + lane * ST;
            if (s < n) {
                if (features)
                    for (int f = 0; f < MLP_IN; f++)
                        row[f] = f2bf(features[size_t(s) * MLP_IN + f]);
                else
                    encode_row(queries + size_t(s) * MLP_REC_FLOATS, a.lo, a.inv_ext, row);
            } else {
                for (int f = 0; f < MLP_IN; f++)
                    row[f] = 0;
            }
        }
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
        f32x4_t y[2];
        forward_tile<true>(wbuf, a.wb, a.params, y);
        for (int mt = 0; mt < 2; mt++)
            for (int i = 0; i < 4; i++) {
                const uint32_t q = base + uint32_t(mt * 16 + 4 * g + i);
                if (r >= 3 || q >= n)
                    continue;
                if (features) {
                    out[size_t(q) * 3 + r] = y[mt][i];
                } else {
                    const float* rec = queries + size_t(q) * MLP_REC_FLOATS;
                    const uint32_t tgt = query_slot[q];
                    float4* px = (tgt & TARGET_ENDED) ? &rad_out[tgt & ~TARGET_ENDED] : &next_rad[tgt];
                    float* dst = reinterpret_cast<float*>(px) + r;
                    *dst += rec[19] * rec[13 + r] * fmaxf(y[mt][i], 0.0f);
                }
            }
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
    }
}

// Encoded features of explicit records (13 floats each), as floats (bf16 values).
__global__ void k_encode(CacheArgs a, const float* rec13, uint32_t n, float* out)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        uint16_t row[MLP_IN];
        encode_row(rec13 + size_t(i) * 13, a.lo, a.inv_ext, row);
        for (int f = 0; f < MLP_IN; f++)
            out[size_t(i) * MLP_IN + f] = bf2f(row[f]);
    }
}

// Position of fp32 parameter i in the layout: a weight W[o][k] of layer l (true), or a bias.
__device__ __forceinline__ bool weight_of(int i, int& l, int& o, int& k)
{
#pragma unroll
    for (int j = 0; j < MLP_LAYERS; j++) {
        if (i < LAYOUT.b[j]) {
            const int e = i - LAYOUT.w[j];
            l = j;
            o = e / mlp_in(j);
            k = e - o * mlp_in(j);
            return true;
        }
        if (i < LAYOUT.b[j] + mlp_out(j))
            return false;
    }
    return false;
}

// Sum of the partial-gradient rows [0, rows) of parameter i, added to grads[i]: a block takes 64
// parameters, its four waves every fourth row each (16 loads in flight per lane), and the
// waves' sums are added in wave order (the same result on every run).
__device__ __forceinline__ float gradient_sum(const float* grads, const float* partials, uint32_t rows, int i,
                                              int count)
{
    __shared__ float part[4][64];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    float sum = 0.0f;
    if (i < count) {
        uint32_t b = uint32_t(w);
        for (; b + 60 < rows; b += 64) {
            float v[16];
#pragma unroll
            for (int j = 0; j < 16; j++)
                v[j] = partials[size_t(b + 4 * j) * PSTRIDE + i];
#pragma unroll
            for (int j = 0; j < 16; j++)
                sum += v[j];
        }
        for (; b < rows; b += 4)
            sum += partials[size_t(b) * PSTRIDE + i];
    }
    part[w][lane] = sum;
    __syncthreads();
    return i < count ? grads[i] + part[0][lane] + part[1][lane] + part[2][lane] + part[3][lane] : 0.0f;
}

// One Adam step on the gradient grads + the sum of the partial rows [0, rows) (blocks of 256
// threads, 64 parameters each; wave 0 updates). With ema != nullptr also the moving average of
// the weights, e <- d e + (1 - d) w from e = 0, and its bias-corrected value e / (1 - d^t)
// into ema_hat (dt = d^t). The bf16 copies of each updated weight (wb, wt; ema_wb of the
// average) are written in the same pass; their zero padding never changes. grads is cleared;
// the batch loss moves from loss_acc[0] to loss_acc[1] and the accumulator is cleared for the
// next batch. With adam = false the sum only goes to grads (the gradient of a batch).
__global__ __launch_bounds__(256) void k_apply_gradients(float* params, float* grads, const float* partials,
                                                         uint32_t rows, float* m, float* v, int count, int adam,
                                                         float lr, float b1t, float b2t, float* ema, float* ema_hat,
                                                         float d, float dt, uint16_t* wb, uint16_t* wt,
                                                         uint16_t* ema_wb, float* loss_acc)
{
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const float gr = gradient_sum(grads, partials, rows, i, count);
    if (threadIdx.x >= 64 || i >= count)
        return;
    if (!adam) {
        grads[i] = gr;
        return;
    }
    if (i == 0) {
        loss_acc[1] = loss_acc[0];
        loss_acc[0] = 0.0f;
    }
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float mi = b1 * m[i] + (1.0f - b1) * gr;
    const float vi = b2 * v[i] + (1.0f - b2) * gr * gr;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / (1.0f - b1t), vh = vi / (1.0f - b2t);
    const float w = params[i] - lr * mh / (sqrtf(vh) + eps);
    params[i] = w;
    grads[i] = 0.0f;
    float eh = 0.0f;
    if (ema) {
        const float e = d * ema[i] + (1.0f - d) * w;
        ema[i] = e;
        eh = e / (1.0f - dt);
        ema_hat[i] = eh;
    }
    int l, o, k;
    if (weight_of(i, l, o, k)) {
        wb[LAYOUT.wb[l] + o * mlp_in(l) + k] = f2bf(w);
        wt[LAYOUT.wt[l] + k * mlp_outt(l) + o] = f2bf(w);
        if (ema)
            ema_wb[LAYOUT.wb[l] + o * mlp_in(l) + k] = f2bf(eh);
    }
}

// bf16 copies of the weights in both layouts (padding rows/columns are zero); wt may be null.
__global__ void k_sync_bf16(const float* params, uint16_t* wb, uint16_t* wt)
{
    for (int l = 0; l < MLP_LAYERS; l++) {
        const int K = mlp_in(l), NO = mlp_outp(l), NT = mlp_outt(l), NV = mlp_out(l);
        const float* W = params + LAYOUT.w[l];
        for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < NO * K; e += gridDim.x * blockDim.x) {
            const int o = e / K, k = e % K;
            wb[LAYOUT.wb[l] + e] = o < NV ? f2bf(W[o * K + k]) : uint16_t(0);
        }
        for (int e = blockIdx.x * blockDim.x + threadIdx.x; wt && e < K * NT; e += gridDim.x * blockDim.x) {
            const int k = e / NT, o = e % NT;
            wt[LAYOUT.wt[l] + e] = o < NV ? f2bf(W[o * K + k]) : uint16_t(0);
        }
    }
}

CacheArgs args_of(const CacheState& c)
{
    CacheArgs a;
    a.params = c.params;
    a.wb = c.wb;
    a.wt = c.wt;
    a.partials = c.partials;
    a.loss_acc = c.loss_acc;
    for (int k = 0; k < 3; k++) {
        a.lo[k] = c.bounds_lo[k];
        const float ext = c.bounds_hi[k] - c.bounds_lo[k];
        a.inv_ext[k] = ext > 0.0f ? 1.0f / ext : 1.0f;
    }
    return a;
}

// The weights the renderer's queries read: the trained ones, or their moving average.
CacheArgs query_args_of(const CacheState& c)
{
    CacheArgs a = args_of(c);
    if (c.ema_decay > 0.0f) {
        a.params = c.ema_hat;
        a.wb = c.ema_wb;
        a.wt = nullptr;
    }
    return a;
}

}  // namespace

int mlp_param_count() { return LAYOUT.count; }

void cache_free(CacheState& c)
{
    void* ptrs[] = {c.params, c.grads, c.adam_m, c.adam_v, c.wb, c.wt, c.records, c.queries,
                    c.query_slot, c.counters, c.loss_acc, c.ema, c.ema_hat, c.ema_wb, c.partials};
    for (void* p : ptrs)
        if (p)
            (void)hipFree(p);
    const float lo[3] = {c.bounds_lo[0], c.bounds_lo[1], c.bounds_lo[2]};
    const float hi[3] = {c.bounds_hi[0], c.bounds_hi[1], c.bounds_hi[2]};
    c = CacheState{};
    memcpy(c.bounds_lo, lo, sizeof(lo));
    memcpy(c.bounds_hi, hi, sizeof(hi));
}

hipError_t cache_sync_weights(CacheState& c, hipStream_t st)
{
    // (the average's copy too, whether or not queries read it: Adam steps then keep its weights
    // current and its padding stays zero)
    hipLaunchKernelGGL(k_sync_bf16, dim3(64), dim3(256), 0, st, c.params, c.wb, c.wt);
    hipLaunchKernelGGL(k_sync_bf16, dim3(64), dim3(256), 0, st, c.ema_hat, c.ema_wb, nullptr);
    return hipGetLastError();
}

// Restarts the moving average at the current weights (no step averaged yet).
static hipError_t cache_reset_ema(CacheState& c, hipStream_t st)
{
    c.ema_steps = 0;
    hipError_t e = hipMemsetAsync(c.ema, 0, size_t(LAYOUT.count) * 4, st);
    if (e != hipSuccess)
        return e;
    return hipMemcpyAsync(c.ema_hat, c.params, size_t(LAYOUT.count) * 4, hipMemcpyDeviceToDevice, st);
}

// He-uniform initialisation from a Splitmix32 stream; biases zero.
hipError_t cache_init(CacheState& c, uint32_t pool_slots, uint32_t record_capacity, hipStream_t st)
{
    hipError_t e = hipSuccess;
    auto A = [&](auto*& p, size_t bytes) {
        if (e == hipSuccess)
            e = hipMalloc((void**)&p, std::max<size_t>(bytes, 16));
    };
    if (!c.params) {
        A(c.params, size_t(LAYOUT.count) * 4);
        A(c.grads, size_t(LAYOUT.count) * 4);
        A(c.adam_m, size_t(LAYOUT.count) * 4);
        A(c.adam_v, size_t(LAYOUT.count) * 4);
        A(c.wb, size_t(LAYOUT.wb_count) * 2);
        A(c.wt, size_t(LAYOUT.wt_count) * 2);
        A(c.counters, 8 * 4);
        A(c.loss_acc, 2 * 4);  // [0] the batch being trained, [1] the last one stepped
        A(c.partials, size_t(PARTIAL_ROWS) * PSTRIDE * 4);
        A(c.ema, size_t(LAYOUT.count) * 4);
        A(c.ema_hat, size_t(LAYOUT.count) * 4);
        A(c.ema_wb, size_t(LAYOUT.wb_count) * 2);
    }
    if (e == hipSuccess && record_capacity > c.record_capacity) {
        if (c.records)
            (void)hipFree(c.records);
        c.records = nullptr;
        A(c.records, size_t(record_capacity) * MLP_REC_FLOATS * 4);
        c.record_capacity = record_capacity;
    }
    if (e == hipSuccess && pool_slots > c.query_capacity) {
        if (c.queries)
            (void)hipFree(c.queries);
        if (c.query_slot)
            (void)hipFree(c.query_slot);
        c.queries = nullptr;
        c.query_slot = nullptr;
        A(c.queries, size_t(pool_slots) * MLP_REC_FLOATS * 4);
        A(c.query_slot, size_t(pool_slots) * 4);
        c.query_capacity = pool_slots;
    }
    if (e != hipSuccess)
        return e;
    std::vector<float> p(LAYOUT.count, 0.0f);
    uint32_t state = c.seed;
    auto next = [&]() {  // Splitmix32
        uint32_t z = (state += 0x9e3779b9u);
        z = (z ^ (z >> 16)) * 0x85ebca6bu;
        z = (z ^ (z >> 13)) * 0xc2b2ae35u;
        return z ^ (z >> 16);
    };
    for (int l = 0; l < MLP_LAYERS; l++) {
        const float lim = std::sqrt(6.0f / float(mlp_in(l)));
        for (int i = 0; i < mlp_out(l) * mlp_in(l); i++)
            p[LAYOUT.w[l] + i] = (float(next()) * 0x1p-32f * 2.0f - 1.0f) * lim;
    }
    if ((e = hipMemcpyAsync(c.params, p.data(), p.size() * 4, hipMemcpyHostToDevice, st)) != hipSuccess)
        return e;
    if ((e = hipMemsetAsync(c.grads, 0, p.size() * 4, st)) != hipSuccess)
        return e;
    if ((e = hipMemsetAsync(c.adam_m, 0, p.size() * 4, st)) != hipSuccess)
        return e;
    if ((e = hipMemsetAsync(c.adam_v, 0, p.size() * 4, st)) != hipSuccess)
        return e;
    if ((e = hipMemsetAsync(c.counters, 0, 8 * 4, st)) != hipSuccess)
        return e;
    if ((e = hipMemsetAsync(c.loss_acc, 0, 2 * 4, st)) != hipSuccess)
        return e;
    if ((e = cache_reset_ema(c, st)) != hipSuccess)
        return e;
    c.steps_done = 0;
    c.configured = true;
    return cache_sync_weights(c, st);
}

hipError_t cache_set_params(CacheState& c, const float* host, hipStream_t st)
{
    hipError_t e = hipMemcpyAsync(c.params, host, size_t(LAYOUT.count) * 4, hipMemcpyHostToDevice, st);
    if (e != hipSuccess)
        return e;
    if ((e = cache_reset_ema(c, st)) != hipSuccess)
        return e;
    return cache_sync_weights(c, st);
}

hipError_t cache_get_query_params(CacheState& c, float* host, hipStream_t st)
{
    const float* src = c.ema_decay > 0.0f ? c.ema_hat : c.params;
    hipError_t e = hipMemcpyAsync(host, src, size_t(LAYOUT.count) * 4, hipMemcpyDeviceToHost, st);
    if (e != hipSuccess)
        return e;
    return hipStreamSynchronize(st);
}

hipError_t cache_get_params(CacheState& c, float* host, hipStream_t st)
{
    hipError_t e = hipMemcpyAsync(host, c.params, size_t(LAYOUT.count) * 4, hipMemcpyDeviceToHost, st);
    if (e != hipSuccess)
        return e;
    return hipStreamSynchronize(st);
}

hipError_t cache_forward(CacheState& c, const float* d_features, uint32_t n, float* d_out, hipStream_t st)
{
    const unsigned blocks = unsigned(std::min<uint64_t>((uint64_t(n) + 4 * WS - 1) / (4 * WS) + 1, 4096));
    hipLaunchKernelGGL(k_mlp_infer, dim3(blocks), dim3(256), 0, st, args_of(c), nullptr, nullptr, nullptr, nullptr,
                       nullptr, d_features, n, d_out);
    return hipGetLastError();
}

hipError_t cache_encode(CacheState& c, const float* d_rec13, uint32_t n, float* d_out, hipStream_t st)
{
    const unsigned blocks = unsigned(std::min<uint64_t>((uint64_t(n) + 255) / 256, 1024));
    hipLaunchKernelGGL(k_encode, dim3(std::max(1u, blocks)), dim3(256), 0, st, args_of(c), d_rec13, n, d_out);
    return hipGetLastError();
}

hipError_t cache_query(CacheState& c, const uint32_t* d_count, float4* next_rad, float4* rad_out, unsigned blocks,
                       hipStream_t st)
{
    hipLaunchKernelGGL(k_mlp_infer, dim3(blocks), dim3(256), 0, st, query_args_of(c), c.queries, c.query_slot, d_count,
                       next_rad, rad_out, nullptr, 0u, nullptr);
    return hipGetLastError();
}

// One training launch over n samples into the partial-gradient rows; returns the rows used.
static uint32_t launch_train(CacheState& c, const TrainSrc& src, uint32_t n, hipStream_t st)
{
    const uint32_t rows = uint32_t(std::max<uint64_t>(1, std::min<uint64_t>((uint64_t(n) + WS - 1) / WS, PARTIAL_ROWS)));
    hipLaunchKernelGGL(k_mlp_train, dim3(rows), dim3(256), 0, st, args_of(c), src, n);
    return rows;
}

// Sums the partial rows into c.grads (adam = false), or takes the Adam step on that sum.
static hipError_t apply_gradients(CacheState& c, uint32_t rows, bool adam, hipStream_t st)
{
    float b1t = 0.0f, b2t = 0.0f, dt = 0.0f;
    const bool ema = adam && c.ema_decay > 0.0f;
    if (adam) {
        c.steps_done++;
        b1t = std::pow(0.9f, float(c.steps_done));
        b2t = std::pow(0.999f, float(c.steps_done));
        if (ema)
            c.ema_steps++;
        dt = ema ? float(std::pow(double(c.ema_decay), double(c.ema_steps))) : 0.0f;
    }
    const unsigned blocks = unsigned((LAYOUT.count + 63) / 64);
    hipLaunchKernelGGL(k_apply_gradients, dim3(blocks), dim3(256), 0, st, c.params, c.grads, c.partials, rows,
                       c.adam_m, c.adam_v, LAYOUT.count, adam ? 1 : 0, c.lr, b1t, b2t, ema ? c.ema : nullptr,
                       c.ema_hat, c.ema_decay, dt, c.wb, c.wt, c.ema_wb, c.loss_acc);
    return hipGetLastError();
}

// Gradients of the batch added into c.grads and the loss into loss_acc[0] (device).
hipError_t cache_backward(CacheState& c, const float* d_features, const float* d_targets, uint32_t n,
                          uint32_t n_rec, uint32_t start, uint32_t stride, hipStream_t st)
{
    hipError_t e = hipMemsetAsync(c.loss_acc, 0, 4, st);
    if (e != hipSuccess)
        return e;
    TrainSrc src{c.records, n_rec, nullptr, c.record_capacity, start, stride, d_features, d_targets};
    const uint32_t rows = launch_train(c, src, n, st);
    if ((e = hipGetLastError()) != hipSuccess)
        return e;
    return apply_gradients(c, rows, false, st);
}

// Adam step on the gradient in c.grads.
hipError_t cache_adam(CacheState& c, hipStream_t st) { return apply_gradients(c, 0, true, st); }

// Steps on the records written so far in the current job (count read on the device; only
// completed records count): lets the network learn while a job's paths are still in flight.
hipError_t cache_train_inflight(CacheState& c, hipStream_t st)
{
    hipError_t e;
    for (uint32_t k = 0; k < c.train_steps; k++) {
        TrainSrc src{c.records, 0, c.counters, c.record_capacity, (c.steps_done + 1) * 0x6C8E9CF5u, 0, nullptr,
                     nullptr};
        const uint32_t rows = launch_train(c, src, c.batch, st);
        if ((e = hipGetLastError()) != hipSuccess)
            return e;
        if ((e = apply_gradients(c, rows, true, st)) != hipSuccess)
            return e;
    }
    return hipSuccess;
}

// Trains on the records collected by the last job: steps minibatches of the batch size,
// each a pseudo-shuffled window (start + i * stride mod n_records, stride coprime with n).
hipError_t cache_train_records(CacheState& c, uint32_t n_records, hipStream_t st)
{
    if (n_records == 0)
        return hipSuccess;
    const uint32_t batch = std::min(c.batch, n_records);
    uint32_t stride = 7919u % n_records;
    auto gcd = [](uint32_t x, uint32_t y) {
        while (y) {
            uint32_t t = x % y;
            x = y;
            y = t;
        }
        return x;
    };
    if (stride == 0)
        stride = 1;
    while (gcd(stride, n_records) != 1)
        stride++;
    hipError_t e;
    for (uint32_t k = 0; k < c.train_steps; k++) {
        const uint32_t start = uint32_t((uint64_t(c.steps_done + k) * 2654435761u) % n_records);
        TrainSrc src{c.records, n_records, nullptr, c.record_capacity, start, stride, nullptr, nullptr};
        const uint32_t rows = launch_train(c, src, batch, st);
        if ((e = hipGetLastError()) != hipSuccess)
            return e;
        if ((e = apply_gradients(c, rows, true, st)) != hipSuccess)
            return e;
    }
    return hipSuccess;
}

}  // namespace ptk
