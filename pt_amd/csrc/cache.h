// Learned radiance cache: a small MLP trained online on the tracer's own path suffixes.
//
// Network: 96 inputs (frequency-encoded position + direction, normal, roughness, albedo),
// five hidden layers of 64 ReLU units, 3 linear outputs (RGB). Weights are kept as fp32
// master copies (Adam state alongside) and as bf16 copies in two layouts for the MFMA
// kernels: row-major W[out][in] (forward) and W^T[in][out] (input gradients). Optionally the
// renderer's queries read an exponential moving average of the weights (fp32 and bf16
// W[out][in]), which damps the step-to-step jitter of training while the paths use it.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace ptk {

constexpr int MLP_IN = 96;
constexpr int MLP_H = 64;
constexpr int MLP_OUT = 3;
constexpr int MLP_LAYERS = 6;
constexpr int MLP_REC_FLOATS = 20;  // record: pos3 dir3 normal3 rough1 albedo3 | T-or-target3 | L_before3 | pad

// Per-layer geometry: input width, output width, padded output width of the bf16 copies.
__host__ __device__ constexpr int mlp_in(int l) { return l == 0 ? MLP_IN : MLP_H; }
__host__ __device__ constexpr int mlp_out(int l) { return l == MLP_LAYERS - 1 ? MLP_OUT : MLP_H; }
__host__ __device__ constexpr int mlp_outp(int l) { return l == MLP_LAYERS - 1 ? 16 : MLP_H; }   // W rows
__host__ __device__ constexpr int mlp_outt(int l) { return l == MLP_LAYERS - 1 ? 32 : MLP_H; }   // W^T cols

// Offsets (in floats) of the fp32 parameter vector: for each layer W[out][in] then b[out]
// (true output widths, no padding). Total 22720 weights + 323 biases.
struct MlpLayout {
    int w[MLP_LAYERS], b[MLP_LAYERS];
    int wb[MLP_LAYERS], wt[MLP_LAYERS];  // offsets (bf16 elements) in the two bf16 copies
    int count, wb_count, wt_count;
    __host__ __device__ constexpr MlpLayout() : w{}, b{}, wb{}, wt{}, count(0), wb_count(0), wt_count(0)
    {
        int o = 0, ob = 0, ot = 0;
        for (int l = 0; l < MLP_LAYERS; l++) {
            w[l] = o;
            o += mlp_out(l) * mlp_in(l);
            b[l] = o;
            o += mlp_out(l);
            wb[l] = ob;
            ob += mlp_outp(l) * mlp_in(l);
            wt[l] = ot;
            ot += mlp_in(l) * mlp_outt(l);
        }
        count = o;
        wb_count = ob;
        wt_count = ot;
    }
};

struct CacheState {
    bool configured = false;
    bool query = false;          // terminate non-training paths at the 2nd bounce
    uint32_t train_stride = 32;  // every n-th path is a training path
    uint32_t train_steps = 4;
    uint32_t batch = 16384;
    float lr = 1e-3f;
    uint32_t seed = 1;
    uint32_t frame_spp = 1;      // samples per render job while training
    float cont_prob = 0.0f;      // control-variate continuation probability (0: terminate)
    uint32_t query_vertex = 2;   // path vertex where records are taken and queries answered
    float ema_decay = 0.0f;      // > 0: queries read the moving average of the weights
    float query_max_pdf = 0.0f;  // > 0: query only paths whose last direction had a pdf below it
    uint32_t ema_steps = 0;      // steps averaged since the average restarted
    uint32_t steps_done = 0;
    float last_loss = 0.0f;
    uint64_t records_total = 0;
    uint32_t last_records = 0;
    uint64_t queries_total = 0;
    float bounds_lo[3] = {0, 0, 0}, bounds_hi[3] = {1, 1, 1};

    float* params = nullptr;     // fp32 master
    float* grads = nullptr;
    float* adam_m = nullptr;
    float* adam_v = nullptr;
    uint16_t* wb = nullptr;      // bf16 W[out][in]
    uint16_t* wt = nullptr;      // bf16 W^T[in][out]
    float* records = nullptr;    // training records (MLP_REC_FLOATS each)
    uint32_t record_capacity = 0;
    float* queries = nullptr;    // query records of one iteration
    uint32_t* query_slot = nullptr;  // radiance target of each query
    uint32_t query_capacity = 0;
    uint32_t* counters = nullptr;  // [0] training records, [1] queries
    float* loss_acc = nullptr;   // [0] loss of the batch being trained, [1] of the last Adam step
    float* partials = nullptr;   // per-block partial gradients of a training launch
    float* ema = nullptr;        // moving average of the weights (from zero)
    float* ema_hat = nullptr;    // its bias-corrected value, ema / (1 - decay^ema_steps)
    uint16_t* ema_wb = nullptr;  // bf16 W[out][in] of ema_hat
};

}  // namespace ptk
