// pt_nrc.hpp -- neural radiance cache: network shape, parameter and fragment layout,
// host object and launchers (nrc.hip).
//
// Network: 64 inputs -> 4 hidden layers of 64 (ReLU) -> 3 outputs (radiance), no
// biases (input 59 is the constant 1).  Inputs: position in the scene box, 6
// frequencies x (sin, cos) per axis (36); the outgoing direction in real spherical
// harmonics of bands 0-3 (16); shading normal (3); roughness (1); base colour (3); 1;
// zeros to 64.  BF16 weights and activations on MFMA (v_mfma_f32_16x16x32_bf16), FP32
// accumulation, FP32 master weights trained with Adam.
#pragma once

#include <stdint.h>

#include "pt_host.hpp"

namespace bipt {

constexpr int kNrcIn = 64;
constexpr int kNrcHidden = 64;
constexpr int kNrcOutPad = 16;                // output layer padded to one 16-wide MFMA tile
constexpr int kNrcOut = 3;
constexpr int kNrcLayers = 5;
// FP32 master weights W_l[k][i] (input k, output i), row-major, layer after layer
constexpr int kNrcParams = 4 * kNrcIn * kNrcHidden + kNrcHidden * kNrcOutPad;   // 17,408
constexpr int kNrcFwd = kNrcParams;                                             // BF16 forward fragments
constexpr int kNrcBwd = 4 * kNrcIn * kNrcHidden + kNrcHidden * 32;              // BF16 backward fragments
constexpr float kNrcLossEps = 0.01f;          // relative L2: (y - t)^2 / (y^2 + eps)

struct NrcInferArgs {
    const NrcRecord* rec;
    const uint32_t* count_in;   // entries (null: num)
    uint32_t num;               // launch capacity
    const void* fwd;            // forward fragments (bf16)
    const float* bounds;        // scene box: lo xyz, hi xyz
    PathState ps;               // radiance kind 0 and throughput of the queried paths
    float* raw_out;             // non-null: raw network outputs, 3 per entry (all entries)
};
void launch_nrc_infer(const NrcInferArgs& a, hipStream_t s);

struct NrcTrainArgs {
    const NrcRecord* rec;
    const float* target;        // 3 per sample
    uint32_t row0, rows;        // the step's samples
    const void* fwd;
    const void* bwd;
    const float* bounds;
    float* grad_part;           // kNrcParams per 64-sample chunk
    float* loss_part;           // per chunk: loss sum, valid samples
    float scale;                // gradient of the mean loss per sample term
    const float* clamp;         // device scalar: upper bound of every target channel (null: none)
};
void launch_nrc_train(const NrcTrainArgs& a, hipStream_t s);

struct NrcAdamArgs {
    float* w;
    float* m;
    float* v;
    void* fwd;
    void* bwd;
    const float* grad_part;
    uint32_t chunks;
    float lr, b1, b2, eps, c1, c2;   // c1 = 1 / (1 - b1^t), c2 = 1 / (1 - b2^t)
    uint32_t update;                 // 0: only rebuild the fragments from w
};
void launch_nrc_adam(const NrcAdamArgs& a, hipStream_t s);

}  // namespace bipt

// A radiance cache bound to one context's scene.
struct bi_pt_nrc : bipt::Owner {
    bi_pt_ctx* ctx = nullptr;
    uint32_t train_samples = 65536, train_steps = 4, cache_bounce = 2, use_after = 0, flags = 0;
    float lr = 3e-3f, b1 = 0.9f, b2 = 0.99f, eps = 1e-15f;
    float clamp_factor = 16.0f;   // targets clamped to this multiple of their mean luminance (<= 0: none)
    bipt::DevBuf<float> clamp_part;   // per-block luminance sums and counts, then the bound
    bipt::DevBuf<float> w, m, v, targets, grad_part, loss_part, bounds;
    bipt::DevBuf<uint16_t> fwd, bwd;
    bipt::DevBuf<bipt::NrcRecord> train_rec, query;
    uint32_t grad_chunks = 0;     // chunk capacity of grad_part / loss_part
    uint64_t steps = 0, frames = 0;
    uint32_t last_chunks = 0;     // loss chunks written by the last frame's steps
    uint32_t last_grad_chunks = 0;   // chunks of the last step's gradient partials
};

namespace bipt {
// The training of one frame: train_steps Adam steps over the frame's training samples.
bi_pt_status nrc_train_frame(bi_pt_nrc* nrc, hipStream_t s);
// Query buffer for waves of `cap` entries.
bi_pt_status nrc_ensure_query(bi_pt_nrc* nrc, uint32_t cap);
}  // namespace bipt
