// nrc.hip -- neural radiance cache on gfx950 matrix cores: input encoding, inference
// into the path accumulators, online training (forward and backward in BF16 with FP32
// accumulation, per-chunk weight-gradient partials summed in a fixed order), Adam, and
// the C API (bi_pt_nrc_*, bi_pt_render_nrc).  No reference counterpart: the reference
// pass traces every path to its last segment (path_tracing.cpp:290-343).
//
// MFMA v_mfma_f32_16x16x32_bf16, per wave of 64 lanes (lane l, g = l / 16):
//   A (16 x 32): row l % 16, columns 8g .. 8g + 7 (8 bf16 per lane)
//   B (32 x 16): column l % 16, rows 8g .. 8g + 7
//   D (16 x 16): column l % 16, rows 4g .. 4g + 3 (4 floats per lane)
// A layer out = X W is computed transposed, out^T = W^T X^T, with the weights as A
// (fragments pre-arranged so a lane loads its 8 values with one 16-B load) and the
// activations as B: a lane reads 8 consecutive features of one sample row from LDS and
// receives 4 consecutive output features of one row, so both LDS accesses are
// contiguous.  Activations stay in LDS, one 64-sample tile per wave.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "pt_kernels.hpp"
#include "pt_nrc.hpp"

namespace bipt {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kLd = 72;   // bf16 per LDS row: 64 features + 8 pad (144 B, 16-B aligned rows)

__device__ __forceinline__ void wave_fence() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ f32x4 mfma(const bf16x8& a, const bf16x8& b, const f32x4& c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Fragment (in units of 8 bf16) of forward layer l, output tile t, input step s.
__device__ __forceinline__ int fwd_frag(int l, int t, int s, int lane) {
    return l < 4 ? l * 512 + (t * 2 + s) * 64 + lane : 2048 + s * 64 + lane;
}
// Backward fragment of layer l, input tile tp, output step sp.
__device__ __forceinline__ int bwd_frag(int l, int tp, int sp, int lane) {
    return l < 4 ? l * 512 + (tp * 2 + sp) * 64 + lane : 2048 + tp * 64 + lane;
}

// Input encoding of one record into 64 floats.
__device__ __forceinline__ void encode(const NrcRecord& r, const float* bounds, float* x) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float lo = bounds[a], ext = bounds[3 + a] - bounds[a];
        const float u = ext > 0.0f ? (r.p[a] - lo) / ext : 0.5f;
#pragma unroll
        for (int f = 0; f < 6; ++f) {
            float sv, cv;
            sincospif(u * (float)(1 << f), &sv, &cv);
            x[a * 12 + 2 * f] = sv;
            x[a * 12 + 2 * f + 1] = cv;
        }
    }
    const float dx = r.wo[0], dy = r.wo[1], dz = r.wo[2];
    const float xx = dx * dx, yy = dy * dy, zz = dz * dz;
    float* h = x + 36;
    h[0] = 0.28209479177387814f;
    h[1] = -0.4886025119029199f * dy;
    h[2] = 0.4886025119029199f * dz;
    h[3] = -0.4886025119029199f * dx;
    h[4] = 1.0925484305920792f * (dx * dy);
    h[5] = -1.0925484305920792f * (dy * dz);
    h[6] = 0.9461746957575601f * zz - 0.31539156525252005f;
    h[7] = -1.0925484305920792f * (dx * dz);
    h[8] = 0.5462742152960396f * (xx - yy);
    h[9] = -0.5900435899266435f * (dy * (3.0f * xx - yy));
    h[10] = 2.890611442640554f * ((dx * dy) * dz);
    h[11] = -0.4570457994644658f * (dy * (5.0f * zz - 1.0f));
    h[12] = 0.3731763325901154f * (dz * (5.0f * zz - 3.0f));
    h[13] = -0.4570457994644658f * (dx * (5.0f * zz - 1.0f));
    h[14] = 1.445305721320277f * (dz * (xx - yy));
    h[15] = -0.5900435899266435f * (dx * (xx - 3.0f * yy));
    x[52] = r.n[0];
    x[53] = r.n[1];
    x[54] = r.n[2];
    x[55] = r.roughness;
    x[56] = r.albedo[0];
    x[57] = r.albedo[1];
    x[58] = r.albedo[2];
    x[59] = 1.0f;
    x[60] = x[61] = x[62] = x[63] = 0.0f;
}

__device__ __forceinline__ void store_row(__bf16* row, const float* x) {
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        bf16x8 v;
#pragma unroll
        for (int e = 0; e < 8; ++e) v[e] = (__bf16)x[8 * c + e];
        reinterpret_cast<bf16x8*>(row)[c] = v;
    }
}

// Hidden layer l on the wave's 64-row tile: Y = relu(X W_l), X and Y in LDS (may alias).
__device__ __forceinline__ void hidden_layer(const bf16x8* fwd, int l, const __bf16* X, __bf16* Y, int lane) {
    f32x4 acc[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int t = 0; t < 4; ++t) acc[r][t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    const int rl = lane & 15, g = lane >> 4;
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        bf16x8 w[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) w[t] = fwd[fwd_frag(l, t, s, lane)];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const bf16x8 xb = *reinterpret_cast<const bf16x8*>(X + (16 * r + rl) * kLd + 32 * s + 8 * g);
#pragma unroll
            for (int t = 0; t < 4; ++t) acc[r][t] = mfma(w[t], xb, acc[r][t]);
        }
    }
    wave_fence();
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            bf16x4 v;
#pragma unroll
            for (int e = 0; e < 4; ++e) v[e] = (__bf16)fmaxf(acc[r][t][e], 0.0f);
            *reinterpret_cast<bf16x4*>(Y + (16 * r + rl) * kLd + 16 * t + 4 * g) = v;
        }
    wave_fence();
}

// Output layer: o[r][e] on lanes 0..15 is output e of row 16 r + lane.
__device__ __forceinline__ void output_layer(const bf16x8* fwd, const __bf16* X, f32x4* o, int lane) {
    const int rl = lane & 15, g = lane >> 4;
#pragma unroll
    for (int r = 0; r < 4; ++r) o[r] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const bf16x8 w = fwd[fwd_frag(4, 0, s, lane)];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const bf16x8 xb = *reinterpret_cast<const bf16x8*>(X + (16 * r + rl) * kLd + 32 * s + 8 * g);
            o[r] = mfma(w, xb, o[r]);
        }
    }
}

// Inference: one wave per 64 entries.  Entries whose tag names a path add throughput x
// max(output, 0) as the path's radiance of kind 0.
__global__ __launch_bounds__(256) void k_nrc_infer(NrcInferArgs a) {
    __shared__ __align__(16) __bf16 s_x[4][64 * kLd];
    __shared__ uint32_t s_tag[4][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t count = a.count_in ? *a.count_in : a.num;
    const uint32_t row0 = (blockIdx.x * 4u + (uint32_t)wave) * 64u;
    if (row0 >= count) return;
    __bf16* X = s_x[wave];
    const uint32_t i = row0 + (uint32_t)lane;
    uint32_t tag = kNoPath;
    float x[64];
    bool live = false;
    if (i < count) {
        const NrcRecord r = a.rec[i];
        tag = r.tag;
        live = a.raw_out != nullptr || tag != kNoPath;
        if (live) encode(r, a.bounds, x);
    }
    if (!live)
        for (int k = 0; k < 64; ++k) x[k] = 0.0f;
    store_row(X + lane * kLd, x);
    s_tag[wave][lane] = tag;
    wave_fence();
    const bf16x8* fwd = reinterpret_cast<const bf16x8*>(a.fwd);
    for (int l = 0; l < 4; ++l) hidden_layer(fwd, l, X, X, lane);
    f32x4 o[4];
    output_layer(fwd, X, o, lane);
    if (lane < 16) {
        const size_t cap = a.ps.cap;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const uint32_t row = 16u * r + (uint32_t)lane;
            const uint32_t e = row0 + row;
            if (e >= count) continue;
            if (a.raw_out) {
                a.raw_out[3 * (size_t)e] = o[r][0];
                a.raw_out[3 * (size_t)e + 1] = o[r][1];
                a.raw_out[3 * (size_t)e + 2] = o[r][2];
                continue;
            }
            const uint32_t path = s_tag[wave][row];
            if (path == kNoPath) continue;
            V3 c = v3(a.ps.bx[path] * fmaxf(o[r][0], 0.0f), a.ps.by[path] * fmaxf(o[r][1], 0.0f),
                      a.ps.bz[path] * fmaxf(o[r][2], 0.0f));
            if (!finite3(c)) c = v3(0.0f, 0.0f, 0.0f);
            a.ps.acc[path] = c.x;
            a.ps.acc[cap + path] = c.y;
            a.ps.acc[2 * cap + path] = c.z;
        }
    }
}

// 8 bf16 of column `col` in rows row0 .. row0 + 7.
__device__ __forceinline__ bf16x8 gather_col(const __bf16* M, int row0, int col) {
    bf16x8 v;
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = M[(row0 + e) * kLd + col];
    return v;
}

// The bound of a frame's training targets: factor x the mean luminance of the valid
// targets.  Per 256-slot block the luminance sum and count (a fixed-order tree in LDS),
// then one block adds the block partials in order, so the bound is reproducible.
__global__ __launch_bounds__(256) void k_target_lum(const NrcRecord* rec, const float* target, uint32_t n,
                                                    float* part) {
    __shared__ float s_sum[256], s_cnt[256];
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    float l = 0.0f, c = 0.0f;
    if (i < n && rec[i].tag == 1u) {
        const float r = target[3 * (size_t)i], g = target[3 * (size_t)i + 1], b = target[3 * (size_t)i + 2];
        const float y = (0.2126f * r + 0.7152f * g) + 0.0722f * b;
        if (isfinite(y)) {
            l = y > 0.0f ? y : 0.0f;
            c = 1.0f;
        }
    }
    s_sum[threadIdx.x] = l;
    s_cnt[threadIdx.x] = c;
    __syncthreads();
    for (uint32_t h = 128u; h > 0u; h >>= 1) {
        if (threadIdx.x < h) {
            s_sum[threadIdx.x] += s_sum[threadIdx.x + h];
            s_cnt[threadIdx.x] += s_cnt[threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * (size_t)blockIdx.x] = s_sum[0];
        part[2 * (size_t)blockIdx.x + 1] = s_cnt[0];
    }
}
__global__ __launch_bounds__(256) void k_target_bound(float* part, uint32_t blocks, float factor, float* bound) {
    __shared__ float s_sum[256], s_cnt[256];
    float l = 0.0f, c = 0.0f;
    for (uint32_t b = threadIdx.x; b < blocks; b += 256u) {
        l += part[2 * (size_t)b];
        c += part[2 * (size_t)b + 1];
    }
    s_sum[threadIdx.x] = l;
    s_cnt[threadIdx.x] = c;
    __syncthreads();
    for (uint32_t h = 128u; h > 0u; h >>= 1) {
        if (threadIdx.x < h) {
            s_sum[threadIdx.x] += s_sum[threadIdx.x + h];
            s_cnt[threadIdx.x] += s_cnt[threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) *bound = s_cnt[0] > 0.0f ? factor * (s_sum[0] / s_cnt[0]) : __builtin_inff();
}

// Training on one 64-sample chunk (one wave): forward keeping every layer's input,
// relative-L2 loss gradient, backward; the chunk's weight gradients go to its partial.
__global__ __launch_bounds__(64) void k_nrc_train(NrcTrainArgs a) {
    __shared__ __align__(16) __bf16 s_h[5][64 * kLd];   // layer inputs H_0 .. H_4
    __shared__ __align__(16) __bf16 s_d[64 * kLd];      // gradient of the current layer's output
    __shared__ float s_t[64][4];                        // target rgb, valid
    const int lane = threadIdx.x;
    const int rl = lane & 15, g = lane >> 4;
    const uint32_t chunk = blockIdx.x;
    const uint32_t local = chunk * 64u + (uint32_t)lane;
    float x[64];
    bool valid = false;
    float tg[3] = {0.0f, 0.0f, 0.0f};
    if (local < a.rows) {
        const uint32_t i = a.row0 + local;
        const NrcRecord r = a.rec[i];
        valid = r.tag == 1u;
        if (valid) {
            encode(r, a.bounds, x);
            tg[0] = a.target[3 * (size_t)i];
            tg[1] = a.target[3 * (size_t)i + 1];
            tg[2] = a.target[3 * (size_t)i + 2];
            valid = isfinite(tg[0]) && isfinite(tg[1]) && isfinite(tg[2]);
            if (a.clamp) {
                const float c = *a.clamp;
                for (int k = 0; k < 3; ++k) tg[k] = tg[k] < c ? tg[k] : c;
            }
        }
    }
    if (!valid)
        for (int k = 0; k < 64; ++k) x[k] = 0.0f;
    store_row(s_h[0] + lane * kLd, x);
    s_t[lane][0] = tg[0];
    s_t[lane][1] = tg[1];
    s_t[lane][2] = tg[2];
    s_t[lane][3] = valid ? 1.0f : 0.0f;
    {
        bf16x8 z;
        for (int e = 0; e < 8; ++e) z[e] = (__bf16)0.0f;
        for (int c = 0; c < 8; ++c) reinterpret_cast<bf16x8*>(s_d + lane * kLd)[c] = z;
    }
    wave_fence();
    const bf16x8* fwd = reinterpret_cast<const bf16x8*>(a.fwd);
    const bf16x8* bwd = reinterpret_cast<const bf16x8*>(a.bwd);
    for (int l = 0; l < 4; ++l) hidden_layer(fwd, l, s_h[l], s_h[l + 1], lane);
    f32x4 o[4];
    output_layer(fwd, s_h[4], o, lane);

    // loss: mean over samples and channels of (y - t)^2 / (y^2 + eps), y^2 + eps held constant
    float lsum = 0.0f, nval = 0.0f;
    if (lane < 16) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = 16 * r + lane;
            if (s_t[row][3] == 0.0f) continue;
            bf16x4 gv;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float y = o[r][c], d = y - s_t[row][c], den = y * y + kNrcLossEps;
                lsum += d * d / den;
                gv[c] = (__bf16)(2.0f * d / den * a.scale);
            }
            gv[3] = (__bf16)0.0f;
            *reinterpret_cast<bf16x4*>(s_d + row * kLd) = gv;
            nval += 1.0f;
        }
    }
    wave_fence();
    float* grad = a.grad_part + (size_t)chunk * kNrcParams;

    // output layer: dW4[k][i] = sum_j H4[j][k] dY[j][i]  (k < 64, i < 16)
    {
        bf16x8 bv[2];
#pragma unroll
        for (int s = 0; s < 2; ++s) bv[s] = gather_col(s_d, 32 * s + 8 * g, rl);
#pragma unroll
        for (int tm = 0; tm < 4; ++tm) {
            f32x4 c = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int s = 0; s < 2; ++s) c = mfma(gather_col(s_h[4], 32 * s + 8 * g, 16 * tm + rl), bv[s], c);
#pragma unroll
            for (int e = 0; e < 4; ++e) grad[4 * 4096 + (16 * tm + 4 * g + e) * 16 + rl] = c[e];
        }
    }
    // dH4 = dY W4^T: the output gradient has 16 columns (zero beyond 3), so one
    // 32-deep step whose upper half reads zeros on both sides
    {
        f32x4 dh[4][4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const bf16x8 yb = *reinterpret_cast<const bf16x8*>(s_d + (16 * r + rl) * kLd + 8 * g);
#pragma unroll
            for (int tp = 0; tp < 4; ++tp) dh[r][tp] = mfma(bwd[bwd_frag(4, tp, 0, lane)], yb, f32x4{0.0f, 0.0f, 0.0f, 0.0f});
        }
        wave_fence();
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int tp = 0; tp < 4; ++tp) {
                const int off = (16 * r + rl) * kLd + 16 * tp + 4 * g;
                const bf16x4 hv = *reinterpret_cast<const bf16x4*>(s_h[4] + off);
                bf16x4 v;
#pragma unroll
                for (int e = 0; e < 4; ++e) v[e] = (__bf16)((float)hv[e] > 0.0f ? dh[r][tp][e] : 0.0f);
                *reinterpret_cast<bf16x4*>(s_d + off) = v;
            }
        wave_fence();
    }
    // hidden layers 3 .. 0
    for (int l = 3; l >= 0; --l) {
        const __bf16* H = s_h[l];
        {
            bf16x8 bv[4][2];
#pragma unroll
            for (int tn = 0; tn < 4; ++tn)
#pragma unroll
                for (int s = 0; s < 2; ++s) bv[tn][s] = gather_col(s_d, 32 * s + 8 * g, 16 * tn + rl);
#pragma unroll
            for (int tm = 0; tm < 4; ++tm) {
                bf16x8 av[2];
#pragma unroll
                for (int s = 0; s < 2; ++s) av[s] = gather_col(H, 32 * s + 8 * g, 16 * tm + rl);
#pragma unroll
                for (int tn = 0; tn < 4; ++tn) {
                    f32x4 c = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                    for (int s = 0; s < 2; ++s) c = mfma(av[s], bv[tn][s], c);
#pragma unroll
                    for (int e = 0; e < 4; ++e) grad[l * 4096 + (16 * tm + 4 * g + e) * 64 + 16 * tn + rl] = c[e];
                }
            }
        }
        if (l == 0) break;
        f32x4 dh[4][4];
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int tp = 0; tp < 4; ++tp) dh[r][tp] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            bf16x8 w[4];
#pragma unroll
            for (int tp = 0; tp < 4; ++tp) w[tp] = bwd[bwd_frag(l, tp, s, lane)];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const bf16x8 yb = *reinterpret_cast<const bf16x8*>(s_d + (16 * r + rl) * kLd + 32 * s + 8 * g);
#pragma unroll
                for (int tp = 0; tp < 4; ++tp) dh[r][tp] = mfma(w[tp], yb, dh[r][tp]);
            }
        }
        wave_fence();
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int tp = 0; tp < 4; ++tp) {
                const int off = (16 * r + rl) * kLd + 16 * tp + 4 * g;
                const bf16x4 hv = *reinterpret_cast<const bf16x4*>(H + off);
                bf16x4 v;
#pragma unroll
                for (int e = 0; e < 4; ++e) v[e] = (__bf16)((float)hv[e] > 0.0f ? dh[r][tp][e] : 0.0f);
                *reinterpret_cast<bf16x4*>(s_d + off) = v;
            }
        wave_fence();
    }
    for (int off = 32; off >= 1; off >>= 1) {
        lsum += __shfl_xor(lsum, off);
        nval += __shfl_xor(nval, off);
    }
    if (lane == 0) {
        a.loss_part[2 * (size_t)chunk] = lsum;
        a.loss_part[2 * (size_t)chunk + 1] = nval;
    }
}

// One thread per parameter: gradient = sum of the chunk partials in chunk order; Adam
// update of the FP32 master weight; BF16 copies into both fragment layouts.
__global__ __launch_bounds__(256) void k_nrc_adam(NrcAdamArgs a) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= kNrcParams) return;
    int l, k, i;
    if (p < 4 * 4096) {
        l = p >> 12;
        k = (p >> 6) & 63;
        i = p & 63;
    } else {
        l = 4;
        k = (p - 4 * 4096) >> 4;
        i = (p - 4 * 4096) & 15;
    }
    float w = a.w[p];
    if (a.update) {
        float g = 0.0f;
        for (uint32_t c = 0; c < a.chunks; ++c) g += a.grad_part[(size_t)c * kNrcParams + p];
        const float m = a.b1 * a.m[p] + (1.0f - a.b1) * g;
        const float v = a.b2 * a.v[p] + (1.0f - a.b2) * (g * g);
        a.m[p] = m;
        a.v[p] = v;
        w = w - a.lr * (m * a.c1) / (sqrtf(v * a.c2) + a.eps);
    }
    if (l == 4 && i >= kNrcOut) w = 0.0f;   // padding outputs stay zero
    a.w[p] = w;
    const __bf16 h = (__bf16)w;
    // forward: A = W^T, lane (i % 16) + 16 ((k % 32) / 8), element k % 8
    const int fl = (i & 15) + 16 * ((k & 31) >> 3);
    const int fi = (l < 4 ? l * 4096 + (((i >> 4) * 2 + (k >> 5)) * 64 + fl) * 8
                          : 4 * 4096 + ((k >> 5) * 64 + fl) * 8) + (k & 7);
    reinterpret_cast<__bf16*>(a.fwd)[fi] = h;
    // backward: A = W, lane (k % 16) + 16 ((i % 32) / 8), element i % 8
    const int bl = (k & 15) + 16 * ((i & 31) >> 3);
    const int bi = (l < 4 ? l * 4096 + (((k >> 4) * 2 + (i >> 5)) * 64 + bl) * 8
                          : 4 * 4096 + ((k >> 4) * 64 + bl) * 8) + (i & 7);
    reinterpret_cast<__bf16*>(a.bwd)[bi] = h;
}

}  // namespace

void launch_nrc_infer(const NrcInferArgs& a, hipStream_t s) {
    if (a.num == 0) return;
    const uint32_t blocks = (a.num + 255u) / 256u;
    hipLaunchKernelGGL(k_nrc_infer, dim3(blocks), dim3(256), 0, s, a);
}

void launch_nrc_train(const NrcTrainArgs& a, hipStream_t s) {
    if (a.rows == 0) return;
    hipLaunchKernelGGL(k_nrc_train, dim3((a.rows + 63u) / 64u), dim3(64), 0, s, a);
}

void launch_nrc_adam(const NrcAdamArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_nrc_adam, dim3((kNrcParams + 255) / 256), dim3(256), 0, s, a);
}

bi_pt_status nrc_ensure_query(bi_pt_nrc* nrc, uint32_t cap) {
    if (nrc->query.n >= cap) return BI_PT_OK;
    return dalloc(nrc, nrc->query, cap);
}

namespace {

bi_pt_status ensure_grad(bi_pt_nrc* nrc, uint32_t chunks) {
    if (nrc->grad_chunks >= chunks) return BI_PT_OK;
    BI_TRY(dalloc(nrc, nrc->grad_part, (size_t)chunks * kNrcParams));
    const uint32_t loss_chunks = chunks * std::max(nrc->train_steps, 1u);
    BI_TRY(dalloc(nrc, nrc->loss_part, (size_t)2 * std::max(loss_chunks, chunks)));
    nrc->grad_chunks = chunks;
    return BI_PT_OK;
}

// One Adam step on rows [row0, row0 + rows) of rec / target; loss partials at loss_off chunks.
bi_pt_status train_step(bi_pt_nrc* nrc, const NrcRecord* rec, const float* target, uint32_t row0, uint32_t rows,
                        uint32_t loss_off, const float* clamp, hipStream_t s) {
    const uint32_t chunks = (rows + 63u) / 64u;
    NrcTrainArgs ta{};
    ta.rec = rec;
    ta.target = target;
    ta.row0 = row0;
    ta.rows = rows;
    ta.fwd = nrc->fwd.p;
    ta.bwd = nrc->bwd.p;
    ta.bounds = nrc->bounds.p;
    ta.grad_part = nrc->grad_part.p;
    ta.loss_part = nrc->loss_part.p + 2 * (size_t)loss_off;
    ta.scale = 1.0f / (3.0f * (float)rows);
    ta.clamp = clamp;
    launch_nrc_train(ta, s);
    nrc->steps++;
    NrcAdamArgs aa{};
    aa.w = nrc->w.p;
    aa.m = nrc->m.p;
    aa.v = nrc->v.p;
    aa.fwd = nrc->fwd.p;
    aa.bwd = nrc->bwd.p;
    aa.grad_part = nrc->grad_part.p;
    aa.chunks = chunks;
    aa.lr = nrc->lr;
    aa.b1 = nrc->b1;
    aa.b2 = nrc->b2;
    aa.eps = nrc->eps;
    aa.c1 = (float)(1.0 / (1.0 - std::pow((double)nrc->b1, (double)nrc->steps)));
    aa.c2 = (float)(1.0 / (1.0 - std::pow((double)nrc->b2, (double)nrc->steps)));
    aa.update = 1;
    launch_nrc_adam(aa, s);
    BI_HIP_TRY(nrc, hipGetLastError());
    return BI_PT_OK;
}

bi_pt_status pack(bi_pt_nrc* nrc, hipStream_t s) {
    NrcAdamArgs aa{};
    aa.w = nrc->w.p;
    aa.m = nrc->m.p;
    aa.v = nrc->v.p;
    aa.fwd = nrc->fwd.p;
    aa.bwd = nrc->bwd.p;
    aa.update = 0;
    launch_nrc_adam(aa, s);
    BI_HIP_TRY(nrc, hipGetLastError());
    return BI_PT_OK;
}

}  // namespace

bi_pt_status nrc_train_frame(bi_pt_nrc* nrc, hipStream_t s) {
    const uint32_t per = nrc->train_samples / nrc->train_steps;
    const uint32_t chunks = (per + 63u) / 64u;
    BI_TRY(ensure_grad(nrc, chunks));
    const float* clamp = nullptr;
    if (nrc->clamp_factor > 0.0f) {   // the frame's target bound, before its steps
        const uint32_t blocks = (nrc->train_samples + 255u) / 256u;
        BI_TRY(dalloc(nrc, nrc->clamp_part, (size_t)2 * blocks + 1));
        float* bound = nrc->clamp_part.p + 2 * (size_t)blocks;
        hipLaunchKernelGGL(k_target_lum, dim3(blocks), dim3(256), 0, s, nrc->train_rec.p, nrc->targets.p,
                           nrc->train_samples, nrc->clamp_part.p);
        hipLaunchKernelGGL(k_target_bound, dim3(1), dim3(256), 0, s, nrc->clamp_part.p, blocks, nrc->clamp_factor,
                           bound);
        clamp = bound;
    }
    for (uint32_t k = 0; k < nrc->train_steps; ++k)
        BI_TRY(train_step(nrc, nrc->train_rec.p, nrc->targets.p, k * per, per, k * chunks, clamp, s));
    nrc->last_chunks = chunks * nrc->train_steps;
    nrc->last_grad_chunks = chunks;
    return BI_PT_OK;
}

}  // namespace bipt

using namespace bipt;

extern "C" {

bi_pt_status bi_pt_nrc_create(bi_pt_ctx* ctx, const bi_pt_nrc_desc* desc, bi_pt_nrc** out) {
    if (!out || !ctx) return BI_PT_ERROR_INVALID_ARGUMENT;
    *out = nullptr;
    bi_pt_nrc* n = new (std::nothrow) bi_pt_nrc();
    if (!n) return BI_PT_ERROR_OUT_OF_MEMORY;
    n->ctx = ctx;
    n->device = ctx->device;
    uint32_t seed = 0;
    if (desc) {
        if (desc->train_samples) n->train_samples = desc->train_samples;
        if (desc->train_steps) n->train_steps = desc->train_steps;
        if (desc->cache_bounce) n->cache_bounce = desc->cache_bounce;
        if (desc->learning_rate > 0.0f) n->lr = desc->learning_rate;
        n->use_after = desc->use_after_steps;
        if (desc->target_clamp != 0.0f) n->clamp_factor = desc->target_clamp;
        n->flags = desc->flags;
        seed = desc->seed;
    }
    if (n->train_samples % (64u * n->train_steps) != 0 || n->train_samples > (1u << 24)) {
        delete n;
        return BI_PT_ERROR_INVALID_ARGUMENT;
    }
    auto bad = [&](bi_pt_status st) {
        dfree(n, n->w); dfree(n, n->m); dfree(n, n->v); dfree(n, n->targets); dfree(n, n->grad_part);
        dfree(n, n->loss_part); dfree(n, n->clamp_part); dfree(n, n->bounds); dfree(n, n->fwd); dfree(n, n->bwd); dfree(n, n->train_rec);
        dfree(n, n->query);
        delete n;
        return st;
    };
    if (hipSetDevice(n->device) != hipSuccess) return bad(BI_PT_ERROR_HIP);
    bi_pt_status st = BI_PT_OK;
    if ((st = dalloc(n, n->w, kNrcParams)) != BI_PT_OK || (st = dalloc(n, n->m, kNrcParams)) != BI_PT_OK ||
        (st = dalloc(n, n->v, kNrcParams)) != BI_PT_OK || (st = dalloc(n, n->fwd, kNrcFwd)) != BI_PT_OK ||
        (st = dalloc(n, n->bwd, kNrcBwd)) != BI_PT_OK || (st = dalloc(n, n->bounds, 6)) != BI_PT_OK ||
        (st = dalloc(n, n->train_rec, n->train_samples)) != BI_PT_OK ||
        (st = dalloc(n, n->targets, (size_t)3 * n->train_samples)) != BI_PT_OK)
        return bad(st);
    // He-uniform initial weights from PCG32 (seed, stream 0x4E5243); padding outputs zero
    std::vector<float> w(kNrcParams, 0.0f);
    Pcg r = pcg_keyed(seed, 0u, 0u, 0x43u);
    for (int p = 0; p < kNrcParams; ++p) {
        const bool out_layer = p >= 4 * 4096;
        const int i = out_layer ? (p - 4 * 4096) & 15 : p & 63;
        const float bound = std::sqrt(6.0f / 64.0f);
        const float u = pcg_float(r);
        if (!out_layer || i < kNrcOut) w[p] = (2.0f * u - 1.0f) * bound;
    }
    const float unit_box[6] = {0.0f, 0.0f, 0.0f, 1.0f, 1.0f, 1.0f};
    if (hipMemcpy(n->w.p, w.data(), kNrcParams * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(n->m.p, 0, kNrcParams * 4) != hipSuccess || hipMemset(n->v.p, 0, kNrcParams * 4) != hipSuccess ||
        hipMemset(n->fwd.p, 0, kNrcFwd * 2) != hipSuccess || hipMemset(n->bwd.p, 0, kNrcBwd * 2) != hipSuccess ||
        hipMemcpy(n->bounds.p, unit_box, 24, hipMemcpyHostToDevice) != hipSuccess)
        return bad(BI_PT_ERROR_HIP);
    if (pack(n, nullptr) != BI_PT_OK || hipDeviceSynchronize() != hipSuccess) return bad(BI_PT_ERROR_HIP);
    *out = n;
    return BI_PT_OK;
}

bi_pt_status bi_pt_nrc_destroy(bi_pt_nrc* n) {
    if (!n) return BI_PT_ERROR_INVALID_ARGUMENT;
    (void)hipSetDevice(n->device);
    (void)hipDeviceSynchronize();
    dfree(n, n->w); dfree(n, n->m); dfree(n, n->v); dfree(n, n->targets); dfree(n, n->grad_part);
    dfree(n, n->loss_part); dfree(n, n->clamp_part); dfree(n, n->bounds); dfree(n, n->fwd); dfree(n, n->bwd); dfree(n, n->train_rec);
    dfree(n, n->query);
    delete n;
    return BI_PT_OK;
}

const char* bi_pt_nrc_last_error(const bi_pt_nrc* n) { return n ? n->err.c_str() : "null cache"; }

bi_pt_status bi_pt_render_nrc(bi_pt_nrc* n, const bi_pt_camera* camera, const bi_pt_settings* settings,
                              float* out_rgba32f, float* out_depth, bi_pt_stream stream) {
    if (!n) return BI_PT_ERROR_INVALID_ARGUMENT;
    if (!camera || !settings || !out_rgba32f) return fail(n, BI_PT_ERROR_INVALID_ARGUMENT, "null argument");
    BI_HIP_TRY(n, hipSetDevice(n->device));
    bi_pt_ctx* ctx = n->ctx;
    const bool train = !(n->flags & BI_PT_NRC_FREEZE);
    const bool use = n->steps >= n->use_after;
    bi_pt_settings tst = *settings;
    tst.spp = 1;
    tst.accumulate = 0;
    tst.flags = (settings->flags & BI_PT_FLAG_NO_SKY_NEE) | BI_PT_FLAG_JITTER;
    tst.seed = (uint32_t)(settings->seed * 0x9E3779B1u) ^ (uint32_t)(n->frames * 0x85EBCA77u + 0x7A1Bu);
    JobIn jobs[2];
    jobs[0] = JobIn{ctx, camera, settings, out_rgba32f, out_depth};
    jobs[0].role = use ? kRoleCached : kRoleImage;
    jobs[1] = JobIn{ctx, camera, &tst, nullptr, nullptr};
    jobs[1].role = kRoleTrain;
    RenderOpts opt;
    opt.knobs = Knobs{settings->wave_paths, settings->stack_lds, settings->refill_lanes, settings->work_parts,
                      settings->block_threads};
    opt.reuse_bvh = (settings->flags & BI_PT_FLAG_REUSE_BVH) != 0;
    opt.nrc = n;
    BI_TRY(render_jobs(n, ctx->eng, jobs, train ? 2u : 1u, reinterpret_cast<hipStream_t>(stream), opt, nullptr));
    n->frames++;
    return BI_PT_OK;
}

bi_pt_status bi_pt_nrc_get_stats(bi_pt_nrc* n, bi_pt_nrc_stats* out) {
    if (!n || !out) return BI_PT_ERROR_INVALID_ARGUMENT;
    BI_HIP_TRY(n, hipSetDevice(n->device));
    BI_HIP_TRY(n, hipDeviceSynchronize());
    *out = bi_pt_nrc_stats{};
    out->frames = n->frames;
    out->steps = n->steps;
    if (n->last_chunks) {
        std::vector<float> lp(2 * (size_t)n->last_chunks);
        BI_HIP_TRY(n, hipMemcpy(lp.data(), n->loss_part.p, lp.size() * 4, hipMemcpyDeviceToHost));
        double l = 0.0, c = 0.0;
        for (uint32_t k = 0; k < n->last_chunks; ++k) {
            l += lp[2 * k];
            c += lp[2 * k + 1];
        }
        out->last_loss = c > 0.0 ? (float)(l / (3.0 * c)) : 0.0f;
        out->last_valid_samples = (uint64_t)c;
    }
    return BI_PT_OK;
}

bi_pt_status bi_pt_nrc_get_params(bi_pt_nrc* n, float* host, uint32_t count) {
    if (!n || !host || count != (uint32_t)kNrcParams) return BI_PT_ERROR_INVALID_ARGUMENT;
    BI_HIP_TRY(n, hipSetDevice(n->device));
    BI_HIP_TRY(n, hipDeviceSynchronize());
    BI_HIP_TRY(n, hipMemcpy(host, n->w.p, kNrcParams * 4, hipMemcpyDeviceToHost));
    return BI_PT_OK;
}

bi_pt_status bi_pt_nrc_set_params(bi_pt_nrc* n, const float* host, uint32_t count) {
    if (!n || !host || count != (uint32_t)kNrcParams) return BI_PT_ERROR_INVALID_ARGUMENT;
    BI_HIP_TRY(n, hipSetDevice(n->device));
    BI_HIP_TRY(n, hipDeviceSynchronize());
    BI_HIP_TRY(n, hipMemcpy(n->w.p, host, kNrcParams * 4, hipMemcpyHostToDevice));
    BI_HIP_TRY(n, hipMemset(n->m.p, 0, kNrcParams * 4));
    BI_HIP_TRY(n, hipMemset(n->v.p, 0, kNrcParams * 4));
    n->steps = 0;
    BI_TRY(pack(n, nullptr));
    BI_HIP_TRY(n, hipDeviceSynchronize());
    return BI_PT_OK;
}

bi_pt_status bi_pt_nrc_set_bounds(bi_pt_nrc* n, const float lo_hi[6]) {
    if (!n || !lo_hi) return BI_PT_ERROR_INVALID_ARGUMENT;
    BI_HIP_TRY(n, hipSetDevice(n->device));
    BI_HIP_TRY(n, hipMemcpy(n->bounds.p, lo_hi, 24, hipMemcpyHostToDevice));
    return BI_PT_OK;
}

bi_pt_status bi_pt_nrc_infer(bi_pt_nrc* n, const void* records, uint32_t count, float* out_rgb, bi_pt_stream stream) {
    if (!n || (count && (!records || !out_rgb))) return BI_PT_ERROR_INVALID_ARGUMENT;
    BI_HIP_TRY(n, hipSetDevice(n->device));
    NrcInferArgs a{};
    a.rec = static_cast<const NrcRecord*>(records);
    a.num = count;
    a.fwd = n->fwd.p;
    a.bounds = n->bounds.p;
    a.raw_out = out_rgb;
    launch_nrc_infer(a, reinterpret_cast<hipStream_t>(stream));
    BI_HIP_TRY(n, hipGetLastError());
    return BI_PT_OK;
}

bi_pt_status bi_pt_nrc_train(bi_pt_nrc* n, const void* records, const float* targets, uint32_t count,
                             bi_pt_stream stream) {
    if (!n || !count || !records || !targets) return BI_PT_ERROR_INVALID_ARGUMENT;
    if (count > (1u << 24)) return fail(n, BI_PT_ERROR_INVALID_ARGUMENT, "at most 2^24 samples per step");
    BI_HIP_TRY(n, hipSetDevice(n->device));
    const uint32_t chunks = (count + 63u) / 64u;
    BI_TRY(ensure_grad(n, chunks));
    BI_TRY(train_step(n, static_cast<const NrcRecord*>(records), targets, 0, count, 0, nullptr,
                      reinterpret_cast<hipStream_t>(stream)));
    n->last_chunks = chunks;
    n->last_grad_chunks = chunks;
    return BI_PT_OK;
}

bi_pt_status bi_pt_nrc_debug_samples(bi_pt_nrc* n, float* records, float* targets, uint32_t count) {
    if (!n || !records || !targets || count > n->train_samples) return BI_PT_ERROR_INVALID_ARGUMENT;
    BI_HIP_TRY(n, hipSetDevice(n->device));
    BI_HIP_TRY(n, hipDeviceSynchronize());
    BI_HIP_TRY(n, hipMemcpy(records, n->train_rec.p, (size_t)count * sizeof(NrcRecord), hipMemcpyDeviceToHost));
    BI_HIP_TRY(n, hipMemcpy(targets, n->targets.p, (size_t)count * 3 * sizeof(float), hipMemcpyDeviceToHost));
    return BI_PT_OK;
}

bi_pt_status bi_pt_nrc_debug_gradient(bi_pt_nrc* n, float* host, uint32_t count) {
    if (!n || !host || count != (uint32_t)kNrcParams) return BI_PT_ERROR_INVALID_ARGUMENT;
    BI_HIP_TRY(n, hipSetDevice(n->device));
    BI_HIP_TRY(n, hipDeviceSynchronize());
    const uint32_t chunks = n->last_grad_chunks;
    std::vector<float> part((size_t)chunks * kNrcParams);
    if (chunks) BI_HIP_TRY(n, hipMemcpy(part.data(), n->grad_part.p, part.size() * 4, hipMemcpyDeviceToHost));
    for (int p = 0; p < kNrcParams; ++p) {
        float g = 0.0f;
        for (uint32_t c = 0; c < chunks; ++c) g += part[(size_t)c * kNrcParams + p];
        host[p] = g;
    }
    return BI_PT_OK;
}

}  // extern "C"
