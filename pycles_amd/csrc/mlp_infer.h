// Learned first guess for the saturation adjustment: an MLP
// (s, qt, p0) -> T with layers 3 -> 64 -> 64 -> 1 and tanh hidden units.
//
// Training keeps FP32 master weights in one flat vector (offsets below).
// Inference uses the packed block `Weights`: weights rounded to bf16, biases
// and normalisation in FP32.  The 64x64 hidden layer runs on the matrix cores
// (v_mfma_f32_32x32x16_bf16, FP32 accumulate) for the 64 cells of a wavefront
// at once; the input and output layers and the activations run on the VALU.
#pragma once
#include "common.h"

namespace mlp {

constexpr int NIN = 3;
constexpr int NH = 64;
// flat FP32 parameter vector: W1 [NH][NIN], b1, W2 [NH][NH] ([out][in]), b2, W3, b3
constexpr int OW1 = 0;
constexpr int OB1 = OW1 + NH * NIN;
constexpr int OW2 = OB1 + NH;
constexpr int OB2 = OW2 + NH * NH;
constexpr int OW3 = OB2 + NH;
constexpr int OB3 = OW3 + NH;
constexpr int NPARAM = OB3 + 1;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// x_n = (x - mu) * isd for (s, qt, p0); T = mu_T + sd_T * y
struct Norm {
    float mu[NIN], isd[NIN];
    float mu_T, sd_T;
};

// Packed inference block.  W2B holds the B operands of the hidden layer in
// the order the MFMA reads them: W2B[(n * 4 + kk) * 64 + lane] is the eight
// bf16 W2[32 n + lane % 32][16 kk + 8 (lane / 32) + e], e = 0..7.  The
// pre-activations enter tanh scaled by TANH_C = 2 log2(e) (tanh(x) =
// 1 - 2 / (2^(TANH_C x) + 1)), so W1b and b2c carry that factor.
constexpr float TANH_C = 2.8853900817779268f;
struct Weights {
    Norm nm;
    float4 W1b[NH];      // TANH_C * (W1[k][0..2] rounded to bf16, b1[k]) per hidden unit
    float b2c[NH];       // TANH_C * b2
    float W3[NH];        // rounded to bf16
    float b3;
    alignas(16) bf16x8 W2B[2 * 4 * 64];
};

// Per-wave scratch for the output-layer row sums: 32 rows x 36 (16-byte rows).
constexpr int RED_STRIDE = 36;
constexpr int RED_FLOATS = 32 * RED_STRIDE;

// tanh of x given u = TANH_C * x: 1 - 2 / (2^u + 1).  Saturates to +-1
// through 2^u = inf or 0; no clamp needed.
__device__ __forceinline__ float tanh_scaled(float u) {
    const float e = __builtin_amdgcn_exp2f(u);
    return __builtin_fmaf(-2.0f, __builtin_amdgcn_rcpf(e + 1.0f), 1.0f);
}

// Copy the packed block into LDS (all threads of the block).
__device__ __forceinline__ void stage(const Weights *src, Weights *dst, int tid, int nthr) {
    const uint4 *s = reinterpret_cast<const uint4 *>(src);
    uint4 *d = reinterpret_cast<uint4 *>(dst);
    for (int t = tid; t < (int)(sizeof(Weights) / 16); t += nthr) d[t] = s[t];
}

__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// The network for the 64 cells of this wavefront, lane = cell; (x0, x1, x2)
// are this lane's normalised inputs.  All 64 lanes must call it together.
// red: this wave's RED_FLOATS of LDS scratch (16-byte aligned).
__device__ __forceinline__ float guess_wave(const Weights *w, float *red, float x0, float x1, float x2) {
    const int lane = __lane_id();
    const int half = lane >> 5;
    const int col = lane & 31;
    float out = 0.0f;
#pragma unroll
    for (int m = 0; m < 2; ++m) {
        // this lane builds the A operand rows of cells 32 m + col
        const int src = 32 * m + col;
        const float y0 = __shfl(x0, src, 64);
        const float y1 = __shfl(x1, src, 64);
        const float y2 = __shfl(x2, src, 64);
        bf16x8 A[4];
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const float4 wb = w->W1b[16 * kk + 8 * half + e];
                const float u = __builtin_fmaf(wb.x, y0, __builtin_fmaf(wb.y, y1, __builtin_fmaf(wb.z, y2, wb.w)));
                A[kk][e] = (__bf16)tanh_scaled(u);
            }
        }
        f32x16 acc0 = {}, acc1 = {};
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[kk], w->W2B[(0 * 4 + kk) * 64 + lane], acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[kk], w->W2B[(1 * 4 + kk) * 64 + lane], acc1, 0, 0, 0);
        }
        // D[i][j] sits in lane j % 32 + 32 ((i / 4) % 2), register (i / 8) * 4 + i % 4
        const float b20 = w->b2c[col], b21 = w->b2c[32 + col];
        const float w30 = w->W3[col], w31 = w->W3[32 + col];
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int i = 8 * (r >> 2) + 4 * half + (r & 3);
            const float t0 = tanh_scaled(__builtin_fmaf(acc0[r], TANH_C, b20));
            const float t1 = tanh_scaled(__builtin_fmaf(acc1[r], TANH_C, b21));
            red[i * RED_STRIDE + col] = __builtin_fmaf(w30, t0, w31 * t1);
        }
        wave_sync_lds();
        if (half == m) {
            const float4 *row = reinterpret_cast<const float4 *>(red + col * RED_STRIDE);
            float s0 = w->b3, s1 = 0.0f, s2 = 0.0f, s3 = 0.0f;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const float4 v = row[q];
                s0 += v.x;
                s1 += v.y;
                s2 += v.z;
                s3 += v.w;
            }
            out = (s0 + s1) + (s2 + s3);
        }
        wave_sync_lds();
    }
    return __builtin_fmaf(w->nm.sd_T, out, w->nm.mu_T);
}

__device__ __forceinline__ void normalise(const Weights *w, double s, float qt, float p0, float *x0, float *x1,
                                          float *x2) {
    *x0 = (float)(s - (double)w->nm.mu[0]) * w->nm.isd[0];
    *x1 = (qt - w->nm.mu[1]) * w->nm.isd[1];
    *x2 = (p0 - w->nm.mu[2]) * w->nm.isd[2];
}

}  // namespace mlp
