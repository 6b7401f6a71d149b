// Learned first guess for the saturation adjustment: an MLP
// (s, qt, p0) -> T with layers 3 -> 64 -> 64 -> 1 and tanh hidden units.
//
// Training keeps FP32 master weights in one flat vector (offsets below).
// Inference uses the packed block `Weights` and runs all three layers for the
// 64 cells of a wavefront at once, in the transposed form H^T = W X^T: hidden
// units are the rows of the v_mfma_f32_32x32x16_bf16 products and cells their
// columns.  The weights are then the A operands, the same in every call, and
// the tanh of one layer's result registers is the next layer's B operand as
// it stands (the K order of the hidden layer follows the result layout).  The
// VALU evaluates only the activations and the output dot product.
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

// tanh(x) = 1 - 2 / (2^(TANH_C x) + 1): the pre-activations are formed
// already scaled by TANH_C = 2 log2(e).
constexpr float TANH_C = 2.8853900817779268f;

// Result layout of the 32x32 product: D[i][j] (i = hidden unit, j = cell)
// sits in lane j % 32 + 32 ((i / 4) % 2), register (i / 8) * 4 + i % 4.  So
// lane (c, h) = (lane % 32, lane / 32), register r of a 32-unit block nb holds
// unit 32 nb + d_unit(h, r):
__host__ __device__ constexpr int d_unit(int h, int r) { return 8 * (r / 4) + 4 * h + r % 4; }
// and the hidden layer reads its K index 8 h + e of step q (four steps of 16)
// from register 8 (q % 2) + e of block q / 2, i.e. from unit
__host__ __device__ constexpr int k_unit(int q, int h, int e) { return 32 * (q / 2) + d_unit(h, 8 * (q % 2) + e); }

// Packed inference block (lane = 64 lanes of a wavefront, h = lane / 32).
//  A1[nb][lane]: for h = 0, the eight bf16 (W1[u][0], W1[u][0], W1[u][1],
//     W1[u][1], W1[u][2], W1[u][2], bh, bl) of unit u = 32 nb + lane % 32, with
//     bh + bl = TANH_C b1[u] split in two bf16; zero for h = 1.  The matching
//     B operand holds (y0h, y0l, y1h, y1l, y2h, y2l, 1, 1), yh + yl = TANH_C x.
//  C2[nb][h][r]: TANH_C b2[32 nb + d_unit(h, r)], the hidden layer's start value.
//  A2[nb][q][lane]: the eight bf16 of TANH_C W2[32 nb + lane % 32][k_unit(q, h, e)].
//  W3n[nb][h][r]: -2 bf16(W3[32 nb + d_unit(h, r)]); with b3s = b3 + sum of the
//     bf16 W3, y = b3s + sum_u W3n[u] / (2^z_u + 1) = b3 + sum_u bf16(W3[u]) tanh.
struct Weights {
    Norm nm;
    float b3s;
    alignas(16) bf16x8 A1[2][64];
    alignas(16) float C2[2][2][16];
    alignas(16) float W3n[2][2][16];
    alignas(16) bf16x8 A2[2][4][64];
};

// 1 / (2^u + 1); tanh(x) = 1 - 2 rcp_exp1(TANH_C x).  Saturates through
// 2^u = inf or 0, with no clamp.
__device__ __forceinline__ float rcp_exp1(float u) {
    return __builtin_amdgcn_rcpf(__builtin_amdgcn_exp2f(u) + 1.0f);
}

// Copy the packed block into LDS (all threads of the block).
__device__ __forceinline__ void stage(const Weights *src, Weights *dst, int tid, int nthr) {
    const uint4 *s = reinterpret_cast<const uint4 *>(src);
    uint4 *d = reinterpret_cast<uint4 *>(dst);
    for (int t = tid; t < (int)(sizeof(Weights) / 16); t += nthr) d[t] = s[t];
}

// The network for the 64 cells of this wavefront, lane = cell; (x0, x1, x2)
// are this lane's normalised inputs.  All 64 lanes must call it together.
__device__ __forceinline__ float guess_wave(const Weights *w, float x0, float x1, float x2) {
    const int lane = __lane_id();
    const int h = lane >> 5;
    const int c = lane & 31;
    float out = 0.0f;
#pragma unroll
    for (int m = 0; m < 2; ++m) {
        // the weights are read from LDS again for each half: kept in
        // registers across both, the hidden layer's 64 would spill
        asm volatile("" ::: "memory");
        // cells 32 m + c: their inputs, scaled and split, in the lanes h = 0
        const int src = 32 * m + c;
        const float y[3] = {TANH_C * __shfl(x0, src, 64), TANH_C * __shfl(x1, src, 64),
                            TANH_C * __shfl(x2, src, 64)};
        bf16x8 B1;
#pragma unroll
        for (int v = 0; v < 3; ++v) {
            const __bf16 hi = (__bf16)y[v];
            const __bf16 lo = (__bf16)(y[v] - (float)hi);
            B1[2 * v] = h ? (__bf16)0.0f : hi;
            B1[2 * v + 1] = h ? (__bf16)0.0f : lo;
        }
        B1[6] = h ? (__bf16)0.0f : (__bf16)1.0f;
        B1[7] = B1[6];
        bf16x8 B2[4];
#pragma unroll
        for (int nb = 0; nb < 2; ++nb) {
            const f32x16 z = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w->A1[nb][lane], B1, f32x16{}, 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 16; ++r)
                B2[2 * nb + r / 8][r % 8] = (__bf16)__builtin_fmaf(-2.0f, rcp_exp1(z[r]), 1.0f);
        }
        float part[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int nb = 0; nb < 2; ++nb) {
            const float4 *cz = reinterpret_cast<const float4 *>(w->C2[nb][h]);
            f32x16 z;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const float4 t = cz[v];
                z[4 * v] = t.x;
                z[4 * v + 1] = t.y;
                z[4 * v + 2] = t.z;
                z[4 * v + 3] = t.w;
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) z = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w->A2[nb][q][lane], B2[q], z, 0, 0, 0);
            const float4 *w3 = reinterpret_cast<const float4 *>(w->W3n[nb][h]);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const float4 t = w3[v];
                part[0] = __builtin_fmaf(t.x, rcp_exp1(z[4 * v]), part[0]);
                part[1] = __builtin_fmaf(t.y, rcp_exp1(z[4 * v + 1]), part[1]);
                part[2] = __builtin_fmaf(t.z, rcp_exp1(z[4 * v + 2]), part[2]);
                part[3] = __builtin_fmaf(t.w, rcp_exp1(z[4 * v + 3]), part[3]);
            }
        }
        // the two lane halves hold the two halves of each cell's sum
        const float half_sum = (part[0] + part[1]) + (part[2] + part[3]);
        const float sum = half_sum + __shfl_xor(half_sum, 32, 64);
        if (h == m) out = w->b3s + sum;
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
