// Learned first guess for the saturation adjustment: an MLP
// (s, qt, p0) -> T with layers 3 -> 64 -> 64 -> 1, tanh hidden units.
// Weights are stored as bf16 (the trained FP32 master weights rounded), biases
// and the input/output normalisation as FP32.  The thermodynamics kernel
// stages the weights in LDS once per block and evaluates the network for the
// cells that reach the saturated branch.
#pragma once
#include <hip/hip_bf16.h>

#include "common.h"

namespace mlp {

constexpr int NIN = 3;
constexpr int NH = 64;

// Device-side weight block (one allocation, filled by pycles_mlp_pack).
struct Weights {
    float mu[NIN], sd[NIN];  // input normalisation
    float mu_T, sd_T;        // output de-normalisation
    float b1[NH], b2[NH], b3;
    __hip_bfloat16 W1[NH * NIN];  // [out][in]
    __hip_bfloat16 W2[NH * NH];   // [out][in]
    __hip_bfloat16 W3[NH];
};

struct Smem {
    float W1[NH * NIN];
    float b1[NH];
    float W2t[NH * NH];  // [in][out]
    float b2[NH];
    float W3[NH];
    float b3, mu[NIN], sd[NIN], mu_T, sd_T;
};

__device__ __forceinline__ float tanh_f(float x) {
    // tanh(x) = 1 - 2/(exp(2x)+1), saturating for large |x|
    const float e = __expf(2.0f * fminf(fmaxf(x, -15.0f), 15.0f));
    return 1.0f - 2.0f / (e + 1.0f);
}

__device__ __forceinline__ void stage(const Weights *w, Smem *m, int tid, int nthr) {
    for (int t = tid; t < NH * NIN; t += nthr) m->W1[t] = __bfloat162float(w->W1[t]);
    for (int t = tid; t < NH * NH; t += nthr) {
        const int o = t / NH, i = t % NH;
        m->W2t[i * NH + o] = __bfloat162float(w->W2[t]);
    }
    for (int t = tid; t < NH; t += nthr) {
        m->b1[t] = w->b1[t];
        m->b2[t] = w->b2[t];
        m->W3[t] = __bfloat162float(w->W3[t]);
    }
    if (tid == 0) {
        m->b3 = w->b3;
        for (int i = 0; i < NIN; ++i) {
            m->mu[i] = w->mu[i];
            m->sd[i] = w->sd[i];
        }
        m->mu_T = w->mu_T;
        m->sd_T = w->sd_T;
    }
}

// Forward pass for one cell (FP32 arithmetic on bf16-rounded weights).
__device__ __forceinline__ float guess(const Smem *m, float s, float qt, float p0) {
    const float x0 = (s - m->mu[0]) / m->sd[0];
    const float x1 = (qt - m->mu[1]) / m->sd[1];
    const float x2 = (p0 - m->mu[2]) / m->sd[2];
    float h1[NH];
#pragma unroll
    for (int o = 0; o < NH; ++o)
        h1[o] = tanh_f(m->b1[o] + m->W1[o * NIN] * x0 + m->W1[o * NIN + 1] * x1 + m->W1[o * NIN + 2] * x2);
    float acc[NH];
#pragma unroll
    for (int o = 0; o < NH; ++o) acc[o] = m->b2[o];
    for (int i = 0; i < NH; ++i) {
        const float hi = h1[i];
#pragma unroll
        for (int o = 0; o < NH; ++o) acc[o] += m->W2t[i * NH + o] * hi;
    }
    float y = m->b3;
#pragma unroll
    for (int o = 0; o < NH; ++o) y += m->W3[o] * tanh_f(acc[o]);
    return m->mu_T + m->sd_T * y;
}

}  // namespace mlp
