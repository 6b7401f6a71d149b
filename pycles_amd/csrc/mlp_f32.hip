// The learned first guess of the saturation adjustment: training pairs from
// the FP64 solver, the network's training (forward, backward and Adam on the
// GPU), packing for inference, and two evaluation kernels.
//
// Training pairs (s, qt, p0) -> T: qt ~ U(0, 25 g/kg), p0 ~ U(500, 1020 hPa),
// T ~ U(230, 310 K) give the equilibrium state (saturated where qt > qv*(T)),
// whose entropy (ThermodynamicsSA.pyx:138-157) is the input s; the target is
// what the FP64 saturation adjustment (eos_c, thermodynamics_sa.h:17-66)
// returns for (p0, s, qt).
//
// Training: mean-squared error on the normalised temperature, minibatches of
// `batch` samples drawn without replacement through a per-epoch affine
// permutation of [0, n) (n a power of two), Adam with a cosine learning-rate
// schedule.  Each block of 128 samples writes its partial gradient; a second
// kernel sums the partials in block order and applies Adam, so a run is
// bitwise reproducible.
#include "common.h"
#include "mlp_infer.h"
#include "thermo_funcs.h"

namespace {

using namespace mlp;

constexpr int TB = 128;            // samples per gradient block
constexpr int PSTRIDE = NPARAM + 1;  // partial gradient + loss per block
constexpr int NORM_BLOCKS = 256;

__device__ __forceinline__ unsigned long long mix64(unsigned long long x) {
    x += 0x9E3779B97F4A7C15ull;
    unsigned long long z = x;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static unsigned long long mix64_host(unsigned long long x) {
    x += 0x9E3779B97F4A7C15ull;
    unsigned long long z = x;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ double unit(unsigned long long z) { return (double)(z >> 11) * (1.0 / 9007199254740992.0); }

// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void pairs_kernel(thermo::Table lt, long long n, unsigned long long seed,
                                                    int max_iter, int saturated_only, double *out) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned long long key = mix64(seed);
    double qt, p0, T, qvs;
    // saturated_only: draw again (at most 64 times) until qt > qv*(T), i.e.
    // sample the same box conditioned on saturation
    for (int a = 0;; ++a) {
        const unsigned long long c = ((unsigned long long)i * 64ull + (unsigned long long)a) * 3ull;
        qt = fmax(0.025 * unit(mix64(key ^ c)), 1e-11);
        p0 = 5.0e4 + 5.2e4 * unit(mix64(key ^ (c + 1)));
        T = 230.0 + 80.0 * unit(mix64(key ^ (c + 2)));
        qvs = thermo::qv_star_c(p0, qt, thermo::lookup(lt, T));
        if (!saturated_only || qt > qvs || a == 63) break;
    }
    const double ql = qt > qvs ? qt - qvs : 0.0;
    const double qv = qt - ql;
    const double L = thermo::latent_heat_variable(T);
    const double s = thermo::sd_c(thermo::pd_c(p0, qt, qv), T) * (1.0 - qt) +
                     thermo::sv_c(thermo::pv_c(p0, qt, qv), T) * qt + thermo::sc_c(L, T) * ql;
    double Ts, qvo, qlo, qio;
    thermo::eos_sa(lt, p0, s, qt, &Ts, &qvo, &qlo, &qio, max_iter);
    out[4 * i] = s;
    out[4 * i + 1] = qt;
    out[4 * i + 2] = p0;
    out[4 * i + 3] = Ts;
}

// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void init_kernel(float *P, float *m, float *v, unsigned long long seed) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= NPARAM) return;
    float a = 0.0f;
    if (p < OB1) a = sqrtf(6.0f / (NIN + NH));
    else if (p >= OW2 && p < OB2) a = sqrtf(6.0f / (NH + NH));
    else if (p >= OW3 && p < OB3) a = sqrtf(6.0f / (NH + 1));
    const double u = unit(mix64(mix64(seed) ^ (unsigned long long)p));
    P[p] = a * (float)(2.0 * u - 1.0);
    if (m) m[p] = 0.0f;
    if (v) v[p] = 0.0f;
}

// mean and standard deviation of the four columns (FP64 partial sums per block)
__global__ __launch_bounds__(256) void norm_partial_kernel(const double *data, long long n, double *part) {
    __shared__ double red[256][8];
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const double x = data[4 * i + c];
            acc[c] += x;
            acc[4 + c] += x * x;
        }
    }
    for (int c = 0; c < 8; ++c) red[threadIdx.x][c] = acc[c];
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h)
            for (int c = 0; c < 8; ++c) red[threadIdx.x][c] += red[threadIdx.x + h][c];
        __syncthreads();
    }
    if (threadIdx.x < 8) part[blockIdx.x * 8 + threadIdx.x] = red[0][threadIdx.x];
}

__global__ void norm_final_kernel(const double *part, int nblk, long long n, Norm *nm) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double sum[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int b = 0; b < nblk; ++b)
        for (int c = 0; c < 8; ++c) sum[c] += part[b * 8 + c];
    double mu[4], sd[4];
    for (int c = 0; c < 4; ++c) {
        mu[c] = sum[c] / (double)n;
        sd[c] = sqrt(fmax(sum[4 + c] / (double)n - mu[c] * mu[c], 1e-30));
    }
    for (int c = 0; c < 3; ++c) {
        nm->mu[c] = (float)mu[c];
        nm->isd[c] = (float)(1.0 / sd[c]);
    }
    nm->mu_T = (float)mu[3];
    nm->sd_T = (float)sd[3];
}

__device__ __forceinline__ void load_inputs(const double *d, const Norm &nm, float *x0, float *x1, float *x2,
                                            float *t) {
    *x0 = (float)(d[0] - (double)nm.mu[0]) * nm.isd[0];
    *x1 = ((float)d[1] - nm.mu[1]) * nm.isd[1];
    *x2 = ((float)d[2] - nm.mu[2]) * nm.isd[2];
    *t = (float)((d[3] - (double)nm.mu_T) / (double)nm.sd_T);
}

__device__ __forceinline__ float tanh_exact(float x) { return tanhf(x); }

// ---------------------------------------------------------------------------
// Partial gradient of sum_b 0.5 (y_b - t_b)^2 * inv_batch over this block's
// 128 samples, written to part[blockIdx.x][0..NPARAM) with the loss sum
// (unscaled) at part[blockIdx.x][NPARAM].
__global__ __launch_bounds__(TB) void grad_kernel(const float *__restrict__ P, const Norm *__restrict__ nmp,
                                                  const double *__restrict__ data, unsigned long long n_mask,
                                                  unsigned long long pa, unsigned long long pc, long long start,
                                                  float inv_batch, float *__restrict__ part) {
    __shared__ float sW2t[NH * NH];  // [in][out]
    __shared__ float sW1[NH * NIN], sb1[NH], sb2[NH], sW3[NH];
    __shared__ float sH[TB][NH + 1];
    __shared__ float sD[TB][NH + 1];
    __shared__ float sX[TB][4];
    __shared__ float sL[TB];
    const int t = threadIdx.x;
    for (int q = t; q < NH * NH; q += TB) {
        const int o = q / NH, i = q % NH;
        sW2t[i * NH + o] = P[OW2 + q];
    }
    for (int q = t; q < NH * NIN; q += TB) sW1[q] = P[OW1 + q];
    for (int q = t; q < NH; q += TB) {
        sb1[q] = P[OB1 + q];
        sb2[q] = P[OB2 + q];
        sW3[q] = P[OW3 + q];
    }
    const float b3 = P[OB3];
    const Norm nm = *nmp;
    const unsigned long long g = (unsigned long long)(start + (long long)blockIdx.x * TB + t);
    const unsigned long long idx = (pa * g + pc) & n_mask;
    float x0, x1, x2, tgt;
    load_inputs(data + 4 * idx, nm, &x0, &x1, &x2, &tgt);
    __syncthreads();
    // forward
    float h1[NH], a2[NH];
#pragma unroll
    for (int o = 0; o < NH; ++o) {
        h1[o] = tanh_exact(sb1[o] + sW1[3 * o] * x0 + sW1[3 * o + 1] * x1 + sW1[3 * o + 2] * x2);
        sH[t][o] = h1[o];
        a2[o] = sb2[o];
    }
    for (int i = 0; i < NH; ++i) {
        const float hi = h1[i];
        const float4 *w = reinterpret_cast<const float4 *>(sW2t + i * NH);
#pragma unroll
        for (int o4 = 0; o4 < NH / 4; ++o4) {
            const float4 ww = w[o4];
            a2[4 * o4] += ww.x * hi;
            a2[4 * o4 + 1] += ww.y * hi;
            a2[4 * o4 + 2] += ww.z * hi;
            a2[4 * o4 + 3] += ww.w * hi;
        }
    }
    float y = b3;
#pragma unroll
    for (int o = 0; o < NH; ++o) {
        a2[o] = tanh_exact(a2[o]);  // h2
        y += sW3[o] * a2[o];
    }
    const float e = y - tgt;
    const float dy = e * inv_batch;
    sL[t] = 0.5f * e * e;
    sX[t][3] = dy;
#pragma unroll
    for (int o = 0; o < NH; ++o) sD[t][o] = dy * a2[o];
    __syncthreads();
    float *pp = part + (size_t)blockIdx.x * PSTRIDE;
    if (t < NH) {
        float acc = 0.0f;
        for (int s = 0; s < TB; ++s) acc += sD[s][t];
        pp[OW3 + t] = acc;
    } else if (t == NH) {
        float acc = 0.0f;
        for (int s = 0; s < TB; ++s) acc += sX[s][3];
        pp[OB3] = acc;
    } else if (t == NH + 1) {
        float acc = 0.0f;
        for (int s = 0; s < TB; ++s) acc += sL[s];
        pp[NPARAM] = acc;
    }
    __syncthreads();
    // dz2 = dy W3 (1 - h2^2), kept in a2
#pragma unroll
    for (int o = 0; o < NH; ++o) {
        a2[o] = dy * sW3[o] * (1.0f - a2[o] * a2[o]);
        sD[t][o] = a2[o];
    }
    __syncthreads();
    {
        // dW2[o][i] = sum_s dz2[s][o] h1[s][i]: this thread's 4 x 8 tile
        const int o0 = (t >> 3) * 4, i0 = (t & 7) * 8;
        float acc[4][8];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 8; ++b) acc[a][b] = 0.0f;
        for (int s = 0; s < TB; ++s) {
            float dv[4], hv[8];
#pragma unroll
            for (int a = 0; a < 4; ++a) dv[a] = sD[s][o0 + a];
#pragma unroll
            for (int b = 0; b < 8; ++b) hv[b] = sH[s][i0 + b];
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 8; ++b) acc[a][b] += dv[a] * hv[b];
        }
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 8; ++b) pp[OW2 + (o0 + a) * NH + i0 + b] = acc[a][b];
        if (t < NH) {
            float acc2 = 0.0f;
            for (int s = 0; s < TB; ++s) acc2 += sD[s][t];
            pp[OB2 + t] = acc2;
        }
    }
    __syncthreads();
    // dz1[i] = (sum_o W2[o][i] dz2[o]) (1 - h1[i]^2)
    for (int i = 0; i < NH; ++i) {
        const float4 *w = reinterpret_cast<const float4 *>(sW2t + i * NH);
        float acc = 0.0f;
#pragma unroll
        for (int o4 = 0; o4 < NH / 4; ++o4) {
            const float4 ww = w[o4];
            acc += ww.x * a2[4 * o4] + ww.y * a2[4 * o4 + 1] + ww.z * a2[4 * o4 + 2] + ww.w * a2[4 * o4 + 3];
        }
        sD[t][i] = acc * (1.0f - h1[i] * h1[i]);
    }
    sX[t][0] = x0;
    sX[t][1] = x1;
    sX[t][2] = x2;
    __syncthreads();
    if (t < NH) {
        float db = 0.0f, dw0 = 0.0f, dw1 = 0.0f, dw2 = 0.0f;
        for (int s = 0; s < TB; ++s) {
            const float d = sD[s][t];
            db += d;
            dw0 += d * sX[s][0];
            dw1 += d * sX[s][1];
            dw2 += d * sX[s][2];
        }
        pp[OB1 + t] = db;
        pp[OW1 + 3 * t] = dw0;
        pp[OW1 + 3 * t + 1] = dw1;
        pp[OW1 + 3 * t + 2] = dw2;
    }
}

// Sum the partials in block order; then Adam (m, v non-null) or store (grad).
__global__ __launch_bounds__(256) void reduce_adam_kernel(float *P, float *m, float *v, const float *part, int nblk,
                                                          float lr, float c1, float c2, float *grad,
                                                          double *loss_acc) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > NPARAM) return;
    float gsum = 0.0f;
    for (int b = 0; b < nblk; ++b) gsum += part[(size_t)b * PSTRIDE + p];
    if (p == NPARAM) {
        if (loss_acc) *loss_acc += (double)gsum;
        return;
    }
    if (grad) grad[p] = gsum;
    if (m && v) {
        const float b1 = 0.9f, b2 = 0.999f;
        const float mm = b1 * m[p] + (1.0f - b1) * gsum;
        const float vv = b2 * v[p] + (1.0f - b2) * gsum * gsum;
        m[p] = mm;
        v[p] = vv;
        P[p] -= lr * (mm * c1) / (sqrtf(vv * c2) + 1e-8f);
    }
}

// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void forward_kernel(const float *__restrict__ P, const Norm *__restrict__ nmp,
                                                      const double *__restrict__ data, long long n, float *out) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Norm nm = *nmp;
    float x0, x1, x2, tgt;
    load_inputs(data + 4 * i, nm, &x0, &x1, &x2, &tgt);
    float h1[NH];
#pragma unroll
    for (int o = 0; o < NH; ++o)
        h1[o] = tanh_exact(P[OB1 + o] + P[OW1 + 3 * o] * x0 + P[OW1 + 3 * o + 1] * x1 + P[OW1 + 3 * o + 2] * x2);
    float y = P[OB3];
    for (int o = 0; o < NH; ++o) {
        float a = P[OB2 + o];
#pragma unroll
        for (int q = 0; q < NH; ++q) a += P[OW2 + o * NH + q] * h1[q];
        y += P[OW3 + o] * tanh_exact(a);
    }
    out[i] = nm.mu_T + nm.sd_T * y;
}

__device__ __forceinline__ float bf16_round(float x) { return (float)(__bf16)x; }

__global__ __launch_bounds__(256) void pack_kernel(const float *__restrict__ P, const Norm *__restrict__ nmp,
                                                   Weights *w) {
    const int t = threadIdx.x;
    if (t == 0) {
        w->nm = *nmp;
        float b3s = P[OB3];
        for (int u = 0; u < NH; ++u) b3s += bf16_round(P[OW3 + u]);
        w->b3s = b3s;
    }
    for (int q = t; q < 2 * 64; q += blockDim.x) {
        const int lane = q % 64, nb = q / 64;
        const int u = 32 * nb + lane % 32;
        bf16x8 a = {};
        if (lane < 32) {
            const float cb = TANH_C * P[OB1 + u];
            const __bf16 bh = (__bf16)cb;
#pragma unroll
            for (int v = 0; v < 3; ++v) a[2 * v] = a[2 * v + 1] = (__bf16)P[OW1 + 3 * u + v];
            a[6] = bh;
            a[7] = (__bf16)(cb - (float)bh);
        }
        w->A1[nb][lane] = a;
    }
    for (int q = t; q < 2 * 2 * 16; q += blockDim.x) {
        const int r = q % 16, h = (q / 16) % 2, nb = q / 32;
        const int u = 32 * nb + d_unit(h, r);
        w->C2[nb][h][r] = TANH_C * P[OB2 + u];
        w->W3n[nb][h][r] = -2.0f * bf16_round(P[OW3 + u]);
    }
    for (int q = t; q < 2 * 4 * 64; q += blockDim.x) {
        const int lane = q % 64, qq = (q / 64) % 4, nb = q / 256;
        const int row = 32 * nb + lane % 32, h = lane / 32;
        bf16x8 b;
#pragma unroll
        for (int e = 0; e < 8; ++e) b[e] = (__bf16)(TANH_C * P[OW2 + row * NH + k_unit(qq, h, e)]);
        w->A2[nb][qq][lane] = b;
    }
}

// The inference path of the thermodynamics kernel on a list of inputs: each
// wavefront evaluates 64 samples with guess_wave (lanes past n repeat the
// last sample and store nothing).
__global__ __launch_bounds__(256) void infer_kernel(const Weights *__restrict__ wg, const double *__restrict__ data,
                                                    long long n, float *out) {
    __shared__ Weights w;
    stage(wg, &w, threadIdx.x, blockDim.x);
    __syncthreads();
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long ic = i < n ? i : n - 1;
    const double *d = data + 4 * ic;
    float x0, x1, x2;
    normalise(&w, d[0], (float)d[1], (float)d[2], &x0, &x1, &x2);
    const float T = guess_wave(&w, x0, x1, x2);
    if (i < n) out[i] = T;
}

int check_stream_args(const void *p, const char *what) {
    if (!p) {
        pycles_set_error("mlp: NULL %s", what);
        return PYCLES_ERR_ARG;
    }
    return PYCLES_OK;
}

}  // namespace

extern "C" int pycles_sa_training_pairs(const pycles_lookup *lt, long long n, unsigned long long seed, int max_iter,
                                        int saturated_only, double *out, void *stream) {
    if (!lt || !lt->y || lt->n < 2 || !out || n < 1) {
        pycles_set_error("sa_training_pairs: bad arguments");
        return PYCLES_ERR_ARG;
    }
    thermo::Table t{lt->y, lt->x_min, lt->dx, lt->dxi, (int)lt->n};
    hipLaunchKernelGGL(pairs_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, t, n,
                       seed, max_iter, saturated_only, out);
    PYCLES_CHECK_LAUNCH();
    return PYCLES_OK;
}

extern "C" size_t pycles_mlp_param_count(void) { return NPARAM; }
extern "C" size_t pycles_mlp_norm_bytes(void) { return sizeof(Norm); }
extern "C" size_t pycles_mlp_weights_bytes(void) { return sizeof(Weights); }

extern "C" size_t pycles_mlp_workspace_bytes(int batch) {
    const size_t g = (size_t)(batch > 0 ? (batch + TB - 1) / TB : 1) * PSTRIDE * sizeof(float);
    const size_t nb = (size_t)NORM_BLOCKS * 8 * sizeof(double);
    return g > nb ? g : nb;
}

extern "C" int pycles_mlp_init(float *params, float *adam_m, float *adam_v, unsigned long long seed, void *stream) {
    if (check_stream_args(params, "params")) return PYCLES_ERR_ARG;
    hipLaunchKernelGGL(init_kernel, dim3((NPARAM + 255) / 256), dim3(256), 0, (hipStream_t)stream, params, adam_m,
                       adam_v, seed);
    PYCLES_CHECK_LAUNCH();
    return PYCLES_OK;
}

extern "C" int pycles_mlp_normalisation(const double *data, long long n, void *norm, void *workspace,
                                        void *stream) {
    if (!data || !norm || !workspace || n < 2) {
        pycles_set_error("mlp_normalisation: bad arguments");
        return PYCLES_ERR_ARG;
    }
    hipStream_t st = (hipStream_t)stream;
    double *part = (double *)workspace;
    hipLaunchKernelGGL(norm_partial_kernel, dim3(NORM_BLOCKS), dim3(256), 0, st, data, n, part);
    hipLaunchKernelGGL(norm_final_kernel, dim3(1), dim3(64), 0, st, part, NORM_BLOCKS, n, (Norm *)norm);
    PYCLES_CHECK_LAUNCH();
    return PYCLES_OK;
}

extern "C" int pycles_mlp_grad(const float *params, const void *norm, const double *data, long long n, float *grad,
                               double *loss, void *workspace, void *stream) {
    if (!params || !norm || !data || !grad || !workspace || n < TB || n % TB) {
        pycles_set_error("mlp_grad: bad arguments (n must be a positive multiple of %d)", TB);
        return PYCLES_ERR_ARG;
    }
    hipStream_t st = (hipStream_t)stream;
    const int nblk = (int)(n / TB);
    // identity order: idx = g & mask with mask >= n - 1
    unsigned long long mask = 1;
    while (mask < (unsigned long long)n) mask <<= 1;
    hipLaunchKernelGGL(grad_kernel, dim3(nblk), dim3(TB), 0, st, params, (const Norm *)norm, data, mask - 1, 1ull,
                       0ull, 0ll, 1.0f / (float)n, (float *)workspace);
    hipLaunchKernelGGL(reduce_adam_kernel, dim3((PSTRIDE + 255) / 256), dim3(256), 0, st, (float *)nullptr,
                       (float *)nullptr, (float *)nullptr, (const float *)workspace, nblk, 0.0f, 0.0f, 0.0f, grad,
                       loss);
    PYCLES_CHECK_LAUNCH();
    return PYCLES_OK;
}

extern "C" int pycles_mlp_train(float *params, float *adam_m, float *adam_v, const void *norm, const double *data,
                                long long n, int batch, int epochs, float lr_max, float lr_min,
                                unsigned long long seed, long long step0, long long total_steps,
                                double *epoch_loss, void *workspace, void *stream) {
    if (!params || !adam_m || !adam_v || !norm || !data || !workspace || epochs < 1) {
        pycles_set_error("mlp_train: bad arguments");
        return PYCLES_ERR_ARG;
    }
    if (n < TB || (n & (n - 1)) || batch < TB || batch % TB || batch > n) {
        pycles_set_error("mlp_train: n must be a power of two and batch a multiple of %d in [%d, n]", TB, TB);
        return PYCLES_ERR_ARG;
    }
    hipStream_t st = (hipStream_t)stream;
    const long long steps = n / batch;
    const int nblk = batch / TB;
    if (total_steps <= 0) total_steps = steps * epochs;
    long long t = step0;
    for (int e = 0; e < epochs; ++e) {
        const unsigned long long pa = mix64_host(seed * 0x100000001B3ull + 2 * e) | 1ull;
        const unsigned long long pc = mix64_host(seed * 0x100000001B3ull + 2 * e + 1);
        for (long long s = 0; s < steps; ++s) {
            hipLaunchKernelGGL(grad_kernel, dim3(nblk), dim3(TB), 0, st, params, (const Norm *)norm, data,
                               (unsigned long long)(n - 1), pa, pc, s * batch, 1.0f / (float)batch,
                               (float *)workspace);
            ++t;
            const double frac = (double)(t - 1) / (double)total_steps;
            const double lr = lr_min + 0.5 * (lr_max - lr_min) * (1.0 + cos(PC_PI * fmin(frac, 1.0)));
            const float c1 = (float)(1.0 / (1.0 - pow(0.9, (double)t)));
            const float c2 = (float)(1.0 / (1.0 - pow(0.999, (double)t)));
            hipLaunchKernelGGL(reduce_adam_kernel, dim3((PSTRIDE + 255) / 256), dim3(256), 0, st, params, adam_m,
                               adam_v, (const float *)workspace, nblk, (float)lr, c1, c2, (float *)nullptr,
                               epoch_loss ? epoch_loss + e : nullptr);
        }
        PYCLES_CHECK_LAUNCH();
    }
    return PYCLES_OK;
}

extern "C" int pycles_mlp_forward(const float *params, const void *norm, const double *data, long long n, float *out,
                                  void *stream) {
    if (!params || !norm || !data || !out || n < 1) {
        pycles_set_error("mlp_forward: bad arguments");
        return PYCLES_ERR_ARG;
    }
    hipLaunchKernelGGL(forward_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, params,
                       (const Norm *)norm, data, n, out);
    PYCLES_CHECK_LAUNCH();
    return PYCLES_OK;
}

extern "C" int pycles_mlp_pack(const float *params, const void *norm, void *weights, void *stream) {
    if (!params || !norm || !weights || ((uintptr_t)weights & 15)) {
        pycles_set_error("mlp_pack: bad arguments (weights must be 16-byte aligned)");
        return PYCLES_ERR_ARG;
    }
    hipLaunchKernelGGL(pack_kernel, dim3(1), dim3(256), 0, (hipStream_t)stream, params, (const Norm *)norm,
                       (Weights *)weights);
    PYCLES_CHECK_LAUNCH();
    return PYCLES_OK;
}

extern "C" int pycles_mlp_infer(const void *weights, const double *data, long long n, float *out, void *stream) {
    if (!weights || !data || !out || n < 1 || ((uintptr_t)weights & 15)) {
        pycles_set_error("mlp_infer: bad arguments");
        return PYCLES_ERR_ARG;
    }
    hipLaunchKernelGGL(infer_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       (const Weights *)weights, data, n, out);
    PYCLES_CHECK_LAUNCH();
    return PYCLES_OK;
}
