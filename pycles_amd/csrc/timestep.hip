// SSP-RK3 stage update with the z boundary fill, the z boundary fill alone,
// level means, CFL maximum and layout conversion.
//
//   TimeStepping.update_third     TimeStepping.pyx:131-151
//   set_bcs (sym / asym)          prognostic_variables.h:205-270
//   ParallelMPI.HorizontalMean    ParallelMPI.pyx:196-241
//   compute_cfl_max               TimeStepping.pyx:222-263
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include "common.h"

static thread_local char g_err[1024] = "";

void pycles_set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

extern "C" const char *pycles_last_error(void) { return g_err; }
extern "C" const char *pycles_version(void) { return "pycles_amd 0.1 (gfx950)"; }
extern "C" int pycles_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int pycles_check_dims(const pycles_dims *d) {
    if (!d) {
        pycles_set_error("dims is NULL");
        return PYCLES_ERR_ARG;
    }
    for (int i = 0; i < 3; ++i)
        if (d->n[i] < 1 || d->n[i] > (1 << 20)) {
            pycles_set_error("dims.n[%d] = %lld out of range", i, (long long)d->n[i]);
            return PYCLES_ERR_ARG;
        }
    if (d->gw < 2 || d->gw > 8 || d->gw > d->n[2]) {
        pycles_set_error("dims.gw = %lld out of range", (long long)d->gw);
        return PYCLES_ERR_ARG;
    }
    const long long npts = d->n[0] * d->n[1] * (d->n[2] + 2 * d->gw);
    if (npts >= (1LL << 40)) {
        pycles_set_error("domain too large");
        return PYCLES_ERR_ARG;
    }
    return PYCLES_OK;
}

// Mirror one column's ghost levels after its interior is final.
template <typename T>
__device__ __forceinline__ void fill_column_bcs(T *f, long long col, long long plane, int gw, int nzg,
                                                double bc) {
    if (bc == 1.0) {
        for (int m = 0; m < gw; ++m) {
            f[(long long)(gw - 1 - m) * plane + col] = f[(long long)(gw + m) * plane + col];
            f[(long long)(nzg - gw + m) * plane + col] = f[(long long)(nzg - gw - 1 - m) * plane + col];
        }
    } else if (bc == 2.0) {
        for (int m = 0; m < gw; ++m) {
            f[(long long)(gw - 1 - m) * plane + col] = f[(long long)gw * plane + col];
            f[(long long)(nzg - gw + m) * plane + col] = f[(long long)(nzg - gw - 1) * plane + col];
        }
    } else {
        const T fac = (T)bc;
        f[(long long)(gw - 1) * plane + col] = T(0);
        f[(long long)(nzg - gw) * plane + col] = T(0);
        for (int m = 1; m < gw; ++m) {
            f[(long long)(gw - 1 - m) * plane + col] = fac * f[(long long)(gw - 1 + m) * plane + col];
            f[(long long)(nzg - gw + m) * plane + col] = fac * f[(long long)(nzg - gw - m) * plane + col];
        }
    }
}

struct PtrPack {
    void *v[8];
    void *t[8];
    void *c[8];
    double bc[8];
};

// One thread per column: interior update in k, then the ghost fill of the
// same column, so no thread reads a level another thread writes.
// Grid (columns/256, nv): blockIdx.y selects the variable.

template <typename T, int RK>
__global__ __launch_bounds__(256) void rk3_kernel(Geo g, PtrPack pk, double dtd, int zero_tend) {
    const long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= g.plane) return;
    const int v = blockIdx.y;
    const T dt = (T)dtd;
    {
        T *val = (T *)pk.v[v];
        T *ten = (T *)pk.t[v];
        T *cpy = (T *)pk.c[v];
        const long long P = g.plane;
        int k = g.gw;
        const int k1 = g.nz + g.gw;
        // four levels per iteration: independent loads in flight
        for (; k + 3 < k1; k += 4) {
            const long long o = (long long)k * P + col;
            T phi[4], tn[4], cp[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                phi[u] = val[o + u * P];
                tn[u] = ten[o + u * P];
                cp[u] = RK == 0 ? T(0) : cpy[o + u * P];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (RK == 0) cpy[o + u * P] = phi[u];
                val[o + u * P] = rk_update<T, RK>(phi[u], tn[u], cp[u], dt);
                if (zero_tend) ten[o + u * P] = T(0);
            }
        }
        for (; k < k1; ++k) {
            const long long o = (long long)k * P + col;
            const T phi = val[o];
            const T tn = ten[o];
            const T cp = RK == 0 ? T(0) : cpy[o];
            if (RK == 0) cpy[o] = phi;
            val[o] = rk_update<T, RK>(phi, tn, cp, dt);
            if (zero_tend) ten[o] = T(0);
        }
        if (zero_tend) {
            // the reference zeroes every tendency, ghost levels included
            // (the SA buoyancy writes w tendencies on low ghost levels)
            for (int m = 0; m < g.gw; ++m) {
                ten[(long long)m * g.plane + col] = T(0);
                ten[(long long)(g.nzg - 1 - m) * g.plane + col] = T(0);
            }
        }
        fill_column_bcs(val, col, g.plane, g.gw, g.nzg, pk.bc[v]);
    }
}

template <typename T>
__global__ __launch_bounds__(256) void bcs_kernel(Geo g, int nv, PtrPack pk) {
    const long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= g.plane) return;
    for (int v = 0; v < nv; ++v) fill_column_bcs((T *)pk.v[v], col, g.plane, g.gw, g.nzg, pk.bc[v]);
}

static int pack(int nv, void *const *values, void *const *tend, void *const *copies, const double *bc,
                PtrPack *pk) {
    if (nv < 1 || nv > 8) {
        pycles_set_error("nv = %d out of range [1, 8]", nv);
        return PYCLES_ERR_ARG;
    }
    memset(pk, 0, sizeof(*pk));
    for (int v = 0; v < nv; ++v) {
        pk->v[v] = values[v];
        pk->t[v] = tend ? tend[v] : nullptr;
        pk->c[v] = copies ? copies[v] : nullptr;
        pk->bc[v] = bc[v];
        if (!pk->v[v] || (tend && !pk->t[v]) || (copies && !pk->c[v])) {
            pycles_set_error("NULL field pointer for variable %d", v);
            return PYCLES_ERR_ARG;
        }
    }
    return PYCLES_OK;
}

extern "C" int pycles_rk3_stage(const pycles_dims *dims, int nv, void *const *values, void *const *tend,
                                void *const *copies, const double *bc_type, int rk_step, double dt,
                                int precision, int zero_tendencies, void *stream) {
    int rc = pycles_check_dims(dims);
    if (rc) return rc;
    if (rk_step < 0 || rk_step > 2 || !tend || !copies || !values || !bc_type) {
        pycles_set_error("rk3_stage: bad arguments");
        return PYCLES_ERR_ARG;
    }
    PtrPack pk;
    if ((rc = pack(nv, values, tend, copies, bc_type, &pk))) return rc;
    Geo g = make_geo(dims);
    const int B = 256;
    const dim3 grid((unsigned)((g.plane + B - 1) / B), nv);
    hipStream_t st = (hipStream_t)stream;
    if (precision != PYCLES_F64 && precision != PYCLES_F32) {
        pycles_set_error("bad precision %d", precision);
        return PYCLES_ERR_ARG;
    }
    const int zt = zero_tendencies ? 1 : 0;
#define RK_LAUNCH(T, R) hipLaunchKernelGGL((rk3_kernel<T, R>), grid, dim3(B), 0, st, g, pk, dt, zt)
    if (precision == PYCLES_F64) {
        if (rk_step == 0) RK_LAUNCH(double, 0);
        else if (rk_step == 1) RK_LAUNCH(double, 1);
        else RK_LAUNCH(double, 2);
    } else {
        if (rk_step == 0) RK_LAUNCH(float, 0);
        else if (rk_step == 1) RK_LAUNCH(float, 1);
        else RK_LAUNCH(float, 2);
    }
#undef RK_LAUNCH
    PYCLES_CHECK_LAUNCH();
    return PYCLES_OK;
}

extern "C" int pycles_apply_z_bcs(const pycles_dims *dims, int nv, void *const *values,
                                  const double *bc_type, int precision, void *stream) {
    int rc = pycles_check_dims(dims);
    if (rc) return rc;
    if (!values || !bc_type) {
        pycles_set_error("apply_z_bcs: NULL argument");
        return PYCLES_ERR_ARG;
    }
    PtrPack pk;
    if ((rc = pack(nv, values, nullptr, nullptr, bc_type, &pk))) return rc;
    Geo g = make_geo(dims);
    const int B = 256;
    const unsigned grid = (unsigned)((g.plane + B - 1) / B);
    hipStream_t st = (hipStream_t)stream;
    if (precision == PYCLES_F64)
        hipLaunchKernelGGL(bcs_kernel<double>, dim3(grid), dim3(B), 0, st, g, nv, pk);
    else if (precision == PYCLES_F32)
        hipLaunchKernelGGL(bcs_kernel<float>, dim3(grid), dim3(B), 0, st, g, nv, pk);
    else {
        pycles_set_error("bad precision %d", precision);
        return PYCLES_ERR_ARG;
    }
    PYCLES_CHECK_LAUNCH();
    return PYCLES_OK;
}

// ---------------------------------------------------------------------------
// Level means: partial sums per (level, split) in FP64, then a fixed-order
// sum of the splits, so the result does not depend on scheduling.
#define MEAN_SPLITS 16

template <typename T>
__global__ __launch_bounds__(256) void level_partial_kernel(Geo g, const T *f, double *partial,
                                                            const double *base) {
    const int k = blockIdx.x;
    const int sp = blockIdx.y;
    const long long n = g.plane;
    const long long per = (n + MEAN_SPLITS - 1) / MEAN_SPLITS;
    const long long lo = sp * per;
    const long long hi = lo + per < n ? lo + per : n;
    const T *row = f + (long long)k * n;
    // four independent sums in flight per thread (fixed order: deterministic)
    const int B = blockDim.x;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    long long c = lo + threadIdx.x;
    for (; c + 3 * B < hi; c += 4 * B) {
        a0 += (double)row[c];
        a1 += (double)row[c + B];
        a2 += (double)row[c + 2 * B];
        a3 += (double)row[c + 3 * B];
    }
    for (; c < hi; c += B) a0 += (double)row[c];
    double acc = (a0 + a1) + (a2 + a3);
    // wavefront sum by lane shuffles, then the block's four wavefronts
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    __shared__ double red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) partial[(long long)k * MEAN_SPLITS + sp] = (red[0] + red[1]) + (red[2] + red[3]);
    (void)base;
}

__global__ void level_final_kernel(Geo g, const double *partial, double *mean, const double *base) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= g.nzg) return;
    double s = 0.0;
    for (int sp = 0; sp < MEAN_SPLITS; ++sp) s += partial[(long long)k * MEAN_SPLITS + sp];
    // multiply by 1/(n0*n1) as the reference does
    double m = s * (1.0 / (double)((long long)g.ny * g.nx));
    if (base) m += base[k];
    mean[k] = m;
}

int pycles_level_mean_ws(const Geo &g, const void *field, double *mean, double *partial,
                         const double *base, int precision, hipStream_t st) {
    if (precision == PYCLES_F64)
        hipLaunchKernelGGL(level_partial_kernel<double>, dim3(g.nzg, MEAN_SPLITS), dim3(256), 0, st, g,
                           (const double *)field, partial, base);
    else
        hipLaunchKernelGGL(level_partial_kernel<float>, dim3(g.nzg, MEAN_SPLITS), dim3(256), 0, st, g,
                           (const float *)field, partial, base);
    PYCLES_CHECK_LAUNCH();
    hipLaunchKernelGGL(level_final_kernel, dim3((g.nzg + 255) / 256), dim3(256), 0, st, g, partial, mean, base);
    PYCLES_CHECK_LAUNCH();
    return PYCLES_OK;
}

// Standalone entry: one block per level (no workspace needed).
template <typename T>
__global__ __launch_bounds__(256) void level_mean_kernel(Geo g, const T *f, double *mean) {
    const int k = blockIdx.x;
    const T *row = f + (long long)k * g.plane;
    double acc = 0.0;
    for (long long c = threadIdx.x; c < g.plane; c += blockDim.x) acc += (double)row[c];
    __shared__ double red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) mean[k] = red[0] * (1.0 / (double)((long long)g.ny * g.nx));
}

extern "C" int pycles_level_mean(const pycles_dims *dims, const void *field, double *mean, int precision,
                                 void *stream) {
    int rc = pycles_check_dims(dims);
    if (rc) return rc;
    if (!field || !mean || (precision != PYCLES_F64 && precision != PYCLES_F32)) {
        pycles_set_error("level_mean: bad arguments");
        return PYCLES_ERR_ARG;
    }
    Geo g = make_geo(dims);
    hipStream_t st = (hipStream_t)stream;
    if (precision == PYCLES_F64)
        hipLaunchKernelGGL(level_mean_kernel<double>, dim3(g.nzg), dim3(256), 0, st, g, (const double *)field, mean);
    else
        hipLaunchKernelGGL(level_mean_kernel<float>, dim3(g.nzg), dim3(256), 0, st, g, (const float *)field, mean);
    PYCLES_CHECK_LAUNCH();
    return PYCLES_OK;
}

// ---------------------------------------------------------------------------
// CFL maximum.  Values are non-negative, so their IEEE bit patterns order
// like unsigned integers and an integer atomic max is exact.
__global__ void cfl_init_kernel(double *out) { *out = 0.0; }

__global__ void cfl_finish_kernel(double *out) { *out = *out + 1e-11; }

template <typename T>
__global__ __launch_bounds__(256) void cfl_kernel(Geo g, const T *u, const T *v, const T *w, double dt,
                                                  double *out) {
    const long long n = g.plane * g.nz;
    double m = 0.0;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
        const long long o = t + (long long)g.gw * g.plane;
        const double c = dt * (fabs((double)u[o]) * g.dxi[0] + fabs((double)v[o]) * g.dxi[1] +
                               fabs((double)w[o]) * g.dxi[2]);
        m = cfl_fold(m, c);
    }
    cfl_block_max(m, out);
}

extern "C" int pycles_cfl_max(const pycles_dims *dims, const void *u, const void *v, const void *w, double dt,
                              double *out, int precision, void *stream) {
    int rc = pycles_check_dims(dims);
    if (rc) return rc;
    if (!u || !v || !w || !out) {
        pycles_set_error("cfl_max: NULL argument");
        return PYCLES_ERR_ARG;
    }
    Geo g = make_geo(dims);
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(cfl_init_kernel, dim3(1), dim3(1), 0, st, out);
    const long long n = g.plane * g.nz;
    unsigned grid = (unsigned)((n + 255) / 256);
    if (grid > 4096) grid = 4096;
    if (precision == PYCLES_F64)
        hipLaunchKernelGGL(cfl_kernel<double>, dim3(grid), dim3(256), 0, st, g, (const double *)u,
                           (const double *)v, (const double *)w, dt, out);
    else if (precision == PYCLES_F32)
        hipLaunchKernelGGL(cfl_kernel<float>, dim3(grid), dim3(256), 0, st, g, (const float *)u,
                           (const float *)v, (const float *)w, dt, out);
    else {
        pycles_set_error("bad precision %d", precision);
        return PYCLES_ERR_ARG;
    }
    hipLaunchKernelGGL(cfl_finish_kernel, dim3(1), dim3(1), 0, st, out);
    PYCLES_CHECK_LAUNCH();
    return PYCLES_OK;
}

// ---------------------------------------------------------------------------
// Layout conversion.  Reference layout index: (i*nlg1 + j)*nlg2 + k with
// i, j, k ghosted; device (k*ny + j)*nx + i over x/y interior.
template <typename T>
__global__ void ingress_kernel(Geo g, const double *src, T *dst, const double *base) {
    const long long n = g.plane * g.nzg;
    const long long nlg1 = g.ny + 2 * g.gw, nlg2 = g.nzg;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(t % g.nx);
        const int j = (int)((t / g.nx) % g.ny);
        const int k = (int)(t / g.plane);
        const long long r = ((long long)(i + g.gw) * nlg1 + (j + g.gw)) * nlg2 + k;
        double v = src[r];
        if (base) v -= base[k];
        dst[t] = (T)v;
    }
}

template <typename T>
__global__ void egress_kernel(Geo g, const T *src, double *dst, const double *base) {
    const long long nlg0 = g.nx + 2 * g.gw, nlg1 = g.ny + 2 * g.gw, nlg2 = g.nzg;
    const long long n = nlg0 * nlg1 * nlg2;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
        const int k = (int)(t % nlg2);
        const int j = (int)((t / nlg2) % nlg1);
        const int i = (int)(t / (nlg1 * nlg2));
        int ii = (i - g.gw) % g.nx;
        if (ii < 0) ii += g.nx;
        int jj = (j - g.gw) % g.ny;
        if (jj < 0) jj += g.ny;
        double v = (double)src[gidx(g, ii, jj, k)];
        if (base) v += base[k];
        dst[t] = v;
    }
}

extern "C" int pycles_layout_ingress(const pycles_dims *dims, const double *ref_layout, void *dev,
                                     const double *base, int precision, void *stream) {
    int rc = pycles_check_dims(dims);
    if (rc) return rc;
    if (!ref_layout || !dev) {
        pycles_set_error("layout_ingress: NULL array");
        return PYCLES_ERR_ARG;
    }
    if (precision != PYCLES_F64 && precision != PYCLES_F32) {
        pycles_set_error("layout_ingress: bad precision %d", precision);
        return PYCLES_ERR_ARG;
    }
    Geo g = make_geo(dims);
    hipStream_t st = (hipStream_t)stream;
    if (precision == PYCLES_F64)
        hipLaunchKernelGGL(ingress_kernel<double>, dim3(2048), dim3(256), 0, st, g, ref_layout, (double *)dev, base);
    else
        hipLaunchKernelGGL(ingress_kernel<float>, dim3(2048), dim3(256), 0, st, g, ref_layout, (float *)dev, base);
    PYCLES_CHECK_LAUNCH();
    return PYCLES_OK;
}

extern "C" int pycles_layout_egress(const pycles_dims *dims, const void *dev, double *ref_layout,
                                    const double *base, int precision, void *stream) {
    int rc = pycles_check_dims(dims);
    if (rc) return rc;
    if (!ref_layout || !dev) {
        pycles_set_error("layout_egress: NULL array");
        return PYCLES_ERR_ARG;
    }
    if (precision != PYCLES_F64 && precision != PYCLES_F32) {
        pycles_set_error("layout_egress: bad precision %d", precision);
        return PYCLES_ERR_ARG;
    }
    Geo g = make_geo(dims);
    hipStream_t st = (hipStream_t)stream;
    if (precision == PYCLES_F64)
        hipLaunchKernelGGL(egress_kernel<double>, dim3(2048), dim3(256), 0, st, g, (const double *)dev, ref_layout, base);
    else
        hipLaunchKernelGGL(egress_kernel<float>, dim3(2048), dim3(256), 0, st, g, (const float *)dev, ref_layout, base);
    PYCLES_CHECK_LAUNCH();
    return PYCLES_OK;
}

// ---------------------------------------------------------------------------
// The CFL value on the host, after the stream's work: the time-step
// adjustment's one synchronisation per step (TimeStepping.pyx:265-266).
extern "C" int pycles_cfl_read(const double *dev_cfl, double *host_cfl, void *stream) {
    if (!dev_cfl || !host_cfl) {
        pycles_set_error("cfl_read: NULL argument");
        return PYCLES_ERR_ARG;
    }
    hipStream_t st = (hipStream_t)stream;
    hipError_t e = hipMemcpyAsync(host_cfl, dev_cfl, sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        pycles_set_error("cfl_read: %s", hipGetErrorString(e));
        return PYCLES_ERR_HIP;
    }
    const double h = *host_cfl;
    if (!(h - h == 0.0)) {  // NaN or infinite
        pycles_set_error("CFL number is not finite (%g): a velocity is NaN or infinite", *host_cfl);
        return PYCLES_ERR_NONFINITE;
    }
    return PYCLES_OK;
}

// Chunk layout of the z-chunked kernels (see the header).
extern "C" int pycles_launch_chunks(int kernel, const pycles_dims *dims, int *nchunks, int *chunk_len) {
    int rc = pycles_check_dims(dims);
    if (rc) return rc;
    if (!nchunks || !chunk_len) {
        pycles_set_error("launch_chunks: NULL argument");
        return PYCLES_ERR_ARG;
    }
    const Geo g = make_geo(dims);
    ZSplit z;
    switch (kernel) {
    case 0: z = correct_zsplit(g.plane, g.nz, g.gw); break;
    case 1: z = advy_zsplit(g.nx, g.ny, g.nz); break;
    case 2: z = advz_zsplit(g.nx, g.ny, g.nz); break;
    case 3: z = thermo_zsplit(g.nx, g.ny, g.nzg); break;
    case 4: z = advyz_zsplit(g.nx, g.ny, g.nz); break;
    default:
        pycles_set_error("launch_chunks: unknown kernel %d", kernel);
        return PYCLES_ERR_ARG;
    }
    *nchunks = z.nch;
    *chunk_len = z.len;
    return PYCLES_OK;
}
