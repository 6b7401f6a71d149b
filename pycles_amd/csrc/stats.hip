// On-device horizontal statistics: per-level moments, maxima and minima of a
// field or of a product of two fields, optionally conditional on a mask, and
// the liquid-water statistics of the saturation adjustment.
//
//   HorizontalMean / MeanofSquares / MeanofCubes  ParallelMPI.pyx:196-326
//   HorizontalMaximum / HorizontalMinimum         ParallelMPI.pyx:327-389
//   HorizontalMeanConditional (+ofSquares)        ParallelMPI.pyx:430-540
//   ThermodynamicsSA.liquid_stats                 ThermodynamicsSA.pyx:395-492
//
// Every level is split over STAT_SPLITS blocks; each block reduces its share
// of the level in a fixed order into FP64 partials, and a second kernel adds
// the partials in split order, so the result does not depend on scheduling.
// The partials live in a caller-provided workspace.
#include <mutex>

#include "common.h"
#include "thermo_funcs.h"

extern int pycles_check_dims(const pycles_dims *d);

#define STAT_SPLITS 16
#define STAT_NQ 6  // sum v m, sum v^2 m, sum v^3 m, max, min, sum m
#define STAT_BLOCK 256

namespace {

template <typename T>
__global__ __launch_bounds__(STAT_BLOCK) void stats_partial_kernel(Geo g, const T *a, const double *base_a,
                                                                   const T *b, const double *base_b,
                                                                   const T *mask, double *partial) {
    const int k = blockIdx.x;
    const int sp = blockIdx.y;
    const long long n = g.plane;
    const long long per = (n + STAT_SPLITS - 1) / STAT_SPLITS;
    const long long lo = sp * per;
    const long long hi = lo + per < n ? lo + per : n;
    const long long off = (long long)k * n;
    const double ba = base_a ? base_a[k] : 0.0;
    const double bb = b && base_b ? base_b[k] : 0.0;
    double s1 = 0.0, s2 = 0.0, s3 = 0.0, sm = 0.0;
    double mx = -9e12, mn = 9e12;  // the reference's initial extremes
    for (long long c = lo + threadIdx.x; c < hi; c += STAT_BLOCK) {
        double v = (double)a[off + c] + ba;
        if (b) v *= (double)b[off + c] + bb;
        const double m = mask ? (double)mask[off + c] : 1.0;
        s1 += v * m;
        s2 += v * v * m;
        s3 += v * v * v * m;
        sm += m;
        if (m != 0.0) {
            mx = fmax(mx, v);
            mn = fmin(mn, v);
        }
    }
    // fixed-order tree over the block
    __shared__ double red[STAT_NQ][STAT_BLOCK];
    red[0][threadIdx.x] = s1;
    red[1][threadIdx.x] = s2;
    red[2][threadIdx.x] = s3;
    red[3][threadIdx.x] = mx;
    red[4][threadIdx.x] = mn;
    red[5][threadIdx.x] = sm;
    __syncthreads();
    for (int s = STAT_BLOCK / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            const int t = threadIdx.x;
            red[0][t] += red[0][t + s];
            red[1][t] += red[1][t + s];
            red[2][t] += red[2][t + s];
            red[3][t] = fmax(red[3][t], red[3][t + s]);
            red[4][t] = fmin(red[4][t], red[4][t + s]);
            red[5][t] += red[5][t + s];
        }
        __syncthreads();
    }
    if (threadIdx.x < STAT_NQ) partial[((long long)k * STAT_SPLITS + sp) * STAT_NQ + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void stats_final_kernel(Geo g, const double *partial, int masked, double *out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= g.nzg) return;
    double q[STAT_NQ] = {0.0, 0.0, 0.0, -9e12, 9e12, 0.0};
    for (int sp = 0; sp < STAT_SPLITS; ++sp) {
        const double *p = partial + ((long long)k * STAT_SPLITS + sp) * STAT_NQ;
        q[0] += p[0];
        q[1] += p[1];
        q[2] += p[2];
        q[3] = fmax(q[3], p[3]);
        q[4] = fmin(q[4], p[4]);
        q[5] += p[5];
    }
    // unconditional means multiply by 1/(n0*n1) as the reference does;
    // conditional ones divide by max(sum m, 1)
    const double inv = 1.0 / (double)((long long)g.ny * g.nx);
    for (int r = 0; r < 3; ++r) out[r * g.nzg + k] = masked ? q[r] / fmax(q[5], 1.0) : q[r] * inv;
    out[3 * g.nzg + k] = q[3];
    out[4 * g.nzg + k] = q[4];
    out[5 * g.nzg + k] = q[5];
}

// Liquid-water statistics: one thread per column marching the interior
// levels.  Per-level cloudy-cell counts and the cloudy-column count are
// integers, gathered per block in LDS and added to the global counters once
// per block (integer adds are exact, so the order does not matter); the
// liquid water path is summed per block in a fixed order and the block sums
// in block order.  Dynamic LDS: nz + 1 counters.
template <typename T>
__global__ __launch_bounds__(STAT_BLOCK) void cloud_kernel(Geo g, const T *ql, const double *rho0_half,
                                                           unsigned long long *counts, double *lwp_part) {
    extern __shared__ unsigned int lcount[];
    __shared__ double red[STAT_BLOCK];
    for (int k = threadIdx.x; k <= g.nz; k += STAT_BLOCK) lcount[k] = 0u;
    __syncthreads();
    const long long c = (long long)blockIdx.x * STAT_BLOCK + threadIdx.x;
    double lwp = 0.0;
    int cloudy = 0;
    if (c < g.plane) {
        for (int k = 0; k < g.nz; ++k) {
            const double q = (double)ql[(long long)(k + g.gw) * g.plane + c];
            if (q > 0.0) {
                cloudy = 1;
                atomicAdd(&lcount[k], 1u);
            }
            // ThermodynamicsSA.pyx:478-480 weights interior level k with
            // rho0_half[k], the ghosted array's k-th entry
            lwp += rho0_half[k] * q * g.dx[2];
        }
    }
    if (cloudy) atomicAdd(&lcount[g.nz], 1u);
    red[threadIdx.x] = lwp;
    __syncthreads();
    for (int s = STAT_BLOCK / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) lwp_part[blockIdx.x] = red[0];
    for (int k = threadIdx.x; k <= g.nz; k += STAT_BLOCK)
        if (lcount[k]) atomicAdd(&counts[k], (unsigned long long)lcount[k]);
}

__global__ void zero_u64_kernel(unsigned long long *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// out: cloud fraction profile [nz], then cloud cover, cloud base, cloud top,
// liquid water path (ThermodynamicsSA.pyx:433-489).  Base and top are the
// centre heights z_half = (k + 1/2) dz of the lowest and highest level with a
// cloudy cell; 99999.9 and -99999.9 when there is none, as the reference's
// initial values.
__global__ void cloud_final_kernel(Geo g, const unsigned long long *counts, const double *lwp_part, int nblk,
                                   double *out) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const double inv = 1.0 / (double)((long long)g.ny * g.nx);
    double cb = 99999.9, ct = -99999.9;
    for (int k = 0; k < g.nz; ++k) {
        out[k] = (double)counts[k] * inv;
        if (counts[k]) {
            const double z = ((double)k + 0.5) * g.dx[2];
            cb = fmin(cb, z);
            ct = fmax(ct, z);
        }
    }
    double lwp = 0.0;
    for (int b = 0; b < nblk; ++b) lwp += lwp_part[b];
    // the column counts and the path are divided by n0*n1 as the reference
    // does (:461, :484); the fraction profile adds 1/(n0*n1) per cloudy cell
    // there, which the count times that reciprocal gives to a few ulps
    const double div = (double)((long long)g.ny * g.nx);
    out[g.nz] = (double)counts[g.nz] / div;
    out[g.nz + 1] = cb;
    out[g.nz + 2] = ct;
    out[g.nz + 3] = lwp / div;
}

// Fields of ThermodynamicsSA.stats_io (ThermodynamicsSA.pyx:256-389) at
// every stored point: the entropy potential temperature thetas_c(s, qt), the
// dry potential temperature theta_c(p0_half, T) and the relative humidity
// pv_c(p0_half, qt, qv) / pv*(T) (thermodynamic_functions.h:9-12, 23-25,
// 39-41; lookup.h:48-54), in the field precision.
template <typename TS>
__global__ __launch_bounds__(256) void sa_stat_fields_kernel(Geo g, const double *p0_half, thermo::Table lt,
                                                             const TS *s, const TS *qt, const TS *T, const TS *qv,
                                                             const double *sbase, const double *qbase,
                                                             TS *thetas, TS *theta, TS *rh) {
    const long long n = g.plane * g.nzg;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        const int k = (int)(t / g.plane);
        const double sv = (double)s[t] + (sbase ? sbase[k] : 0.0);
        const double q = (double)qt[t] + (qbase ? qbase[k] : 0.0);
        const double Tv = (double)T[t];
        if (thetas)
            thetas[t] = (TS)(PC_T_TILDE * exp((sv - (1.0 - q) * PC_SD_TILDE - q * PC_SV_TILDE) / thermo::cpm_c(q)));
        if (theta) theta[t] = (TS)(Tv / thermo::exner_c(p0_half[k]));
        if (rh) {
            const double pv_star = thermo::lookup(lt, Tv);
            const double pv = thermo::pv_c(p0_half[k], q, (double)qv[t]);
            rh[t] = (TS)(pv / pv_star);
        }
    }
}

}  // namespace

extern "C" int pycles_sa_stat_fields(const pycles_dims *dims, const pycles_ref *ref, const pycles_lookup *lt,
                                     const void *s, const void *qt, const void *T, const void *qv,
                                     const pycles_scalar_base *base, int precision, void *thetas, void *theta,
                                     void *rh, void *stream) {
    int rc = pycles_check_dims(dims);
    if (rc) return rc;
    if (!ref || !ref->p0_half || !lt || !lt->y || !s || !qt || !T || !qv) {
        pycles_set_error("sa_stat_fields: NULL argument");
        return PYCLES_ERR_ARG;
    }
    if (precision != PYCLES_F64 && precision != PYCLES_F32) {
        pycles_set_error("sa_stat_fields: bad precision %d", precision);
        return PYCLES_ERR_ARG;
    }
    const Geo g = make_geo(dims);
    thermo::Table t{lt->y, lt->x_min, lt->dx, lt->dxi, (int)lt->n};
    const double *sb = (precision == PYCLES_F32 && base) ? base->base[0] : nullptr;
    const double *qb = (precision == PYCLES_F32 && base) ? base->base[1] : nullptr;
    hipStream_t st = (hipStream_t)stream;
    if (precision == PYCLES_F64)
        hipLaunchKernelGGL(sa_stat_fields_kernel<double>, dim3(4096), dim3(256), 0, st, g, ref->p0_half, t,
                           (const double *)s, (const double *)qt, (const double *)T, (const double *)qv, sb, qb,
                           (double *)thetas, (double *)theta, (double *)rh);
    else
        hipLaunchKernelGGL(sa_stat_fields_kernel<float>, dim3(4096), dim3(256), 0, st, g, ref->p0_half, t,
                           (const float *)s, (const float *)qt, (const float *)T, (const float *)qv, sb, qb,
                           (float *)thetas, (float *)theta, (float *)rh);
    PYCLES_CHECK_LAUNCH();
    return PYCLES_OK;
}

static size_t stats_ws_bytes(const Geo &g) { return sizeof(double) * (size_t)g.nzg * STAT_SPLITS * STAT_NQ; }

extern "C" size_t pycles_horizontal_stats_workspace_bytes(const pycles_dims *dims) {
    if (pycles_check_dims(dims)) return 0;
    const Geo g = make_geo(dims);
    return stats_ws_bytes(g);
}

static int stats_args(const pycles_dims *dims, const void *a, int precision, double *out) {
    int rc = pycles_check_dims(dims);
    if (rc) return rc;
    if (!a || !out) {
        pycles_set_error("horizontal_stats: NULL field or output");
        return PYCLES_ERR_ARG;
    }
    if (precision != PYCLES_F64 && precision != PYCLES_F32) {
        pycles_set_error("horizontal_stats: bad precision %d", precision);
        return PYCLES_ERR_ARG;
    }
    return PYCLES_OK;
}

static int stats_launch(const Geo &g, const void *a, const double *base_a, const void *b, const double *base_b,
                        const void *mask, int precision, double *out, double *partial, hipStream_t st) {
    const dim3 grid(g.nzg, STAT_SPLITS);
    if (precision == PYCLES_F64)
        hipLaunchKernelGGL(stats_partial_kernel<double>, grid, dim3(STAT_BLOCK), 0, st, g, (const double *)a, base_a,
                           (const double *)b, base_b, (const double *)mask, partial);
    else
        hipLaunchKernelGGL(stats_partial_kernel<float>, grid, dim3(STAT_BLOCK), 0, st, g, (const float *)a, base_a,
                           (const float *)b, base_b, (const float *)mask, partial);
    PYCLES_CHECK_LAUNCH();
    hipLaunchKernelGGL(stats_final_kernel, dim3((g.nzg + 255) / 256), dim3(256), 0, st, g, partial, mask ? 1 : 0, out);
    PYCLES_CHECK_LAUNCH();
    return PYCLES_OK;
}

extern "C" int pycles_horizontal_stats(const pycles_dims *dims, const void *a, const double *base_a, const void *b,
                                       const double *base_b, const void *mask, int precision, double *out,
                                       void *workspace, void *stream) {
    int rc = stats_args(dims, a, precision, out);
    if (rc) return rc;
    if (!workspace) {
        pycles_set_error("horizontal_stats: NULL workspace");
        return PYCLES_ERR_ARG;
    }
    const Geo g = make_geo(dims);
    return stats_launch(g, a, base_a, b, base_b, mask, precision, out, (double *)workspace, (hipStream_t)stream);
}

// No-allocation variant of pycles_horizontal_stats for callers that cannot
// allocate device memory.  Its partials live in a file-scope device array,
// one copy per GPU, whose address each device resolves once (with that
// device current) by hipGetSymbolAddress; host code passes only that
// address, never the array's name.  Every (level, split) slot is written by
// stats_partial_kernel before stats_final_kernel reads it, so nothing needs
// clearing between calls.  All calls on one GPU share the array: calls that
// overlap on one GPU (two streams, two host threads) race.
#define STATS_NOALLOC_MAX_LEVELS 2048
#define STATS_NOALLOC_MAX_DEVICES 64
static __device__ __align__(256) double scratch[(size_t)STATS_NOALLOC_MAX_LEVELS * STAT_SPLITS * STAT_NQ];
static double *g_scratch_dev[STATS_NOALLOC_MAX_DEVICES];
static std::mutex g_scratch_mu;

static int stats_scratch(double **out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess || dev < 0 || dev >= STATS_NOALLOC_MAX_DEVICES) {
        pycles_set_error("horizontal_stats_noalloc: no usable current device (%d, %s)", dev, hipGetErrorString(e));
        return PYCLES_ERR_HIP;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            pycles_set_error("horizontal_stats_noalloc: hipGetSymbolAddress: %s", hipGetErrorString(e));
            return PYCLES_ERR_HIP;
        }
        g_scratch_dev[dev] = (double *)ptr;
    }
    *out = g_scratch_dev[dev];
    return PYCLES_OK;
}

extern "C" int pycles_horizontal_stats_noalloc(const pycles_dims *dims, const void *a, const double *base_a,
                                               const void *b, const double *base_b, const void *mask,
                                               int precision, double *out, void *stream) {
    int rc = stats_args(dims, a, precision, out);
    if (rc) return rc;
    const Geo g = make_geo(dims);
    if (g.nzg > STATS_NOALLOC_MAX_LEVELS) {
        pycles_set_error("horizontal_stats_noalloc: %d levels (with ghosts) exceed the scratch limit of %d",
                         g.nzg, STATS_NOALLOC_MAX_LEVELS);
        return PYCLES_ERR_UNSUPPORTED;
    }
    double *partial = nullptr;
    rc = stats_scratch(&partial);
    if (rc) return rc;
    return stats_launch(g, a, base_a, b, base_b, mask, precision, out, partial, (hipStream_t)stream);
}

extern "C" size_t pycles_cloud_stats_workspace_bytes(const pycles_dims *dims) {
    if (pycles_check_dims(dims)) return 0;
    const Geo g = make_geo(dims);
    const size_t nblk = (size_t)((g.plane + STAT_BLOCK - 1) / STAT_BLOCK);
    return sizeof(unsigned long long) * (size_t)(g.nz + 1) + sizeof(double) * nblk;
}

extern "C" int pycles_cloud_stats(const pycles_dims *dims, const double *rho0_half, const void *ql, int precision,
                                  double *out, void *workspace, void *stream) {
    int rc = pycles_check_dims(dims);
    if (rc) return rc;
    if (!rho0_half || !ql || !out || !workspace) {
        pycles_set_error("cloud_stats: NULL argument");
        return PYCLES_ERR_ARG;
    }
    if (precision != PYCLES_F64 && precision != PYCLES_F32) {
        pycles_set_error("cloud_stats: bad precision %d", precision);
        return PYCLES_ERR_ARG;
    }
    const Geo g = make_geo(dims);
    hipStream_t st = (hipStream_t)stream;
    unsigned long long *counts = (unsigned long long *)workspace;
    double *part = (double *)(counts + (g.nz + 1));
    const int nblk = (int)((g.plane + STAT_BLOCK - 1) / STAT_BLOCK);
    hipLaunchKernelGGL(zero_u64_kernel, dim3((g.nz + 1 + 255) / 256), dim3(256), 0, st, counts, g.nz + 1);
    const size_t lds = sizeof(unsigned int) * (size_t)(g.nz + 1);
    if (lds > 48 * 1024) {
        pycles_set_error("cloud_stats: %d levels exceed the kernel's LDS counters", g.nz);
        return PYCLES_ERR_UNSUPPORTED;
    }
    if (precision == PYCLES_F64)
        hipLaunchKernelGGL(cloud_kernel<double>, dim3(nblk), dim3(STAT_BLOCK), lds, st, g, (const double *)ql,
                           rho0_half, counts, part);
    else
        hipLaunchKernelGGL(cloud_kernel<float>, dim3(nblk), dim3(STAT_BLOCK), lds, st, g, (const float *)ql,
                           rho0_half, counts, part);
    PYCLES_CHECK_LAUNCH();
    hipLaunchKernelGGL(cloud_final_kernel, dim3(1), dim3(1), 0, st, g, counts, part, nblk, out);
    PYCLES_CHECK_LAUNCH();
    return PYCLES_OK;
}
