// Anelastic pressure projection on one GPU (whole domain).
//
// The reference (PressureSolver.pyx:38-81, PressureFFTParallel.pyx:150-220)
// removes the level mean of w, forms div(rho0 u) with second-order
// differences, takes complex FFTs of the real divergence along x then y
// (scipy.fftpack.fft), solves one real tridiagonal system per (kx, ky) for the
// real and imaginary parts (SparseSolvers.pyx:24-43), zeroes the (0,0) mode,
// inverts the FFTs, keeps the real part, fills the pressure halo and
// subtracts grad p from u, v, w.  MPI pencil transposes move data between the
// FFT and solve directions.
//
// Here the domain is whole on one GPU, so there are no transposes: the
// divergence of a row is computed and Fourier-transformed in LDS in one
// kernel, and because the input is real only the half spectrum
// kx = 0..nx/2 is kept (the other half is its complex conjugate, so the
// solves there are redundant).  The y transform, the per-column Thomas
// sweep and the inverse transforms each take one pass over the half
// spectrum [nz][ny][nx/2+1] (kx fastest).  The Thomas recurrence runs in FP64
// in both precisions (its low-wavenumber systems are the ill-conditioned
// ones).  The correction marches each column in z so it can also fill the z
// ghost levels of u, v, w and reduce the CFL number.
#include <math.h>
#include <string.h>

#include <vector>

#include "common.h"

extern int pycles_check_dims(const pycles_dims *d);
extern int pycles_level_mean_ws(const Geo &g, const void *field, double *mean, double *partial,
                                const double *base, int precision, hipStream_t st);

struct pycles_pressure_plan {
    Geo g;
    int precision;
    int nxh;
    int gx;          // kx columns per y-FFT block
    void *spec;      // [nz][ny][nxh] complex, field precision
    double *xs;      // Thomas forward-sweep values, complex FP64
    double *cs;      // Thomas forward-sweep coefficients, FP64
    void *twx;       // exp(-2 pi i m / nx), m < nx/2, interleaved (field precision)
    void *twy;
    double *kx2, *ky2, *rho_h, *dd, *a, *c;
    double *mean, *partial, *cfl;
    int logx, logy, loggx;
};

static int ilog2(long long n) {
    int l = 0;
    while ((1LL << l) < n) ++l;
    return ((1LL << l) == n) ? l : -1;
}

template <typename T> struct C2 { T x, y; };

__device__ __forceinline__ unsigned brev(unsigned v, int bits) { return __brev(v) >> (32 - bits); }

// In-place radix-2 decimation-in-time FFT over LDS arrays re/im holding
// `batch` transforms of length N (element n of transform b at n*stride+b),
// already in bit-reversed order.  tw[m] = exp(-2 pi i m/N), m < N/2; sign
// +1 selects the inverse kernel (conjugate twiddles).  Butterflies are
// evaluated as in a textbook DIT sweep: x_b <- x_a - w x_b, x_a <- x_a + w x_b.
// batch and N are powers of two (logB = log2(batch)); tw points to LDS.
template <typename T>
__device__ void lds_fft(T *re, T *im, int N, int logN, int stride, int logB, const C2<T> *tw, int inverse) {
    const int batch = 1 << logB;
    const int nb = (N >> 1) << logB;
    for (int s = 1; s <= logN; ++s) {
        const int h = 1 << (s - 1);
        const int tstep = N >> s;  // twiddle index stride
        for (int t = threadIdx.x; t < nb; t += blockDim.x) {
            const int b = t & (batch - 1);
            const int bf = t >> logB;
            const int grp = bf >> (s - 1), pos = bf & (h - 1);
            const int ia = (grp * 2 * h + pos) * stride + b;
            const int ib = ia + h * stride;
            const C2<T> w = tw[pos * tstep];
            const T wr = w.x, wi = inverse ? -w.y : w.y;
            const T xr = re[ib] * wr - im[ib] * wi;
            const T xi = re[ib] * wi + im[ib] * wr;
            const T ar = re[ia], ai = im[ia];
            re[ib] = ar - xr;
            im[ib] = ai - xi;
            re[ia] = ar + xr;
            im[ia] = ai + xi;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Level mean of w is computed before this kernel.  One block per R = 2^logR
// consecutive rows (j0 .. j0+R-1, k'): divergence (PressureSolver.pyx:138-175,
// zeroed then x, y, z contributions added in that order), forward FFT along x
// of the R rows together (interleaved in LDS), half spectra written out.
template <typename TS>
__global__ __launch_bounds__(256) void div_fftx_kernel(Geo g, const TS *u, const TS *v, const TS *w,
                                                       const double *alpha0, const double *alpha0_half,
                                                       const double *mean, TS *div_out, C2<TS> *spec,
                                                       const C2<TS> *tw, int logx, int nxh, int logR) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int R = 1 << logR;
    C2<TS> *twl = reinterpret_cast<C2<TS> *>(smem_raw);
    TS *re = reinterpret_cast<TS *>(twl + g.nx / 2);
    TS *im = re + g.nx * R;
    for (int m = threadIdx.x; m < g.nx / 2; m += blockDim.x) twl[m] = tw[m];
    const int kk = blockIdx.y;
    const int k = kk + g.gw;
    const double ah = alpha0_half[k];
    const double a0 = alpha0[k], a0m = alpha0[k - 1];
    const double mk = mean[k], mkm = mean[k - 1];
    const double dxi0 = 1.0 / g.dx[0], dxi1 = 1.0 / g.dx[1], dxi2 = 1.0 / g.dx[2];
    for (int t = threadIdx.x; t < (g.nx << logR); t += blockDim.x) {
        const int i = t & (g.nx - 1);
        const int rb = t >> logx;
        const int j = (blockIdx.x << logR) + rb;
        const int jm = j == 0 ? g.ny - 1 : j - 1;
        const int im1 = i == 0 ? g.nx - 1 : i - 1;
        const long long o = gidx(g, i, j, k);
        TS d;
        if (sizeof(TS) == 8) {
            double dv = 0.0;
            dv += ((double)u[o] / ah - (double)u[gidx(g, im1, j, k)] / ah) * dxi0;
            dv += ((double)v[o] / ah - (double)v[gidx(g, i, jm, k)] / ah) * dxi1;
            const double wk = (double)w[o] - mk;
            const double wkm = (double)w[o - g.plane] - mkm;
            dv += ((wk) / a0 - (wkm) / a0m) * dxi2;
            d = (TS)dv;
        } else {
            // FP32 path: rho0*u in FP32 except the level means and profiles
            const float rh = (float)(1.0 / ah);
            float dv = ((float)u[o] - (float)u[gidx(g, im1, j, k)]) * (float)(dxi0)*rh;
            dv += ((float)v[o] - (float)v[gidx(g, i, jm, k)]) * (float)(dxi1)*rh;
            const float wk = (float)((double)w[o] - mk);
            const float wkm = (float)((double)w[o - g.plane] - mkm);
            dv += (wk * (float)(1.0 / a0) - wkm * (float)(1.0 / a0m)) * (float)dxi2;
            d = (TS)dv;
        }
        if (div_out) div_out[o] = d;
        const unsigned r = (brev((unsigned)i, logx) << logR) + rb;
        re[r] = d;
        im[r] = TS(0);
    }
    __syncthreads();
    lds_fft<TS>(re, im, g.nx, logx, R, logR, twl, 0);
    for (int rb = 0; rb < R; ++rb) {
        const int j = (blockIdx.x << logR) + rb;
        C2<TS> *row = spec + ((long long)kk * g.ny + j) * nxh;
        for (int m = threadIdx.x; m < nxh; m += blockDim.x) row[m] = C2<TS>{re[(m << logR) + rb], im[(m << logR) + rb]};
    }
}

// FFT along y for gx consecutive kx columns of one level; in place.
template <typename TS>
__global__ __launch_bounds__(256) void ffty_kernel(Geo g, C2<TS> *spec, const C2<TS> *tw, int logy, int nxh,
                                                   int gx, int loggx, int inverse) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C2<TS> *twl = reinterpret_cast<C2<TS> *>(smem_raw);
    TS *re = reinterpret_cast<TS *>(twl + g.ny / 2);
    TS *im = re + (long long)g.ny * gx;
    for (int m = threadIdx.x; m < g.ny / 2; m += blockDim.x) twl[m] = tw[m];
    const int kx0 = blockIdx.x * gx;
    const int kk = blockIdx.y;
    const int ncol = min(gx, nxh - kx0);
    C2<TS> *lev = spec + (long long)kk * g.ny * nxh;
    for (int t = threadIdx.x; t < g.ny * gx; t += blockDim.x) {
        const int y = t >> loggx, b = t & (gx - 1);
        const unsigned r = brev((unsigned)y, logy);
        if (b < ncol) {
            const C2<TS> z = lev[(long long)y * nxh + kx0 + b];
            re[r * gx + b] = z.x;
            im[r * gx + b] = z.y;
        } else {
            re[r * gx + b] = TS(0);
            im[r * gx + b] = TS(0);
        }
    }
    __syncthreads();
    lds_fft<TS>(re, im, g.ny, logy, gx, loggx, twl, inverse);
    const TS sc = inverse ? TS(1.0 / (double)g.ny) : TS(1);
    for (int t = threadIdx.x; t < g.ny * gx; t += blockDim.x) {
        const int y = t >> loggx, b = t & (gx - 1);
        if (b < ncol) {
            TS xr = re[y * gx + b], xi = im[y * gx + b];
            if (inverse) {
                xr *= sc;
                xi *= sc;
            }
            lev[(long long)y * nxh + kx0 + b] = C2<TS>{xr, xi};
        }
    }
}

// One thread per (kx, ky) column: TDMA of SparseSolvers.pyx:24-43 on the real
// and imaginary parts, with the diagonal of compute_diagonal
// (PressureFFTParallel.pyx:129-148); the (0,0) mode is set to zero (:196-200).
template <typename TS>
__global__ __launch_bounds__(256) void thomas_kernel(Geo g, C2<TS> *spec, double *xs, double *cs,
                                                     const double *kx2, const double *ky2, const double *rho_h,
                                                     const double *dd, const double *a, const double *c, int nxh) {
    const long long ncol = (long long)g.ny * nxh;
    const long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= ncol) return;
    const int kx = (int)(col % nxh);
    const int ky = (int)(col / nxh);
    const int nz = g.nz;
    if (kx == 0 && ky == 0) {
        for (int k = 0; k < nz; ++k) spec[(long long)k * ncol + col] = C2<TS>{TS(0), TS(0)};
        return;
    }
    const double kk = kx2[kx] + ky2[ky];
    // forward sweep
    C2<TS> z = spec[col];
    double b = rho_h[0] * kk - dd[0];
    double sc = c[0] / b;
    double xr = (double)z.x / b, xi = (double)z.y / b;
    cs[col] = sc;
    xs[2 * col] = xr;
    xs[2 * col + 1] = xi;
    for (int k = 1; k < nz; ++k) {
        const long long o = (long long)k * ncol + col;
        z = spec[o];
        b = rho_h[k] * kk - dd[k];
        const double m = 1.0 / (b - a[k] * sc);
        sc = c[k] * m;
        xr = ((double)z.x - a[k] * xr) * m;
        xi = ((double)z.y - a[k] * xi) * m;
        cs[o] = sc;
        xs[2 * o] = xr;
        xs[2 * o + 1] = xi;
    }
    // back substitution
    spec[(long long)(nz - 1) * ncol + col] = C2<TS>{(TS)xr, (TS)xi};
    for (int k = nz - 2; k >= 0; --k) {
        const long long o = (long long)k * ncol + col;
        const double s = cs[o];
        xr = xs[2 * o] - s * xr;
        xi = xs[2 * o + 1] - s * xi;
        spec[o] = C2<TS>{(TS)xr, (TS)xi};
    }
}

// Inverse x transform of one row from the half spectrum (Hermitian
// extension), real part times 1/nx, written to p at level k'+gw together with
// its mirrored z ghost level (sym, as communicate_variable does).
template <typename TS>
__global__ __launch_bounds__(256) void ifftx_kernel(Geo g, const C2<TS> *spec, TS *p, const C2<TS> *tw, int logx,
                                                    int nxh, int logR) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int R = 1 << logR;
    C2<TS> *twl = reinterpret_cast<C2<TS> *>(smem_raw);
    TS *re = reinterpret_cast<TS *>(twl + g.nx / 2);
    TS *im = re + g.nx * R;
    for (int m = threadIdx.x; m < g.nx / 2; m += blockDim.x) twl[m] = tw[m];
    const int kk = blockIdx.y;
    for (int t = threadIdx.x; t < (g.nx << logR); t += blockDim.x) {
        const int m = t & (g.nx - 1);
        const int rb = t >> logx;
        const int j = (blockIdx.x << logR) + rb;
        const C2<TS> *row = spec + ((long long)kk * g.ny + j) * nxh;
        C2<TS> z;
        if (m < nxh) z = row[m];
        else {
            z = row[g.nx - m];
            z.y = -z.y;
        }
        const unsigned r = (brev((unsigned)m, logx) << logR) + rb;
        re[r] = z.x;
        im[r] = z.y;
    }
    __syncthreads();
    lds_fft<TS>(re, im, g.nx, logx, R, logR, twl, 1);
    const TS sc = TS(1.0 / (double)g.nx);
    const int k = kk + g.gw;
    for (int t = threadIdx.x; t < (g.nx << logR); t += blockDim.x) {
        const int i = t & (g.nx - 1);
        const int rb = t >> logx;
        const int j = (blockIdx.x << logR) + rb;
        const TS val = re[(i << logR) + rb] * sc;
        p[gidx(g, i, j, k)] = val;
        if (kk < g.gw) p[gidx(g, i, j, g.gw - 1 - kk)] = val;
        if (kk >= g.nz - g.gw) p[gidx(g, i, j, 2 * g.gw + 2 * g.nz - 1 - k)] = val;
    }
}

// second_order_pressure_correction (PressureSolver.pyx:83-110) after the mean
// removal of w, then the z ghost fill of u, v (sym) and w (asym) of this
// column, and the CFL maximum of the corrected interior.
template <typename TS>
__global__ __launch_bounds__(256) void correct_kernel(Geo g, TS *u, TS *v, TS *w, const TS *p, const double *mean,
                                                      double dt, double *cfl) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    double cm = 0.0;
    if (i < g.nx) {
        const int ip = i + 1 == g.nx ? 0 : i + 1;
        const int jp = j + 1 == g.ny ? 0 : j + 1;
        const long long col = (long long)j * g.nx + i;
        for (int k = g.gw; k < g.nz + g.gw; ++k) {
            const long long o = (long long)k * g.plane + col;
            const TS pc = p[o];
            const TS pe = p[gidx(g, ip, j, k)];
            const TS pn = p[gidx(g, i, jp, k)];
            const TS pt = p[o + g.plane];
            TS un, vn, wn;
            if (sizeof(TS) == 8) {
                un = u[o] - (pe - pc) * g.dxi[0];
                vn = v[o] - (pn - pc) * g.dxi[1];
                const double wm = (double)w[o] - mean[k];
                wn = (TS)(wm - ((double)pt - (double)pc) * g.dxi[2]);
            } else {
                un = u[o] - (pe - pc) * (TS)g.dxi[0];
                vn = v[o] - (pn - pc) * (TS)g.dxi[1];
                wn = (TS)((double)w[o] - mean[k]) - (pt - pc) * (TS)g.dxi[2];
            }
            u[o] = un;
            v[o] = vn;
            w[o] = wn;
            if (cfl) {
                const double c = dt * (fabs((double)un) * g.dxi[0] + fabs((double)vn) * g.dxi[1] +
                                       fabs((double)wn) * g.dxi[2]);
                cm = fmax(cm, c);
            }
        }
        for (int m = 0; m < g.gw; ++m) {
            u[gidx(g, i, j, g.gw - 1 - m)] = u[gidx(g, i, j, g.gw + m)];
            u[gidx(g, i, j, g.nzg - g.gw + m)] = u[gidx(g, i, j, g.nzg - g.gw - 1 - m)];
            v[gidx(g, i, j, g.gw - 1 - m)] = v[gidx(g, i, j, g.gw + m)];
            v[gidx(g, i, j, g.nzg - g.gw + m)] = v[gidx(g, i, j, g.nzg - g.gw - 1 - m)];
        }
        w[gidx(g, i, j, g.gw - 1)] = TS(0);
        w[gidx(g, i, j, g.nzg - g.gw)] = TS(0);
        for (int m = 1; m < g.gw; ++m) {
            w[gidx(g, i, j, g.gw - 1 - m)] = TS(-1.0) * w[gidx(g, i, j, g.gw - 1 + m)];
            w[gidx(g, i, j, g.nzg - g.gw + m)] = TS(-1.0) * w[gidx(g, i, j, g.nzg - g.gw - m)];
        }
    }
    if (cfl) {
        __shared__ double red[256];
        red[threadIdx.x] = cm;
        __syncthreads();
        for (int s = blockDim.x / 2; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
            __syncthreads();
        }
        if (threadIdx.x == 0) atomicMax((unsigned long long *)cfl, (unsigned long long)__double_as_longlong(red[0]));
    }
}

__global__ void cfl_zero_kernel(double *c) { *c = 0.0; }
__global__ void cfl_add_kernel(const double *src, double *dst) { *dst = *src + 1e-11; }

// Standalone divergence (interior) in the field precision.
template <typename TS>
__global__ void divergence_kernel(Geo g, const TS *u, const TS *v, const TS *w, const double *alpha0,
                                  const double *alpha0_half, TS *div) {
    const long long n = g.plane * g.nz;
    const double dxi0 = 1.0 / g.dx[0], dxi1 = 1.0 / g.dx[1], dxi2 = 1.0 / g.dx[2];
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(t % g.nx);
        const int j = (int)((t / g.nx) % g.ny);
        const int k = (int)(t / g.plane) + g.gw;
        const int im1 = i == 0 ? g.nx - 1 : i - 1;
        const int jm = j == 0 ? g.ny - 1 : j - 1;
        const long long o = gidx(g, i, j, k);
        const double ah = alpha0_half[k];
        double dv = 0.0;
        dv += ((double)u[o] / ah - (double)u[gidx(g, im1, j, k)] / ah) * dxi0;
        dv += ((double)v[o] / ah - (double)v[gidx(g, i, jm, k)] / ah) * dxi1;
        dv += (((double)w[o]) / alpha0[k] - ((double)w[o - g.plane]) / alpha0[k - 1]) * dxi2;
        div[o] = (TS)dv;
    }
}

// ---------------------------------------------------------------------------
extern "C" int pycles_pressure_plan_create(const pycles_dims *dims, const double *rho0, const double *rho0_half,
                                           int precision, pycles_pressure_plan **out) {
    int rc = pycles_check_dims(dims);
    if (rc) return rc;
    if (!out || !rho0 || !rho0_half || (precision != PYCLES_F64 && precision != PYCLES_F32)) {
        pycles_set_error("pressure_plan_create: bad arguments");
        return PYCLES_ERR_ARG;
    }
    Geo g = make_geo(dims);
    const int logx = ilog2(g.nx), logy = ilog2(g.ny);
    if (logx < 1 || logy < 1 || g.nx > 4096 || g.ny > 4096) {
        pycles_set_error("pressure: nx, ny must be powers of two in [2, 4096] (got %d, %d)", g.nx, g.ny);
        return PYCLES_ERR_UNSUPPORTED;
    }
    if (g.nz < 2) {
        pycles_set_error("pressure: nz must be at least 2");
        return PYCLES_ERR_UNSUPPORTED;
    }
    pycles_pressure_plan *P = new pycles_pressure_plan();
    memset(P, 0, sizeof(*P));
    P->g = g;
    P->precision = precision;
    P->nxh = g.nx / 2 + 1;
    P->logx = logx;
    P->logy = logy;
    const size_t cs = precision == PYCLES_F64 ? 16 : 8;
    const size_t es = cs / 2;
    int gx = 16;
    while (gx > 1 && (size_t)2 * g.ny * gx * es + (size_t)g.ny / 2 * cs > 64 * 1024) gx >>= 1;
    P->gx = gx;
    P->loggx = ilog2(gx);
    const long long nspec = (long long)g.nz * g.ny * P->nxh;
    const int gw = g.gw, nz = g.nz;
    // host tables (same expressions as the reference initializers)
    std::vector<double> kx2(P->nxh), ky2(g.ny), rh(nz), dd(nz), a(nz), c(nz);
    for (int i = 0; i < P->nxh; ++i) {
        const double xi = (i <= g.nx / 2) ? (double)i : (double)(i - g.nx);
        kx2[i] = (2.0 * cos((2.0 * PC_PI / g.nx) * xi) - 2.0) / g.dx[0] / g.dx[0];
    }
    kx2[0] = 0.0;
    for (int j = 0; j < g.ny; ++j) {
        const double yi = (j <= g.ny / 2) ? (double)j : (double)(j - g.ny);
        ky2[j] = (2.0 * cos((2.0 * PC_PI / g.ny) * yi) - 2.0) / g.dx[1] / g.dx[1];
    }
    ky2[0] = 0.0;
    const double dzi = g.dxi[2];
    for (int k = 0; k < nz; ++k) {
        rh[k] = rho0_half[k + gw];
        if (k == 0) dd[k] = (rho0[gw]) * dzi * dzi;
        else if (k == nz - 1) dd[k] = (rho0[k + gw - 1]) * dzi * dzi;
        else dd[k] = (rho0[k + gw] + rho0[k + gw - 1]) * dzi * dzi;
        a[k] = (k == 0) ? 0.0 : dzi * dzi * rho0[k + gw - 1];
        c[k] = (k == nz - 1) ? 0.0 : dzi * dzi * rho0[k + gw];
    }
    std::vector<double> twx(g.nx), twy(g.ny);
    for (int m = 0; m < g.nx / 2; ++m) {
        const double ang = -2.0 * PC_PI * (double)m / (double)g.nx;
        twx[2 * m] = cos(ang);
        twx[2 * m + 1] = sin(ang);
    }
    for (int m = 0; m < g.ny / 2; ++m) {
        const double ang = -2.0 * PC_PI * (double)m / (double)g.ny;
        twy[2 * m] = cos(ang);
        twy[2 * m + 1] = sin(ang);
    }
#define ALLOC(ptr, bytes)                                   \
    do {                                                    \
        if (hipMalloc((void **)&(ptr), (bytes)) != hipSuccess) { \
            pycles_set_error("pressure: hipMalloc of %zu bytes failed", (size_t)(bytes)); \
            pycles_pressure_plan_destroy(P);                \
            return PYCLES_ERR_HIP;                          \
        }                                                   \
    } while (0)
    ALLOC(P->spec, nspec * cs);
    ALLOC(P->xs, nspec * 16);
    ALLOC(P->cs, nspec * 8);
    ALLOC(P->twx, (size_t)g.nx / 2 * cs + cs);
    ALLOC(P->twy, (size_t)g.ny / 2 * cs + cs);
    ALLOC(P->kx2, P->nxh * 8);
    ALLOC(P->ky2, g.ny * 8);
    ALLOC(P->rho_h, nz * 8);
    ALLOC(P->dd, nz * 8);
    ALLOC(P->a, nz * 8);
    ALLOC(P->c, nz * 8);
    ALLOC(P->mean, (size_t)g.nzg * 8);
    ALLOC(P->partial, (size_t)g.nzg * 16 * 8);
    ALLOC(P->cfl, 16);
#undef ALLOC
    hipError_t e = hipSuccess;
    if (precision == PYCLES_F64) {
        e = hipMemcpy(P->twx, twx.data(), g.nx / 2 * 16, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(P->twy, twy.data(), g.ny / 2 * 16, hipMemcpyHostToDevice);
    } else {
        std::vector<float> fx(g.nx), fy(g.ny);
        for (int m = 0; m < g.nx; ++m) fx[m] = (float)twx[m];
        for (int m = 0; m < g.ny; ++m) fy[m] = (float)twy[m];
        e = hipMemcpy(P->twx, fx.data(), g.nx / 2 * 8, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(P->twy, fy.data(), g.ny / 2 * 8, hipMemcpyHostToDevice);
    }
    if (e == hipSuccess) e = hipMemcpy(P->kx2, kx2.data(), P->nxh * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(P->ky2, ky2.data(), g.ny * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(P->rho_h, rh.data(), nz * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(P->dd, dd.data(), nz * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(P->a, a.data(), nz * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(P->c, c.data(), nz * 8, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        pycles_set_error("pressure: table upload failed: %s", hipGetErrorString(e));
        pycles_pressure_plan_destroy(P);
        return PYCLES_ERR_HIP;
    }
    *out = P;
    return PYCLES_OK;
}

extern "C" void pycles_pressure_plan_destroy(pycles_pressure_plan *P) {
    if (!P) return;
    void *ptrs[] = {P->spec, P->xs, P->cs, P->twx, P->twy, P->kx2, P->ky2, P->rho_h,
                    P->dd, P->a, P->c, P->mean, P->partial, P->cfl};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    delete P;
}

template <typename TS>
static int run_pressure(pycles_pressure_plan *P, const pycles_ref *ref, void *u, void *v, void *w, void *p,
                        void *div, double dt, double *cfl_out, hipStream_t st) {
    const Geo &g = P->g;
    int rc = pycles_level_mean_ws(g, w, P->mean, P->partial, nullptr, P->precision, st);
    if (rc) return rc;
    // rows per block for the x transforms: about 2048 points per block
    int logR = 0;
    while ((g.nx << (logR + 1)) <= 2048 && (1 << (logR + 1)) <= g.ny && logR < 4) ++logR;
    const int R = 1 << logR;
    const size_t rowsmem = 2 * sizeof(TS) * g.nx * R + sizeof(C2<TS>) * (g.nx / 2);
    const int bx = 256;
    hipLaunchKernelGGL(div_fftx_kernel<TS>, dim3(g.ny / R, g.nz), dim3(bx), rowsmem, st, g, (const TS *)u,
                       (const TS *)v, (const TS *)w, ref->alpha0, ref->alpha0_half, P->mean, (TS *)div,
                       (C2<TS> *)P->spec, (const C2<TS> *)P->twx, P->logx, P->nxh, logR);
    PYCLES_CHECK_LAUNCH();
    const size_t ysmem = 2 * sizeof(TS) * (size_t)g.ny * P->gx + sizeof(C2<TS>) * (g.ny / 2);
    const dim3 ygrid((P->nxh + P->gx - 1) / P->gx, g.nz);
    hipLaunchKernelGGL(ffty_kernel<TS>, ygrid, dim3(256), ysmem, st, g, (C2<TS> *)P->spec,
                       (const C2<TS> *)P->twy, P->logy, P->nxh, P->gx, P->loggx, 0);
    PYCLES_CHECK_LAUNCH();
    const long long ncol = (long long)g.ny * P->nxh;
    hipLaunchKernelGGL(thomas_kernel<TS>, dim3((unsigned)((ncol + 255) / 256)), dim3(256), 0, st, g,
                       (C2<TS> *)P->spec, P->xs, P->cs, P->kx2, P->ky2, P->rho_h, P->dd, P->a, P->c, P->nxh);
    PYCLES_CHECK_LAUNCH();
    hipLaunchKernelGGL(ffty_kernel<TS>, ygrid, dim3(256), ysmem, st, g, (C2<TS> *)P->spec,
                       (const C2<TS> *)P->twy, P->logy, P->nxh, P->gx, P->loggx, 1);
    PYCLES_CHECK_LAUNCH();
    hipLaunchKernelGGL(ifftx_kernel<TS>, dim3(g.ny / R, g.nz), dim3(bx), rowsmem, st, g, (const C2<TS> *)P->spec,
                       (TS *)p, (const C2<TS> *)P->twx, P->logx, P->nxh, logR);
    PYCLES_CHECK_LAUNCH();
    if (cfl_out) hipLaunchKernelGGL(cfl_zero_kernel, dim3(1), dim3(1), 0, st, P->cfl);
    hipLaunchKernelGGL(correct_kernel<TS>, dim3((g.nx + 255) / 256, g.ny), dim3(256), 0, st, g, (TS *)u, (TS *)v,
                       (TS *)w, (const TS *)p, P->mean, dt, cfl_out ? P->cfl : nullptr);
    PYCLES_CHECK_LAUNCH();
    if (cfl_out) {
        hipLaunchKernelGGL(cfl_add_kernel, dim3(1), dim3(1), 0, st, P->cfl, cfl_out);
        PYCLES_CHECK_LAUNCH();
    }
    return PYCLES_OK;
}

extern "C" int pycles_pressure_update(pycles_pressure_plan *P, const pycles_ref *ref, void *u, void *v, void *w,
                                      void *p, void *div, double dt, double *cfl_out, void *stream) {
    if (!P || !ref || !u || !v || !w || !p) {
        pycles_set_error("pressure_update: NULL argument");
        return PYCLES_ERR_ARG;
    }
    hipStream_t st = (hipStream_t)stream;
    if (P->precision == PYCLES_F64) return run_pressure<double>(P, ref, u, v, w, p, div, dt, cfl_out, st);
    return run_pressure<float>(P, ref, u, v, w, p, div, dt, cfl_out, st);
}

extern "C" int pycles_divergence(const pycles_dims *dims, const pycles_ref *ref, const void *u, const void *v,
                                 const void *w, void *div, int precision, void *stream) {
    int rc = pycles_check_dims(dims);
    if (rc) return rc;
    if (!ref || !u || !v || !w || !div) {
        pycles_set_error("divergence: NULL argument");
        return PYCLES_ERR_ARG;
    }
    Geo g = make_geo(dims);
    hipStream_t st = (hipStream_t)stream;
    if (precision == PYCLES_F64)
        hipLaunchKernelGGL(divergence_kernel<double>, dim3(2048), dim3(256), 0, st, g, (const double *)u,
                           (const double *)v, (const double *)w, ref->alpha0, ref->alpha0_half, (double *)div);
    else
        hipLaunchKernelGGL(divergence_kernel<float>, dim3(2048), dim3(256), 0, st, g, (const float *)u,
                           (const float *)v, (const float *)w, ref->alpha0, ref->alpha0_half, (float *)div);
    PYCLES_CHECK_LAUNCH();
    return PYCLES_OK;
}
