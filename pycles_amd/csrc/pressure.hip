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
    int nxp;         // row stride of the spectrum: nxh rounded up to 16 (128-byte rows in FP32)
    int gx;          // kx columns per y-FFT block
    void *spec;      // [nz][ny][nxp] complex, field precision (columns nxh..nxp-1 unused)
    double *ck;      // Thomas segment carries: [nz / S][3][ny][nxp] (S levels per segment)
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

template <typename T>
__device__ __forceinline__ C2<T> cmul(C2<T> a, C2<T> b) {
    return C2<T>{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ void bfly(C2<T> &a, C2<T> &b) {
    const C2<T> t = a;
    a = C2<T>{t.x + b.x, t.y + b.y};
    b = C2<T>{t.x - b.x, t.y - b.y};
}
// times -i (forward) or +i (inverse)
template <typename T, bool INV>
__device__ __forceinline__ C2<T> rot90(C2<T> a) {
    return INV ? C2<T>{-a.y, a.x} : C2<T>{a.y, -a.x};
}

// In-register DFTs of size 2, 4, 8, natural order in and out, kernel
// exp(-+2 pi i n k / R) (forward: minus).
template <typename T, bool INV>
__device__ __forceinline__ void dft2(C2<T> *v) {
    bfly(v[0], v[1]);
}
template <typename T, bool INV>
__device__ __forceinline__ void dft4(C2<T> *v) {
    bfly(v[0], v[2]);
    bfly(v[1], v[3]);
    v[3] = rot90<T, INV>(v[3]);
    bfly(v[0], v[1]);
    bfly(v[2], v[3]);
    const C2<T> t = v[1];
    v[1] = v[2];
    v[2] = t;
}
template <typename T, bool INV>
__device__ __forceinline__ void dft8(C2<T> *v) {
    C2<T> e[4] = {v[0], v[2], v[4], v[6]};
    C2<T> o[4] = {v[1], v[3], v[5], v[7]};
    dft4<T, INV>(e);
    dft4<T, INV>(o);
    const T h = T(0.70710678118654752440);
    // o[k] *= exp(-+ i pi k / 4)
    o[1] = INV ? C2<T>{h * (o[1].x - o[1].y), h * (o[1].x + o[1].y)}
               : C2<T>{h * (o[1].x + o[1].y), h * (o[1].y - o[1].x)};
    o[2] = rot90<T, INV>(o[2]);
    o[3] = INV ? C2<T>{-h * (o[3].x + o[3].y), h * (o[3].x - o[3].y)}
               : C2<T>{h * (o[3].y - o[3].x), -h * (o[3].x + o[3].y)};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        v[k] = C2<T>{e[k].x + o[k].x, e[k].y + o[k].y};
        v[k + 4] = C2<T>{e[k].x - o[k].x, e[k].y - o[k].y};
    }
}
template <typename T, int R, bool INV>
__device__ __forceinline__ void dft(C2<T> *v) {
    if (R == 2) dft2<T, INV>(v);
    else if (R == 4) dft4<T, INV>(v);
    else dft8<T, INV>(v);
}

// One radix-R pass of a Stockham autosort FFT over S interleaved length-N
// sequences in LDS (element e of sequence b at buf[e * S + b]).  Thread t
// (0 .. N/8-1) of sequence b does the 8/R butterflies q = t + s N/8: it reads
// x[q + r N/R], twiddles by w_N^{k r N/(Ns R)} with k = q mod Ns, transforms,
// and writes to (q - k) R + k + r Ns.  Ns is the length of the sub-transforms
// already combined.  All threads of the block take part (two barriers).
template <typename T, int R, bool INV>
__device__ __forceinline__ void stockham_pass(C2<T> *buf, int S, int b, int t, int N, int Ns, const C2<T> *tw) {
    constexpr int NB = 8 / R;
    const int NR = N / R, N8 = N >> 3;
    C2<T> v[NB][R];
#pragma unroll
    for (int s = 0; s < NB; ++s) {
        const int q = t + s * N8;
#pragma unroll
        for (int r = 0; r < R; ++r) v[s][r] = buf[(q + r * NR) * S + b];
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < NB; ++s) {
        const int q = t + s * N8;
        const int k = q & (Ns - 1);
        if (Ns > 1) {
            const int step = k * (N / (Ns * R));
#pragma unroll
            for (int r = 1; r < R; ++r) {
                C2<T> w = tw[step * r];
                if (INV) w.y = -w.y;
                v[s][r] = cmul(v[s][r], w);
            }
        }
        dft<T, R, INV>(v[s]);
        const int base = (q - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) buf[(base + r * Ns) * S + b] = v[s][r];
    }
    __syncthreads();
}

// Unnormalised DFT (forward: exp(-2 pi i n k/N); inverse: +) of the S
// sequences in buf, N = 2^logN >= 8, in natural order in and out.
// tw[m] = exp(-2 pi i m/N), m < N.
template <typename T, bool INV>
__device__ void stockham_fft(C2<T> *buf, int S, int b, int t, int N, int logN, const C2<T> *tw) {
    int Ns = 1;
    const int rem = logN % 3;
    if (rem == 1) {
        stockham_pass<T, 2, INV>(buf, S, b, t, N, Ns, tw);
        Ns = 2;
    } else if (rem == 2) {
        stockham_pass<T, 4, INV>(buf, S, b, t, N, Ns, tw);
        Ns = 4;
    }
    for (int p = 0; p < logN / 3; ++p) {
        stockham_pass<T, 8, INV>(buf, S, b, t, N, Ns, tw);
        Ns *= 8;
    }
}

// ---------------------------------------------------------------------------
// Level mean of w is computed before this kernel.  One block per 2S
// consecutive rows (j0 .. j0+2S-1) of one level: divergence
// (PressureSolver.pyx:138-175, zeroed then x, y, z contributions added in that
// order), then the forward FFT along x of the rows in pairs (row 2b as the
// real part and row 2b+1 as the imaginary part of one complex sequence),
// separated into the two half spectra.
template <typename TS>
__global__ __launch_bounds__(512) void div_fftx_kernel(Geo g, const TS *u, const TS *v, const TS *w,
                                                       const double *alpha0, const double *alpha0_half,
                                                       const double *mean, TS *div_out, C2<TS> *spec,
                                                       const C2<TS> *tw, int logx, int nxh, int nxp, int logS) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int S = 1 << logS;
    C2<TS> *buf = reinterpret_cast<C2<TS> *>(smem_raw);
    TS *bufr = reinterpret_cast<TS *>(smem_raw);
    const int kk = blockIdx.y;
    const int k = kk + g.gw;
    const int j0 = blockIdx.x << (logS + 1);
    const double ah = alpha0_half[k];
    const double a0 = alpha0[k], a0m = alpha0[k - 1];
    const double mk = mean[k], mkm = mean[k - 1];
    const double dxi0 = 1.0 / g.dx[0], dxi1 = 1.0 / g.dx[1], dxi2 = 1.0 / g.dx[2];
    // four cells per thread at a time: all their loads are issued before the
    // arithmetic and the stores (a store in between would hold back the
    // next cell's loads, which the compiler must assume it aliases)
    const int B = blockDim.x, nt = g.nx << (logS + 1);
    for (int t0 = threadIdx.x; t0 < nt; t0 += 4 * B) {
        TS uc[4], um[4], vc[4], vm[4], wc[4], wm[4];
        long long oo[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int t = min(t0 + q * B, nt - 1);  // a clamped duplicate past the end: loaded, not stored
            const int i = t & (g.nx - 1);
            const int j = j0 + (t >> logx);
            const int jm = j == 0 ? g.ny - 1 : j - 1;
            const int im1 = i == 0 ? g.nx - 1 : i - 1;
            const long long o = gidx(g, i, j, k);
            oo[q] = o;
            uc[q] = u[o];
            um[q] = u[gidx(g, im1, j, k)];
            vc[q] = v[o];
            vm[q] = v[gidx(g, i, jm, k)];
            wc[q] = w[o];
            wm[q] = w[o - g.plane];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int t = t0 + q * B;
            if (t >= nt) continue;
            const int i = t & (g.nx - 1);
            const int rb = t >> logx;
            TS d;
            if (sizeof(TS) == 8) {
                // the profiles are positive normal numbers: div64 is the
                // IEEE quotient (common.h)
                double dv = 0.0;
                dv += (div64((double)uc[q], ah) - div64((double)um[q], ah)) * dxi0;
                dv += (div64((double)vc[q], ah) - div64((double)vm[q], ah)) * dxi1;
                const double wk = (double)wc[q] - mk;
                const double wkm = (double)wm[q] - mkm;
                dv += (div64(wk, a0) - div64(wkm, a0m)) * dxi2;
                d = (TS)dv;
            } else {
                // FP32 path: rho0*u in FP32 except the level means and profiles
                const float rh = (float)(1.0 / ah);
                float dv = ((float)uc[q] - (float)um[q]) * (float)(dxi0)*rh;
                dv += ((float)vc[q] - (float)vm[q]) * (float)(dxi1)*rh;
                const float wk = (float)((double)wc[q] - mk);
                const float wkm = (float)((double)wm[q] - mkm);
                dv += (wk * (float)(1.0 / a0) - wkm * (float)(1.0 / a0m)) * (float)dxi2;
                d = (TS)dv;
            }
            if (div_out) div_out[oo[q]] = d;
            bufr[2 * (i * S + (rb >> 1)) + (rb & 1)] = d;
        }
    }
    __syncthreads();
    const int b = threadIdx.x & (S - 1), t = threadIdx.x >> logS;
    stockham_fft<TS, false>(buf, S, b, t, g.nx, logx, tw);
    // X_a[m] = (Z[m] + conj Z[N-m]) / 2, X_b[m] = -i (Z[m] - conj Z[N-m]) / 2
    for (int q = threadIdx.x; q < nxh * S; q += blockDim.x) {
        const int m = q % nxh, pb = q / nxh;
        const C2<TS> z = buf[m * S + pb];
        const C2<TS> zm = buf[((g.nx - m) & (g.nx - 1)) * S + pb];
        const C2<TS> xa{TS(0.5) * (z.x + zm.x), TS(0.5) * (z.y - zm.y)};
        const C2<TS> xb{TS(0.5) * (z.y + zm.y), TS(-0.5) * (z.x - zm.x)};
        spec[((long long)kk * g.ny + j0 + 2 * pb) * nxp + m] = xa;
        spec[((long long)kk * g.ny + j0 + 2 * pb + 1) * nxp + m] = xb;
    }
}

// FFT along y for gx consecutive kx columns of one level, in place
// (inverse: times 1/ny).
template <typename TS>
__global__ __launch_bounds__(1024) void ffty_kernel(Geo g, C2<TS> *spec, const C2<TS> *tw, int logy, int nxh,
                                                   int nxp, int gx, int loggx, int inverse) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C2<TS> *buf = reinterpret_cast<C2<TS> *>(smem_raw);
    const int kx0 = blockIdx.x * gx;
    const int kk = blockIdx.y;
    const int ncol = min(gx, nxh - kx0);
    C2<TS> *lev = spec + (long long)kk * g.ny * nxp;
    // staged four elements per thread at a time, loads before LDS writes
    const int B = blockDim.x, nq = g.ny * gx;
    for (int q0 = threadIdx.x; q0 < nq; q0 += 4 * B) {
        C2<TS> z[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int q = q0 + u * B;
            const int y = q >> loggx, b = q & (gx - 1);
            z[u] = (q < nq && b < ncol) ? lev[(long long)y * nxp + kx0 + b] : C2<TS>{TS(0), TS(0)};
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (q0 + u * B < nq) buf[q0 + u * B] = z[u];
    }
    __syncthreads();
    const int b = threadIdx.x & (gx - 1), t = threadIdx.x >> loggx;
    if (inverse) stockham_fft<TS, true>(buf, gx, b, t, g.ny, logy, tw);
    else stockham_fft<TS, false>(buf, gx, b, t, g.ny, logy, tw);
    const TS sc = inverse ? TS(1.0 / (double)g.ny) : TS(1);
    for (int q = threadIdx.x; q < g.ny * gx; q += blockDim.x) {
        const int y = q >> loggx, bb = q & (gx - 1);
        if (bb < ncol) {
            C2<TS> z = buf[q];
            if (inverse) {
                z.x *= sc;
                z.y *= sc;
            }
            lev[(long long)y * nxp + kx0 + bb] = z;
        }
    }
}

// Largest nz of the pressure solve: the Thomas kernel holds four FP64 level
// profiles in LDS (64 KB at this size).
constexpr int PRESSURE_MAX_NZ = 2048;

// Levels per segment of the checkpointed Thomas solve: the right-hand sides
// of two segments are held in registers, twice the registers per level in FP64.
template <typename TS> constexpr int thomas_seg() { return sizeof(TS) == 4 ? 16 : 8; }

// 1/d in FP64 from the hardware reciprocal and two Newton steps (within an
// ulp or two of the division; no special values occur: d is a pivot of a
// diagonally dominant system).
__device__ __forceinline__ double rcp64(double d) {
    double r = __builtin_amdgcn_rcp(d);
    double e = __builtin_fma(-d, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-d, r, 1.0);
    return __builtin_fma(r, e, r);
}

// One thread per (kx, ky) column: TDMA of SparseSolvers.pyx:24-43 on the real
// and imaginary parts, with the diagonal of compute_diagonal
// (PressureFFTParallel.pyx:129-148); the (0,0) mode is set to zero (:196-200).
// The recurrence runs in FP64 whatever the storage type TS.  No per-level
// scratch: the forward sweep streams the right-hand sides and keeps only the
// FP64 carry (c', x'r, x'i) at the last level of every S-level segment (3
// doubles per column per segment, in ck).  The back substitution walks the
// segments from the top: it recomputes a segment's forward values from the
// carry below it and the right-hand sides, which are still in place (the
// solution overwrites a segment only after its forward values are
// recomputed), then substitutes down the segment.  The next segment's
// right-hand sides and carry are loaded while the current one is worked on.
// Per column-level: the right-hand side read twice and the solution written
// (24 B in FP32, 48 B in FP64), where a stored scratch of forward values
// would add its write and read back (40 and 80 B).  Both sweeps evaluate the
// same FP64 expressions, so the recomputed values are the forward sweep's,
// bit for bit; the reciprocal is rcp64, not an IEEE division.  The forward
// values enter the back substitution rounded to TS, as the reference's
// in-place arrays would hold them in that precision.
template <typename TS, int S>
__global__ __launch_bounds__(256, 3) void thomas_kernel(Geo g, C2<TS> *spec, double *ck, const double *kx2,
                                                        const double *ky2, const double *rho_h, const double *dd,
                                                        const double *a, const double *c, int nxh, int nxp) {
    // The level profiles (rho_h, dd, a, c) in LDS, 4 nz doubles: read from
    // global memory in the level loops they are vector loads (the stores
    // keep the compiler from the scalar cache), and each wait for one would
    // also wait for every older load and store, right-hand sides included.
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double *lv = reinterpret_cast<double *>(smem_raw);
    const int nz = g.nz;
    for (int t = threadIdx.x; t < nz; t += blockDim.x) {
        lv[t] = rho_h[t];
        lv[nz + t] = dd[t];
        lv[2 * nz + t] = a[t];
        lv[3 * nz + t] = c[t];
    }
    __syncthreads();
    const double *lrho = lv, *ldd = lv + nz, *la = lv + 2 * nz, *lc = lv + 3 * nz;
    const long long ncol = (long long)g.ny * nxp;
    const long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= ncol) return;
    const int kx = (int)(col % nxp);
    const int ky = (int)(col / nxp);
    if (kx >= nxh) return;
    if (kx == 0 && ky == 0) {
        for (int k = 0; k < nz; ++k) spec[(long long)k * ncol + col] = C2<TS>{TS(0), TS(0)};
        return;
    }
    const double kk = kx2[kx] + ky2[ky];
    auto at = [&](int k) { return spec + (long long)k * ncol + col; };
    // one forward step: the carry of level k - 1 to level k
    auto fstep = [&](int k, C2<TS> z, double &sc, double &xr, double &xi) {
        const double b = lrho[k] * kk - ldd[k];
        if (k == 0) {
            const double ib = rcp64(b);
            sc = lc[0] * ib;
            xr = (double)z.x * ib;
            xi = (double)z.y * ib;
        } else {
            const double m = rcp64(b - la[k] * sc);
            sc = lc[k] * m;
            xr = ((double)z.x - la[k] * xr) * m;
            xi = ((double)z.y - la[k] * xi) * m;
        }
    };
    // forward sweep: right-hand sides S levels ahead, loaded at every level
    // (clamped), a carry stored at the end of every whole segment
    double sc = 0.0, xr = 0.0, xi = 0.0;
    C2<TS> zq[S];
#pragma unroll
    for (int u = 0; u < S; ++u) zq[u] = *at(min(u, nz - 1));
    int k0 = 0;
    for (; k0 + S <= nz; k0 += S) {
#pragma unroll
        for (int u = 0; u < S; ++u) {
            const int k = k0 + u;
            const C2<TS> z = zq[u];
            zq[u] = *at(min(k + S, nz - 1));
            fstep(k, z, sc, xr, xi);
        }
        const long long q = 3LL * (k0 / S) * ncol + col;
        ck[q] = sc;
        ck[q + ncol] = xr;
        ck[q + 2 * ncol] = xi;
    }
#pragma unroll
    for (int u = 0; u < S; ++u)
        if (k0 + u < nz) fstep(k0 + u, zq[u], sc, xr, xi);
    // the top level's solution is its forward value
    *at(nz - 1) = C2<TS>{(TS)xr, (TS)xi};
    // back substitution, segment by segment from the top; (xr, xi) carries
    // the solution of the level above, in FP64.  Segment sg needs the carry
    // of segment sg - 1 (none for sg = 0, whose first step starts afresh);
    // the loads of the next segment down use a clamped index, and with a
    // single segment no carry is read.
    const int nseg = (nz + S - 1) / S;
    const int top = nz - 2 < 0 ? 0 : nz - 2;  // the highest level still holding its right-hand side
    C2<TS> zs[S];
    double csc = 0.0, cxr = 0.0, cxi = 0.0;
    auto load_seg = [&](int sg, C2<TS> *z, double &lsc, double &lxr, double &lxi) {
#pragma unroll
        for (int u = 0; u < S; ++u) z[u] = *at(min(sg * S + u, top));
        if (nseg > 1) {
            const long long q = 3LL * max(sg - 1, 0) * ncol + col;
            lsc = ck[q];
            lxr = ck[q + ncol];
            lxi = ck[q + 2 * ncol];
        }
    };
    load_seg(nseg - 1, zs, csc, cxr, cxi);
    for (int sg = nseg - 1; sg >= 0; --sg) {
        const int lo = sg * S;
        C2<TS> zn[S];
        double nsc = 0.0, nxr = 0.0, nxi = 0.0;
        load_seg(max(sg - 1, 0), zn, nsc, nxr, nxi);
        TS sT[S], rT[S], iT[S];
#pragma unroll
        for (int u = 0; u < S; ++u) {
            if (lo + u < nz - 1) {
                fstep(lo + u, zs[u], csc, cxr, cxi);
                sT[u] = (TS)csc;
                rT[u] = (TS)cxr;
                iT[u] = (TS)cxi;
            }
        }
#pragma unroll
        for (int u = S - 1; u >= 0; --u) {
            const int k = lo + u;
            if (k < nz - 1) {
                xr = (double)rT[u] - (double)sT[u] * xr;
                xi = (double)iT[u] - (double)sT[u] * xi;
                *at(k) = C2<TS>{(TS)xr, (TS)xi};
            }
        }
#pragma unroll
        for (int u = 0; u < S; ++u) zs[u] = zn[u];
        csc = nsc;
        cxr = nxr;
        cxi = nxi;
    }
}

// Inverse x transform of 2S rows in pairs: Z[m] = X_a[m] + i X_b[m] with the
// Hermitian extension X[N-m] = conj X[m], one inverse FFT, real part -> row
// 2b and imaginary part -> row 2b+1, times 1/nx, written to p at level k'+gw
// together with the mirrored z ghost level (sym, as communicate_variable does).
template <typename TS>
__global__ __launch_bounds__(512) void ifftx_kernel(Geo g, const C2<TS> *spec, TS *p, const C2<TS> *tw, int logx,
                                                    int nxh, int nxp, int logS) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int S = 1 << logS;
    C2<TS> *buf = reinterpret_cast<C2<TS> *>(smem_raw);
    const TS *bufr = reinterpret_cast<const TS *>(smem_raw);
    const int kk = blockIdx.y;
    const int j0 = blockIdx.x << (logS + 1);
    // staged four elements per thread at a time: their loads are all issued
    // before the LDS writes (one memory round trip per four, not per one)
    const int B = blockDim.x, nq = g.nx << logS;
    for (int q0 = threadIdx.x; q0 < nq; q0 += 4 * B) {
        C2<TS> xa[4], xb[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int q = q0 + u * B;
            if (q < nq) {
                const int m = q & (g.nx - 1);
                const int pb = q >> logx;
                const C2<TS> *ra = spec + ((long long)kk * g.ny + j0 + 2 * pb) * nxp;
                const int mm = m < nxh ? m : g.nx - m;
                xa[u] = ra[mm];
                xb[u] = ra[nxp + mm];
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int q = q0 + u * B;
            if (q < nq) {
                const int m = q & (g.nx - 1);
                const int pb = q >> logx;
                if (m >= nxh) {  // Hermitian extension X[N-m] = conj X[m]
                    xa[u].y = -xa[u].y;
                    xb[u].y = -xb[u].y;
                }
                buf[m * S + pb] = C2<TS>{xa[u].x - xb[u].y, xa[u].y + xb[u].x};
            }
        }
    }
    __syncthreads();
    const int b = threadIdx.x & (S - 1), t = threadIdx.x >> logS;
    stockham_fft<TS, true>(buf, S, b, t, g.nx, logx, tw);
    const TS sc = TS(1.0 / (double)g.nx);
    const int k = kk + g.gw;
    for (int q = threadIdx.x; q < (g.nx << (logS + 1)); q += blockDim.x) {
        const int i = q & (g.nx - 1);
        const int rb = q >> logx;
        const int j = j0 + rb;
        const TS val = bufr[2 * (i * S + (rb >> 1)) + (rb & 1)] * sc;
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
                                                      double dt, double *cfl, int kchunk) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    // chunks of kchunk levels; the last one also takes the remainder, so
    // every chunk holds at least kchunk >= gw levels (or the whole column)
    const int k0 = g.gw + blockIdx.z * kchunk;
    const int k1 = blockIdx.z + 1 == gridDim.z ? g.nz + g.gw : k0 + kchunk;
    double cm = 0.0;
    if (i < g.nx) {
        const int ip = i + 1 == g.nx ? 0 : i + 1;
        const int jp = j + 1 == g.ny ? 0 : j + 1;
        const long long col = (long long)j * g.nx + i;
        const long long ce = (long long)j * g.nx + ip;
        const long long cn = (long long)jp * g.nx + i;
        TS pc = p[(long long)k0 * g.plane + col];
        // four levels at a time: all their loads before their stores (a store
        // would hold back the next level's loads, which it may alias)
        for (int kb = k0; kb < k1; kb += 4) {
            TS uq[4], vq[4], wq[4], peq[4], pnq[4], ptq[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int k = min(kb + q, k1 - 1);  // a clamped duplicate past the end: loaded, not stored
                const long long lo = (long long)k * g.plane;
                const long long o = lo + col;
                peq[q] = p[lo + ce];
                pnq[q] = p[lo + cn];
                ptq[q] = p[o + g.plane];
                uq[q] = u[o];
                vq[q] = v[o];
                wq[q] = w[o];
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int k = kb + q;
                if (k >= k1) break;
                const long long o = (long long)k * g.plane + col;
                const TS pe = peq[q], pn = pnq[q], pt = ptq[q];
                TS un, vn, wn;
                if (sizeof(TS) == 8) {
                    un = uq[q] - (pe - pc) * g.dxi[0];
                    vn = vq[q] - (pn - pc) * g.dxi[1];
                    const double wm = (double)wq[q] - mean[k];
                    wn = (TS)(wm - ((double)pt - (double)pc) * g.dxi[2]);
                } else {
                    un = uq[q] - (pe - pc) * (TS)g.dxi[0];
                    vn = vq[q] - (pn - pc) * (TS)g.dxi[1];
                    wn = (TS)((double)wq[q] - mean[k]) - (pt - pc) * (TS)g.dxi[2];
                }
                u[o] = un;
                v[o] = vn;
                w[o] = wn;
                if (cfl) {
                    const double c = dt * (fabs((double)un) * g.dxi[0] + fabs((double)vn) * g.dxi[1] +
                                           fabs((double)wn) * g.dxi[2]);
                    cm = cfl_fold(cm, c);
                }
                pc = pt;
            }
        }
        // z ghost levels from the corrected boundary levels this thread wrote:
        // the first chunk holds levels gw..2gw-1 and the last nz..nz+gw-1
        // (each chunk has at least gw levels unless it is the whole column)
        if (k0 == g.gw) {
            for (int m = 0; m < g.gw; ++m) {
                u[gidx(g, i, j, g.gw - 1 - m)] = u[gidx(g, i, j, g.gw + m)];
                v[gidx(g, i, j, g.gw - 1 - m)] = v[gidx(g, i, j, g.gw + m)];
            }
            w[gidx(g, i, j, g.gw - 1)] = TS(0);
            for (int m = 1; m < g.gw; ++m) w[gidx(g, i, j, g.gw - 1 - m)] = TS(-1.0) * w[gidx(g, i, j, g.gw - 1 + m)];
        }
        if (k1 == g.nz + g.gw) {
            for (int m = 0; m < g.gw; ++m) {
                u[gidx(g, i, j, g.nzg - g.gw + m)] = u[gidx(g, i, j, g.nzg - g.gw - 1 - m)];
                v[gidx(g, i, j, g.nzg - g.gw + m)] = v[gidx(g, i, j, g.nzg - g.gw - 1 - m)];
            }
            w[gidx(g, i, j, g.nzg - g.gw)] = TS(0);
            for (int m = 1; m < g.gw; ++m)
                w[gidx(g, i, j, g.nzg - g.gw + m)] = TS(-1.0) * w[gidx(g, i, j, g.nzg - g.gw - m)];
        }
    }
    if (cfl) cfl_block_max(cm, cfl);
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
    if (logx < 3 || logy < 3 || g.nx > 4096 || g.ny > 4096) {
        pycles_set_error("pressure: nx, ny must be powers of two in [8, 4096] (got %d, %d)", g.nx, g.ny);
        return PYCLES_ERR_UNSUPPORTED;
    }
    if (g.nz < 2 || g.nz > PRESSURE_MAX_NZ) {
        pycles_set_error("pressure: nz must be in [2, %d] (got %d)", PRESSURE_MAX_NZ, g.nz);
        return PYCLES_ERR_UNSUPPORTED;
    }
    pycles_pressure_plan *P = new pycles_pressure_plan();
    memset(P, 0, sizeof(*P));
    P->g = g;
    P->precision = precision;
    P->nxh = g.nx / 2 + 1;
    P->nxp = (P->nxh + 15) & ~15;
    P->logx = logx;
    P->logy = logy;
    const size_t cs = precision == PYCLES_F64 ? 16 : 8;
    const size_t es = cs / 2;
    // y transform: gx kx columns per block (16 makes 128-byte row segments in
    // FP32), at most 1024 threads and 64 KB
    int gx = 16;
    while (gx > 1 && ((size_t)gx * g.ny / 8 > 1024 || (size_t)gx * g.ny * cs > 64 * 1024)) gx >>= 1;
    P->gx = gx;
    P->loggx = ilog2(gx);
    const long long nspec = (long long)g.nz * g.ny * P->nxp;
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
    // twiddles exp(-2 pi i m / N), m < N, interleaved (re, im)
    std::vector<double> twx(2 * g.nx), twy(2 * g.ny);
    for (int m = 0; m < g.nx; ++m) {
        const double ang = -2.0 * PC_PI * (double)m / (double)g.nx;
        twx[2 * m] = cos(ang);
        twx[2 * m + 1] = sin(ang);
    }
    for (int m = 0; m < g.ny; ++m) {
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
    {
        const int seg = precision == PYCLES_F64 ? thomas_seg<double>() : thomas_seg<float>();
        ALLOC(P->ck, 3 * sizeof(double) * (size_t)(g.nz / seg + 1) * g.ny * P->nxp);
    }
    ALLOC(P->twx, (size_t)g.nx * cs);
    ALLOC(P->twy, (size_t)g.ny * cs);
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
        e = hipMemcpy(P->twx, twx.data(), (size_t)g.nx * 16, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(P->twy, twy.data(), (size_t)g.ny * 16, hipMemcpyHostToDevice);
    } else {
        std::vector<float> fx(2 * g.nx), fy(2 * g.ny);
        for (int m = 0; m < 2 * g.nx; ++m) fx[m] = (float)twx[m];
        for (int m = 0; m < 2 * g.ny; ++m) fy[m] = (float)twy[m];
        e = hipMemcpy(P->twx, fx.data(), (size_t)g.nx * 8, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(P->twy, fy.data(), (size_t)g.ny * 8, hipMemcpyHostToDevice);
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
    void *ptrs[] = {P->spec, P->ck, P->twx, P->twy, P->kx2, P->ky2, P->rho_h,
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
    // x transforms: S complex sequences (2S rows) per block, S nx/8 threads
    int logS = 0;
    while ((g.nx << (logS + 1)) <= 2048 && (2 << (logS + 1)) <= g.ny) ++logS;
    const int S = 1 << logS;
    const size_t rowsmem = sizeof(C2<TS>) * (size_t)g.nx * S;
    const int bx = S * g.nx / 8;
    hipLaunchKernelGGL(div_fftx_kernel<TS>, dim3(g.ny / (2 * S), g.nz), dim3(bx), rowsmem, st, g, (const TS *)u,
                       (const TS *)v, (const TS *)w, ref->alpha0, ref->alpha0_half, P->mean, (TS *)div,
                       (C2<TS> *)P->spec, (const C2<TS> *)P->twx, P->logx, P->nxh, P->nxp, logS);
    PYCLES_CHECK_LAUNCH();
    const size_t ysmem = sizeof(C2<TS>) * (size_t)g.ny * P->gx;
    const dim3 ygrid((P->nxh + P->gx - 1) / P->gx, g.nz);
    const int by = P->gx * g.ny / 8;
    hipLaunchKernelGGL(ffty_kernel<TS>, ygrid, dim3(by), ysmem, st, g, (C2<TS> *)P->spec,
                       (const C2<TS> *)P->twy, P->logy, P->nxh, P->nxp, P->gx, P->loggx, 0);
    PYCLES_CHECK_LAUNCH();
    const long long ncol = (long long)g.ny * P->nxp;
    hipLaunchKernelGGL((thomas_kernel<TS, thomas_seg<TS>()>), dim3((unsigned)((ncol + 255) / 256)), dim3(256),
                       4 * sizeof(double) * (size_t)g.nz, st, g, (C2<TS> *)P->spec, P->ck, P->kx2, P->ky2,
                       P->rho_h, P->dd, P->a, P->c, P->nxh, P->nxp);
    PYCLES_CHECK_LAUNCH();
    hipLaunchKernelGGL(ffty_kernel<TS>, ygrid, dim3(by), ysmem, st, g, (C2<TS> *)P->spec,
                       (const C2<TS> *)P->twy, P->logy, P->nxh, P->nxp, P->gx, P->loggx, 1);
    PYCLES_CHECK_LAUNCH();
    hipLaunchKernelGGL(ifftx_kernel<TS>, dim3(g.ny / (2 * S), g.nz), dim3(bx), rowsmem, st, g,
                       (const C2<TS> *)P->spec, (TS *)p, (const C2<TS> *)P->twx, P->logx, P->nxh, P->nxp, logS);
    PYCLES_CHECK_LAUNCH();
    if (cfl_out) hipLaunchKernelGGL(cfl_zero_kernel, dim3(1), dim3(1), 0, st, P->cfl);
    // z chunks so the grid holds ~16k waves
    const ZSplit zs = correct_zsplit(g.plane, g.nz, g.gw);
    const int nch = zs.nch, kchunk = zs.len;
    hipLaunchKernelGGL(correct_kernel<TS>, dim3((g.nx + 255) / 256, g.ny, nch), dim3(256), 0, st, g, (TS *)u,
                       (TS *)v, (TS *)w, (const TS *)p, P->mean, dt, cfl_out ? P->cfl : nullptr, kchunk);
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
