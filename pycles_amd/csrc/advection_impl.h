// Advection kernels: WENO5/7/9 flux + rho0-weighted flux divergence for the
// scalars and the nine momentum (advected, advecting) pairs, fused per
// direction.  Included by advection_f64.hip and advection_f32.hip, each of
// which instantiates one precision policy with its own floating-point
// contraction setting.
//
// Reference semantics (all cited files under the reference tree):
//   scalar flux      scalar_advection.h:319-578, rho0_half in x/y, rho0 in z
//   momentum flux    momentum_advection_weno.h:113-581: WENO of the advected
//                    component along the advecting direction, centred
//                    interpolation (order-1) of the advecting component along
//                    the advected direction; rho0_half for (u|v by u|v),
//                    rho0_half[k+1] for (w by w), rho0[k] for mixed pairs
//   divergence       flux_divergence.h:4-80: tend -= alpha*(F[c]-F[c-1])*dxi,
//                    alpha0_half for s, qt, u, v and alpha0 for w
// A face flux F[c] sits between cells c and c+1.  Each direction's pass
// computes every face flux exactly once: the y and z passes march along the
// line and keep the previous face flux in registers, the x pass holds a whole
// periodic row in LDS.
#pragma once
#include "common.h"

namespace adv {

// Per-variable slot numbering inside the fused kernels.
enum { VS0 = 0, VS1 = 1, VU = 2, VV = 3, VW = 4 };

struct Fields {
    const void *q[5];  // s0, s1, u, v, w
    void *t[5];        // tendencies
    const double *base[2];
};

struct Prof {
    const double *rho0, *rho0_half, *alpha0, *alpha0_half;
};

// -------------------------------------------------------------------------
// FP64 policy: storage and arithmetic in double, reference operation order.
struct PolicyF64 {
    using TS = double;
    using TC = double;
    using SFlux = double;
    static constexpr bool kBase = false;
    static __device__ __forceinline__ TC load(const TS *p, long long i) { return p[i]; }
    template <int O>
    static __device__ __forceinline__ SFlux scal_flux(const TC *win, const double *, TC U, double rho) {
        return upwind_flux<O, double>(win, U, rho);
    }
    template <int O>
    static __device__ __forceinline__ TC mom_flux(const TC *win, TC U, double rho) {
        return upwind_flux<O, double>(win, U, rho);
    }
    // tend -= alpha*(Fc - Fp)*dxi, evaluated as the reference does.
    static __device__ __forceinline__ TC scal_div(SFlux Fc, SFlux Fp, double alpha, double dxi) {
        return alpha * (Fc - Fp) * dxi;
    }
    static __device__ __forceinline__ TC mom_div(TC Fc, TC Fp, double alpha, double dxi) {
        return alpha * (Fc - Fp) * dxi;
    }
    static __device__ __forceinline__ void acc(TS *p, long long i, TC d) { p[i] -= d; }
};

// -------------------------------------------------------------------------
// FP32 policy.  Scalars are stored as deviations phi' = phi - base(k) from an
// FP64 level profile.  The reconstruction is evaluated on offsets from the
// upwind centre cell, d_m = phi_m - phi_ctr (exact for neighbouring values;
// WENO is invariant to adding a constant), so the large constant part never
// enters FP32 arithmetic.  The face flux is carried as a pair: the small part
// rho*U*(phi'_ctr + R(d)) in FP32 and the base part rho*U*base_ctr in FP64,
// and the flux difference of the base part is formed in FP64 before the
// tendency is rounded to FP32 (compensated flux divergence).
struct SF32 {
    float f;
    double b;
};
struct PolicyF32 {
    using TS = float;
    using TC = float;
    using SFlux = SF32;
    static constexpr bool kBase = true;
    static __device__ __forceinline__ TC load(const TS *p, long long i) { return p[i]; }
    template <int O>
    static __device__ __forceinline__ SFlux scal_flux(const TC *win, const double *bwin, TC U, double rho) {
        constexpr int R = Interp<O, float>::R;
        const bool pos = U >= 0.0f;
        const int ctr = pos ? R - 1 : R;
        const float pc = win[ctr];
        const double bc = bwin[ctr];
        float q[2 * R - 1];
#pragma unroll
        for (int m = 0; m < 2 * R - 1; ++m) {
            const int src = pos ? m : 2 * R - 1 - m;
            q[m] = (win[src] - pc) + (float)(bwin[src] - bc);
        }
        const float r = pc + Interp<O, float>::weno(q);
        SF32 F;
        F.f = (U * r) * (float)rho;
        F.b = ((double)U * bc) * rho;
        return F;
    }
    template <int O>
    static __device__ __forceinline__ TC mom_flux(const TC *win, TC U, double rho) {
        constexpr int R = Interp<O, float>::R;
        const bool pos = U >= 0.0f;
        const float pc = win[pos ? R - 1 : R];
        float q[2 * R - 1];
#pragma unroll
        for (int m = 0; m < 2 * R - 1; ++m) q[m] = (pos ? win[m] : win[2 * R - 1 - m]) - pc;
        const float r = pc + Interp<O, float>::weno(q);
        return (U * r) * (float)rho;
    }
    static __device__ __forceinline__ TC scal_div(SFlux Fc, SFlux Fp, double alpha, double dxi) {
        const double big = Fc.b - Fp.b;
        const float small = Fc.f - Fp.f;
        return (float)((big + (double)small) * (alpha * dxi));
    }
    static __device__ __forceinline__ TC mom_div(TC Fc, TC Fp, double alpha, double dxi) {
        return (float)alpha * (Fc - Fp) * (float)dxi;
    }
    static __device__ __forceinline__ void acc(TS *p, long long i, TC d) { p[i] -= d; }
};

// Load the centred advecting-velocity stencil (order MO) along a periodic
// horizontal direction from global memory.
template <class P, int MO>
__device__ __forceinline__ typename P::TC centred_x(const typename P::TS *f, const Geo &g, int i, int j, int k) {
    using TC = typename P::TC;
    constexpr int RM = Interp<MO, TC>::R;
    TC q[2 * (RM - 1)];
#pragma unroll
    for (int m = 0; m < 2 * (RM - 1); ++m) q[m] = P::load(f, gidx(g, wrapi(i + m - (RM - 2), g.nx), j, k));
    return Interp<MO, TC>::centred(q);
}
template <class P, int MO>
__device__ __forceinline__ typename P::TC centred_y(const typename P::TS *f, const Geo &g, int i, int j, int k) {
    using TC = typename P::TC;
    constexpr int RM = Interp<MO, TC>::R;
    TC q[2 * (RM - 1)];
#pragma unroll
    for (int m = 0; m < 2 * (RM - 1); ++m) q[m] = P::load(f, gidx(g, i, wrapi(j + m - (RM - 2), g.ny), k));
    return Interp<MO, TC>::centred(q);
}
template <class P, int MO>
__device__ __forceinline__ typename P::TC centred_z(const typename P::TS *f, const Geo &g, int i, int j, int k) {
    using TC = typename P::TC;
    constexpr int RM = Interp<MO, TC>::R;
    TC q[2 * (RM - 1)];
#pragma unroll
    for (int m = 0; m < 2 * (RM - 1); ++m) q[m] = P::load(f, gidx(g, i, j, k + m - (RM - 2)));
    return Interp<MO, TC>::centred(q);
}

// -------------------------------------------------------------------------
// z pass: one thread per (i, j) column, marching the faces c = k0-1 .. k1-1.
// Block (64, 4); grid (ceil(nx/64), ceil(ny/4), nchunks).
template <class P, int SO, int MO, int NSC, bool MOM>
__global__ __launch_bounds__(256) void adv_z(Geo g, Prof pr, Fields fl, int kchunk) {
    using TS = typename P::TS;
    using TC = typename P::TC;
    using SFlux = typename P::SFlux;
    constexpr int RS = Interp<SO, TC>::R;
    constexpr int RM = Interp<MO, TC>::R;
    const int i = blockIdx.x * 64 + threadIdx.x;
    const int j = blockIdx.y * 4 + threadIdx.y;
    if (i >= g.nx || j >= g.ny) return;  // no barriers in this kernel
    const int k0 = g.gw + blockIdx.z * kchunk;
    const int k1 = min(k0 + kchunk, g.nz + g.gw);
    const long long col = (long long)j * g.nx + i;
    const TS *ps[2] = {(const TS *)fl.q[0], (const TS *)fl.q[1]};
    const TS *pu = (const TS *)fl.q[VU];
    const TS *pv = (const TS *)fl.q[VV];
    const TS *pw = (const TS *)fl.q[VW];

    TC ws[NSC > 0 ? NSC : 1][2 * RS];
    double bw[NSC > 0 ? NSC : 1][2 * RS];
    TC wsu[2 * RM], wsv[2 * RM], wsw[2 * RM];  // momentum windows
    TC wwS[2 * RS];                               // w at scalar-window levels (only [RS-1] used)
    (void)wwS;

    int c = k0 - 1;
    // initial windows: levels c-R+1 .. c+R
#pragma unroll
    for (int sc = 0; sc < NSC; ++sc)
#pragma unroll
        for (int m = 0; m < 2 * RS; ++m) {
            const int lev = c - (RS - 1) + m;
            ws[sc][m] = P::load(ps[sc], lev * g.plane + col);
            bw[sc][m] = P::kBase ? (fl.base[sc] ? fl.base[sc][lev] : 0.0) : 0.0;
        }
    if (MOM) {
#pragma unroll
        for (int m = 0; m < 2 * RM; ++m) {
            const long long o = (long long)(c - (RM - 1) + m) * g.plane + col;
            wsu[m] = P::load(pu, o);
            wsv[m] = P::load(pv, o);
            wsw[m] = P::load(pw, o);
        }
    }
    SFlux Fs_prev[NSC > 0 ? NSC : 1];
    TC Fu_prev = 0, Fv_prev = 0, Fw_prev = 0;
    const double dzi = g.dxi[2];

    for (; c < k1; ++c) {
        const double rho0c = pr.rho0[c];
        const TC W = P::load(pw, c * g.plane + col);
        SFlux Fs[NSC > 0 ? NSC : 1];
#pragma unroll
        for (int sc = 0; sc < NSC; ++sc) Fs[sc] = P::template scal_flux<SO>(ws[sc], bw[sc], W, rho0c);
        TC Fu = 0, Fv = 0, Fw = 0;
        if (MOM) {
            // u advected by w: U = centred(w along x), rho0[k]
            const TC Uu = centred_x<P, MO>(pw, g, i, j, c);
            Fu = P::template mom_flux<MO>(wsu, Uu, rho0c);
            // v advected by w: U = centred(w along y), rho0[k]
            const TC Uv = centred_y<P, MO>(pw, g, i, j, c);
            Fv = P::template mom_flux<MO>(wsv, Uv, rho0c);
            // w advected by w: U = centred(w along z) from the window, rho0_half[k+1]
            TC q[2 * (RM - 1)];
#pragma unroll
            for (int m = 0; m < 2 * (RM - 1); ++m) q[m] = wsw[m + 1];
            const TC Uw = Interp<MO, TC>::centred(q);
            Fw = P::template mom_flux<MO>(wsw, Uw, pr.rho0_half[c + 1]);
        }
        if (c >= k0) {
            const long long o = c * g.plane + col;
            const double ah = pr.alpha0_half[c];
#pragma unroll
            for (int sc = 0; sc < NSC; ++sc) P::acc((TS *)fl.t[sc], o, P::scal_div(Fs[sc], Fs_prev[sc], ah, dzi));
            if (MOM) {
                P::acc((TS *)fl.t[VU], o, P::mom_div(Fu, Fu_prev, ah, dzi));
                P::acc((TS *)fl.t[VV], o, P::mom_div(Fv, Fv_prev, ah, dzi));
                P::acc((TS *)fl.t[VW], o, P::mom_div(Fw, Fw_prev, pr.alpha0[c], dzi));
            }
        }
#pragma unroll
        for (int sc = 0; sc < NSC; ++sc) Fs_prev[sc] = Fs[sc];
        Fu_prev = Fu;
        Fv_prev = Fv;
        Fw_prev = Fw;
        if (c + 1 < k1) {
            // shift windows and load the next level
#pragma unroll
            for (int sc = 0; sc < NSC; ++sc) {
#pragma unroll
                for (int m = 0; m < 2 * RS - 1; ++m) {
                    ws[sc][m] = ws[sc][m + 1];
                    bw[sc][m] = bw[sc][m + 1];
                }
                const int lev = c + 1 + RS;
                ws[sc][2 * RS - 1] = P::load(ps[sc], lev * g.plane + col);
                bw[sc][2 * RS - 1] = P::kBase ? (fl.base[sc] ? fl.base[sc][lev] : 0.0) : 0.0;
            }
            if (MOM) {
#pragma unroll
                for (int m = 0; m < 2 * RM - 1; ++m) {
                    wsu[m] = wsu[m + 1];
                    wsv[m] = wsv[m + 1];
                    wsw[m] = wsw[m + 1];
                }
                const long long o = (long long)(c + 1 + RM) * g.plane + col;
                wsu[2 * RM - 1] = P::load(pu, o);
                wsv[2 * RM - 1] = P::load(pv, o);
                wsw[2 * RM - 1] = P::load(pw, o);
            }
        }
    }
}

// -------------------------------------------------------------------------
// y pass: one thread per (i, k) with k interior, marching faces c = j0-1 .. j1-1
// (periodic).  Block (64, 4); grid (ceil(nx/64), ceil(nz/4), nchunks).
template <class P, int SO, int MO, int NSC, bool MOM>
__global__ __launch_bounds__(256) void adv_y(Geo g, Prof pr, Fields fl, int jchunk) {
    using TS = typename P::TS;
    using TC = typename P::TC;
    using SFlux = typename P::SFlux;
    constexpr int RS = Interp<SO, TC>::R;
    constexpr int RM = Interp<MO, TC>::R;
    const int i = blockIdx.x * 64 + threadIdx.x;
    const int kk = blockIdx.y * 4 + threadIdx.y;
    if (i >= g.nx || kk >= g.nz) return;
    const int k = kk + g.gw;
    const int j0 = blockIdx.z * jchunk;
    const int j1 = min(j0 + jchunk, g.ny);
    const long long lev = (long long)k * g.plane;
    const TS *ps[2] = {(const TS *)fl.q[0], (const TS *)fl.q[1]};
    const TS *pu = (const TS *)fl.q[VU];
    const TS *pv = (const TS *)fl.q[VV];
    const TS *pw = (const TS *)fl.q[VW];
    const double rh = pr.rho0_half[k];
    const double r0 = pr.rho0[k];
    const double ah = pr.alpha0_half[k];
    const double a0 = pr.alpha0[k];
    const double dyi = g.dxi[1];

    TC ws[NSC > 0 ? NSC : 1][2 * RS];
    double bw[NSC > 0 ? NSC : 1][2 * RS];
    TC wsu[2 * RM], wsv[2 * RM], wsw[2 * RM];
    int c = j0 - 1;
#pragma unroll
    for (int sc = 0; sc < NSC; ++sc) {
        const double b = P::kBase ? (fl.base[sc] ? fl.base[sc][k] : 0.0) : 0.0;
#pragma unroll
        for (int m = 0; m < 2 * RS; ++m) {
            ws[sc][m] = P::load(ps[sc], lev + (long long)wrapi(c - (RS - 1) + m, g.ny) * g.nx + i);
            bw[sc][m] = b;
        }
    }
    if (MOM) {
#pragma unroll
        for (int m = 0; m < 2 * RM; ++m) {
            const long long o = lev + (long long)wrapi(c - (RM - 1) + m, g.ny) * g.nx + i;
            wsu[m] = P::load(pu, o);
            wsv[m] = P::load(pv, o);
            wsw[m] = P::load(pw, o);
        }
    }
    SFlux Fs_prev[NSC > 0 ? NSC : 1];
    TC Fu_prev = 0, Fv_prev = 0, Fw_prev = 0;
    for (; c < j1; ++c) {
        const int cw = wrapi(c, g.ny);
        SFlux Fs[NSC > 0 ? NSC : 1];
        TC V;
        if (MOM) V = wsv[RM - 1];
        else V = P::load(pv, lev + (long long)cw * g.nx + i);
#pragma unroll
        for (int sc = 0; sc < NSC; ++sc) Fs[sc] = P::template scal_flux<SO>(ws[sc], bw[sc], V, rh);
        TC Fu = 0, Fv = 0, Fw = 0;
        if (MOM) {
            const TC Uu = centred_x<P, MO>(pv, g, i, cw, k);  // v along x
            Fu = P::template mom_flux<MO>(wsu, Uu, rh);
            TC q[2 * (RM - 1)];
#pragma unroll
            for (int m = 0; m < 2 * (RM - 1); ++m) q[m] = wsv[m + 1];  // v along y
            const TC Uv = Interp<MO, TC>::centred(q);
            Fv = P::template mom_flux<MO>(wsv, Uv, rh);
            const TC Uw = centred_z<P, MO>(pv, g, i, cw, k);  // v along z
            Fw = P::template mom_flux<MO>(wsw, Uw, r0);
        }
        if (c >= j0) {
            const long long o = lev + (long long)c * g.nx + i;
#pragma unroll
            for (int sc = 0; sc < NSC; ++sc) P::acc((TS *)fl.t[sc], o, P::scal_div(Fs[sc], Fs_prev[sc], ah, dyi));
            if (MOM) {
                P::acc((TS *)fl.t[VU], o, P::mom_div(Fu, Fu_prev, ah, dyi));
                P::acc((TS *)fl.t[VV], o, P::mom_div(Fv, Fv_prev, ah, dyi));
                P::acc((TS *)fl.t[VW], o, P::mom_div(Fw, Fw_prev, a0, dyi));
            }
        }
#pragma unroll
        for (int sc = 0; sc < NSC; ++sc) Fs_prev[sc] = Fs[sc];
        Fu_prev = Fu;
        Fv_prev = Fv;
        Fw_prev = Fw;
        if (c + 1 < j1) {
#pragma unroll
            for (int sc = 0; sc < NSC; ++sc) {
#pragma unroll
                for (int m = 0; m < 2 * RS - 1; ++m) ws[sc][m] = ws[sc][m + 1];
                ws[sc][2 * RS - 1] = P::load(ps[sc], lev + (long long)wrapi(c + 1 + RS, g.ny) * g.nx + i);
            }
            if (MOM) {
#pragma unroll
                for (int m = 0; m < 2 * RM - 1; ++m) {
                    wsu[m] = wsu[m + 1];
                    wsv[m] = wsv[m + 1];
                    wsw[m] = wsw[m + 1];
                }
                const long long o = lev + (long long)wrapi(c + 1 + RM, g.ny) * g.nx + i;
                wsu[2 * RM - 1] = P::load(pu, o);
                wsv[2 * RM - 1] = P::load(pv, o);
                wsw[2 * RM - 1] = P::load(pw, o);
            }
        }
    }
}

// -------------------------------------------------------------------------
// x pass: one block per (j, k) row.  The row of every field (plus a periodic
// halo of H cells) is staged in LDS; for each variable the block computes the
// nx face fluxes into LDS, then updates the nx tendencies.
template <class P, int SO, int MO, int NSC, bool MOM>
__global__ __launch_bounds__(256) void adv_x(Geo g, Prof pr, Fields fl) {
    using TS = typename P::TS;
    using TC = typename P::TC;
    using SFlux = typename P::SFlux;
    constexpr int RS = Interp<SO, TC>::R;
    constexpr int RM = Interp<MO, TC>::R;
    constexpr int H = (MOM && RM > RS) ? RM : RS;
    constexpr int NV = NSC + (MOM ? 3 : 1);  // u row always staged (advecting velocity)
    extern __shared__ __align__(16) unsigned char smem_raw[];
    TC *rows = reinterpret_cast<TC *>(smem_raw);
    const int nx = g.nx;
    const int W = nx + 2 * H;
    SFlux *sflux = reinterpret_cast<SFlux *>(rows + NV * W);
    TC *mflux = reinterpret_cast<TC *>(sflux + nx);
    const int j = blockIdx.x;
    const int k = blockIdx.y + g.gw;
    const long long row = ((long long)k * g.ny + j) * nx;
    const int tid = threadIdx.x;
    const int B = blockDim.x;
    // slots: 0..NSC-1 scalars, NSC = u, NSC+1 = v, NSC+2 = w
    const TS *src[5];
#pragma unroll
    for (int sc = 0; sc < NSC; ++sc) src[sc] = (const TS *)fl.q[sc];
    src[NSC] = (const TS *)fl.q[VU];
    if (MOM) {
        src[NSC + 1] = (const TS *)fl.q[VV];
        src[NSC + 2] = (const TS *)fl.q[VW];
    }
    for (int v = 0; v < NV; ++v)
        for (int t = tid; t < W; t += B) rows[v * W + t] = P::load(src[v], row + wrapi(t - H, nx));
    __syncthreads();
    const double rh = pr.rho0_half[k];
    const double r0 = pr.rho0[k];
    const double ah = pr.alpha0_half[k];
    const double a0 = pr.alpha0[k];
    const double dxi = g.dxi[0];
    const TC *urow = rows + NSC * W + H;  // urow[i] = u at cell i (wrapped halo)

    // scalars
#pragma unroll
    for (int sc = 0; sc < NSC; ++sc) {
        const double b = P::kBase ? (fl.base[sc] ? fl.base[sc][k] : 0.0) : 0.0;
        double bw[2 * RS];
#pragma unroll
        for (int m = 0; m < 2 * RS; ++m) bw[m] = b;
        const TC *srow = rows + sc * W + H;
        for (int c = tid; c < nx; c += B) {
            TC win[2 * RS];
#pragma unroll
            for (int m = 0; m < 2 * RS; ++m) win[m] = srow[c - (RS - 1) + m];
            sflux[c] = P::template scal_flux<SO>(win, bw, urow[c], rh);
        }
        __syncthreads();
        for (int c = tid; c < nx; c += B) {
            const int cm = c == 0 ? nx - 1 : c - 1;
            P::acc((TS *)fl.t[sc], row + c, P::scal_div(sflux[c], sflux[cm], ah, dxi));
        }
        __syncthreads();
    }
    if (MOM) {
        const TS *pu = (const TS *)fl.q[VU];
        for (int a = 0; a < 3; ++a) {  // advected component u, v, w
            const TC *prow = rows + (NSC + a) * W + H;
            const double rho = (a == 2) ? r0 : rh;
            for (int c = tid; c < nx; c += B) {
                TC win[2 * RM];
#pragma unroll
                for (int m = 0; m < 2 * RM; ++m) win[m] = prow[c - (RM - 1) + m];
                TC U;
                if (a == 0) {
                    TC q[2 * (RM - 1)];
#pragma unroll
                    for (int m = 0; m < 2 * (RM - 1); ++m) q[m] = urow[c + m - (RM - 2)];
                    U = Interp<MO, TC>::centred(q);
                } else if (a == 1) {
                    U = centred_y<P, MO>(pu, g, c, j, k);
                } else {
                    U = centred_z<P, MO>(pu, g, c, j, k);
                }
                mflux[c] = P::template mom_flux<MO>(win, U, rho);
            }
            __syncthreads();
            const double al = (a == 2) ? a0 : ah;
            for (int c = tid; c < nx; c += B) {
                const int cm = c == 0 ? nx - 1 : c - 1;
                P::acc((TS *)fl.t[VU + a], row + c, P::mom_div(mflux[c], mflux[cm], al, dxi));
            }
            __syncthreads();
        }
    }
}

template <class P, int SO, int MO, int NSC, bool MOM>
size_t x_smem_bytes(int nx) {
    using TC = typename P::TC;
    constexpr int RS = Interp<SO, TC>::R;
    constexpr int RM = Interp<MO, TC>::R;
    constexpr int H = (MOM && RM > RS) ? RM : RS;
    constexpr int NV = NSC + (MOM ? 3 : 1);
    return sizeof(TC) * (size_t)NV * (nx + 2 * H) + sizeof(typename P::SFlux) * nx + sizeof(TC) * nx + 64;
}

// Launch the three passes in the reference's per-variable order (x, y, z).
template <class P, int SO, int MO, int NSC, bool MOM>
int launch(const Geo &g, const Prof &pr, const Fields &fl, hipStream_t st) {
    const size_t smem = x_smem_bytes<P, SO, MO, NSC, MOM>(g.nx);
    if (smem > 160 * 1024) {
        pycles_set_error("advection: row of %d cells does not fit in LDS", g.nx);
        return PYCLES_ERR_UNSUPPORTED;
    }
    if (smem > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void *)adv_x<P, SO, MO, NSC, MOM>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
        if (e != hipSuccess) {
            pycles_set_error("advection: LDS request %zu bytes refused: %s", smem, hipGetErrorString(e));
            return PYCLES_ERR_UNSUPPORTED;
        }
    }
    int bx = g.nx >= 256 ? 256 : ((g.nx + 63) / 64) * 64;
    hipLaunchKernelGGL((adv_x<P, SO, MO, NSC, MOM>), dim3(g.ny, g.nz), dim3(bx), smem, st, g, pr, fl);
    PYCLES_CHECK_LAUNCH();
    // y: split the march into chunks so the grid holds enough waves.
    const long long ycols = (long long)((g.nx + 63) / 64) * ((g.nz + 3) / 4);
    int ych = 1;
    while (ycols * ych < 2048 && g.ny / (ych * 2) >= 32) ych *= 2;
    const int jchunk = (g.ny + ych - 1) / ych;
    hipLaunchKernelGGL((adv_y<P, SO, MO, NSC, MOM>), dim3((g.nx + 63) / 64, (g.nz + 3) / 4, ych), dim3(64, 4),
                       0, st, g, pr, fl, jchunk);
    PYCLES_CHECK_LAUNCH();
    const long long zcols = (long long)((g.nx + 63) / 64) * ((g.ny + 3) / 4);
    int zch = 1;
    while (zcols * zch < 2048 && g.nz / (zch * 2) >= 32) zch *= 2;
    const int kchunk = (g.nz + zch - 1) / zch;
    hipLaunchKernelGGL((adv_z<P, SO, MO, NSC, MOM>), dim3((g.nx + 63) / 64, (g.ny + 3) / 4, zch), dim3(64, 4),
                       0, st, g, pr, fl, kchunk);
    PYCLES_CHECK_LAUNCH();
    return PYCLES_OK;
}

// Runtime dispatch over the instantiated (order, scalar count, momentum) set.
template <class P>
int dispatch(const Geo &g, const Prof &pr, const Fields &fl, int nsc, bool mom, int so, int mo,
             hipStream_t st) {
#define ADV_CASE(SO_, MO_, NSC_, MOM_) \
    if (so == SO_ && mo == MO_ && nsc == NSC_ && mom == MOM_) return launch<P, SO_, MO_, NSC_, MOM_>(g, pr, fl, st);
    if (!mom) mo = so;
    if (nsc == 0) so = mo;
    ADV_CASE(5, 5, 0, true)
    ADV_CASE(7, 7, 0, true)
    ADV_CASE(9, 9, 0, true)
    ADV_CASE(5, 5, 1, false)
    ADV_CASE(5, 5, 2, false)
    ADV_CASE(7, 7, 1, false)
    ADV_CASE(7, 7, 2, false)
    ADV_CASE(9, 9, 1, false)
    ADV_CASE(9, 9, 2, false)
    ADV_CASE(5, 5, 1, true)
    ADV_CASE(5, 5, 2, true)
    ADV_CASE(7, 7, 1, true)
    ADV_CASE(7, 7, 2, true)
    ADV_CASE(9, 9, 1, true)
    ADV_CASE(9, 9, 2, true)
    ADV_CASE(9, 5, 1, true)
    ADV_CASE(9, 5, 2, true)
#undef ADV_CASE
    pycles_set_error("advection: unsupported combination scalar order %d, momentum order %d, %d scalars", so, mo, nsc);
    return PYCLES_ERR_UNSUPPORTED;
}

}  // namespace adv
