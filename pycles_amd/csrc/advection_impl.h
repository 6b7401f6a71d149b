// Advection kernels: WENO5/7/9 flux + rho0-weighted flux divergence for the
// scalars and the nine momentum (advected, advecting) pairs, fused per
// direction.  Included by advection_f64.hip and advection_f32.hip, each of
// which instantiates one precision policy with its own floating-point
// contraction setting (off for FP64, which reproduces the reference's
// rounding; on for FP32).
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
// periodic row in LDS.  Field offsets are 32-bit byte offsets from the field
// base (one field must stay below 4 GiB), so loads use a scalar base plus a
// 32-bit vector offset; wrapped horizontal neighbour offsets are computed
// once per thread, outside the march.
#pragma once
#include <utility>

#include "common.h"


namespace adv {

// Calls f(integral_constant<int, Q>) for Q = Q0 .. N-1 while it returns true;
// returns false as soon as one call does.
template <int Q, int N, class F>
__device__ __forceinline__ bool phases(F &f) {
    if constexpr (Q == N) {
        return true;
    } else {
        if (!f(std::integral_constant<int, Q>{})) return false;
        return phases<Q + 1, N>(f);
    }
}

enum { VS0 = 0, VS1 = 1, VU = 2, VV = 3, VW = 4 };

struct Fields {
    const void *q[5];  // s0, s1, u, v, w
    void *t[5];        // tendencies
    const double *base[2];
    int tzero;  // scalar, u, v tendencies are zero on entry (x pass writes them)
    int passes;  // bit 0: x, bit 1: y, bit 2: z
    int fuse_yz;  // run the y and z passes as one kernel in FP64 too (FP32: always, where the orders allow)
    // RK3 stage fused into the z pass (rk >= 0): new values to out, stage
    // copies in cp, z boundary fill (sym for scalars, u, v; asym for w)
    int rk;
    double dt;
    void *out[5];
    void *cp[5];
};

struct Prof {
    const double *rho0, *rho0_half, *alpha0, *alpha0_half;
};

template <typename T>
__device__ __forceinline__ T ld(const T *p, unsigned idx) {
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(p) + (size_t)(idx * (unsigned)sizeof(T)));
}
template <typename T>
__device__ __forceinline__ T *at(T *p, unsigned idx) {
    return reinterpret_cast<T *>(reinterpret_cast<char *>(p) + (size_t)(idx * (unsigned)sizeof(T)));
}

// Streamed accesses, non-temporal: the tendencies, stage copies and new
// values that a pass reads or writes once (the field windows, which
// neighbouring lanes and blocks re-read, stay ordinary loads).
template <typename T>
__device__ __forceinline__ T lds_(const T *p, unsigned idx) {
    return __builtin_nontemporal_load(at(const_cast<T *>(p), idx));
}
template <typename T>
__device__ __forceinline__ void sts_(T *p, unsigned idx, T v) {
    __builtin_nontemporal_store(v, at(p, idx));
}

// tend -= d, or tend = 0 - d when the tendency is known to be zero (same bits).
template <typename TS>
__device__ __forceinline__ void sub_into(TS *p, unsigned o, TS d, bool zero) {
    if (zero) *at(p, o) = TS(0) - d;
    else *at(p, o) -= d;
}

// -------------------------------------------------------------------------
// FP64 policy: storage and arithmetic in double, reference operation order.
struct PolicyF64 {
    using TS = double;
    using TC = double;
    using SFlux = double;
    static constexpr bool kBase = false;
    // the y and z passes stay separate unless asked (PYCLES_ADV_FUSE_YZ, for
    // tests): the fused kernel's double-buffered FP64 tiles (86 KB of LDS)
    // leave one block per CU
    static constexpr bool kFuseYZ = false;
    static constexpr int kYZWaves = 1;
    // waves per SIMD the z pass is held to (<= 168 VGPRs): its stage-2 form
    // otherwise takes 171 and runs 2 waves
    static constexpr int kZWaves = 3;
    template <int O, bool BV>
    static __device__ __forceinline__ SFlux scal_flux(const TC *win, const double *, TC U, double rho) {
        return upwind_flux<O, double>(win, U, rho);
    }
    template <int O>
    static __device__ __forceinline__ TC mom_flux(const TC *win, TC U, double rho) {
        return upwind_flux<O, double>(win, U, rho);
    }
    // tend -= alpha*(Fc - Fp)*dxi, evaluated as the reference does.
    template <bool BV>
    static __device__ __forceinline__ TC scal_div(SFlux Fc, SFlux Fp, double alpha, double dxi, double) {
        return alpha * (Fc - Fp) * dxi;
    }
    static __device__ __forceinline__ TC mom_div(TC Fc, TC Fp, double alpha, double dxi) {
        return alpha * (Fc - Fp) * dxi;
    }
    // x pass: the flux kept in LDS, and its divergence from two kept fluxes
    // and the face velocities
    static __device__ __forceinline__ TC x_keep(SFlux F) { return F; }
    static __device__ __forceinline__ TC x_scal_div(TC fc, TC fp, TC, TC, double alpha, double dxi, double) {
        return alpha * (fc - fp) * dxi;
    }
};

// -------------------------------------------------------------------------
// FP32 policy.  Scalars are stored as deviations phi' = phi - base(k) from an
// FP64 level profile.  The reconstruction is evaluated on offsets from the
// upwind centre cell, d_m = phi_m - phi_ctr (WENO is invariant to adding a
// constant, and neighbouring values subtract without rounding), so the large
// constant part never enters FP32 arithmetic.  The face flux is carried as a
// pair: the small part rho*U*(phi'_ctr + R(d)) in FP32 and the base part
// rho*U*base_ctr in FP64, and the base part's flux difference is formed in
// FP64 before the tendency is rounded to FP32 (compensated flux divergence).
// BV: the base varies along the line (z); in x and y it is one value bc, and
// the pair carries U itself: the base part's flux difference is then
// (U_c - U_{c-1}) * bc * rho, with the difference of two FP32 values exact
// in FP64.
struct SF32 {
    float f;
    double b;
};
struct PolicyF32 {
    using TS = float;
    using TC = float;
    using SFlux = SF32;
    static constexpr bool kBase = true;
    static constexpr bool kFuseYZ = true;  // y and z passes fused (WENO5)
    static constexpr int kYZWaves = 3;     // waves per SIMD of the fused pass (its LDS allows 3 blocks per CU)
    static constexpr int kZWaves = 1;      // the z pass's registers as the compiler allocates them
    template <int O, bool BV>
    static __device__ __forceinline__ SFlux scal_flux(const TC *win, const double *bwin, TC U, double rho) {
        constexpr int R = Interp<O, float>::R;
        const bool pos = U >= 0.0f;
        // selects of values with constant indices (an index select would
        // become a compare/select chain over the whole register window)
        const float pc = pos ? win[R - 1] : win[R];
        float q[2 * R - 1];
        double bc = 0.0;
        if (BV) {
            bc = pos ? bwin[R - 1] : bwin[R];
#pragma unroll
            for (int m = 0; m < 2 * R - 1; ++m) {
                const float wv = pos ? win[m] : win[2 * R - 1 - m];
                const double bv = pos ? bwin[m] : bwin[2 * R - 1 - m];
                q[m] = (wv - pc) + (float)(bv - bc);
            }
        } else {
#pragma unroll
            for (int m = 0; m < 2 * R - 1; ++m) q[m] = (pos ? win[m] : win[2 * R - 1 - m]) - pc;
        }
        const float r = pc + Interp<O, float>::weno_c0(q);
        SF32 F;
        F.f = (U * r) * (float)rho;
        F.b = BV ? ((double)U * bc) * rho : (double)U;
        return F;
    }
    template <int O>
    static __device__ __forceinline__ TC mom_flux(const TC *win, TC U, double rho) {
        constexpr int R = Interp<O, float>::R;
        const bool pos = U >= 0.0f;
        const float pc = pos ? win[R - 1] : win[R];
        float q[2 * R - 1];
#pragma unroll
        for (int m = 0; m < 2 * R - 1; ++m) q[m] = (pos ? win[m] : win[2 * R - 1 - m]) - pc;
        const float r = pc + Interp<O, float>::weno_c0(q);
        return (U * r) * (float)rho;
    }
    // brho = bc * rho for a line of constant base (BV false), unused otherwise
    template <bool BV>
    static __device__ __forceinline__ TC scal_div(SFlux Fc, SFlux Fp, double alpha, double dxi, double brho) {
        const float small = Fc.f - Fp.f;
        const double big = BV ? (Fc.b - Fp.b) : (Fc.b - Fp.b) * brho;
        return (float)((big + (double)small) * (alpha * dxi));
    }
    static __device__ __forceinline__ TC mom_div(TC Fc, TC Fp, double alpha, double dxi) {
        return (float)alpha * (Fc - Fp) * (float)dxi;
    }
    // x pass: only the FP32 part is kept; the base part is rebuilt from the
    // face velocities, which is scal_div<false> exactly
    static __device__ __forceinline__ TC x_keep(SFlux F) { return F.f; }
    static __device__ __forceinline__ TC x_scal_div(TC fc, TC fp, TC Uc, TC Up, double alpha, double dxi,
                                                    double brho) {
        SF32 c, m;
        c.f = fc;
        c.b = (double)Uc;
        m.f = fp;
        m.b = (double)Up;
        return scal_div<false>(c, m, alpha, dxi, brho);
    }
};

// Centred interpolation (order MO) from NC = 2(R-1) loads at the given offsets.
template <class P, int MO>
__device__ __forceinline__ typename P::TC centred_at(const typename P::TS *f, const unsigned *off) {
    using TC = typename P::TC;
    constexpr int NC = 2 * (Interp<MO, TC>::R - 1);
    TC q[NC];
#pragma unroll
    for (int m = 0; m < NC; ++m) q[m] = (TC)ld(f, off[m]);
    return Interp<MO, TC>::centred(q);
}

// -------------------------------------------------------------------------
// z pass: one thread per (i, j) column, marching the faces c = k0-1 .. k1-1.
// Block (64, 4); grid (ceil(nx/64), ceil(ny/4), nchunks).
// The RK3 stage of one cell from its final tendency, written to fl.out[v]
// with the z boundary fill the cell feeds (set_bcs, prognostic_variables.h:
// 205-270): sym mirrors the first/last gw levels; asym (w) zeroes the
// boundary level and mirrors with a sign change.
// RKM selects what the z pass stores: 0 the final tendencies, 1 RK3 stage 0
// (writes the stage copy, unless its pointer is NULL), 2 RK3 stage 1 or 2
// (reads it; cpv is the copy, read before the output is written, so the
// output may be the copy array itself).
template <typename TS, int RKM>
__device__ __forceinline__ void rk_store(const Fields &fl, int v, unsigned o, unsigned colo, unsigned plane, int c,
                                         int gw, int nzg, TS vold, TS tn, TS cpv, bool asym) {
    TS *out = (TS *)fl.out[v];
    const TS dt = (TS)fl.dt;
    TS vn;
    if (RKM == 1) {
        if (fl.cp[v]) sts_((TS *)fl.cp[v], o, vold);  // NULL: the caller keeps the input as the copy
        vn = rk_update<TS, 0>(vold, tn, TS(0), dt);
    } else if (fl.rk == 1) {
        vn = rk_update<TS, 1>(vold, tn, cpv, dt);
    } else {
        vn = rk_update<TS, 2>(vold, tn, cpv, dt);
    }
    sts_(out, o, vn);
    if (!asym) {
        if (c < 2 * gw) *at(out, (unsigned)(2 * gw - 1 - c) * plane + colo) = vn;
        if (c >= nzg - 2 * gw) *at(out, (unsigned)(2 * (nzg - gw) - 1 - c) * plane + colo) = vn;
    } else {
        if (c == gw) *at(out, (unsigned)(gw - 1) * plane + colo) = TS(0);
        if (c <= 2 * gw - 2) *at(out, (unsigned)(2 * gw - 2 - c) * plane + colo) = -vn;
        if (c == nzg - gw - 1) *at(out, (unsigned)(nzg - gw) * plane + colo) = TS(0);
        if (c >= nzg - 2 * gw + 1) *at(out, (unsigned)(2 * (nzg - gw) - c) * plane + colo) = -vn;
    }
}

// Software pipelined: at the top of each face the loads of the next window
// levels (clamped to the last ghost level) and of this face's centred w
// stencils (used after the scalar fluxes) are issued, and this cell's
// tendencies and stage copies together after the fluxes; all unconditional,
// so one counted wait covers each group.  The first face (k0 - 1, flux only)
// is peeled off the loop.  The velocity and scalar windows are register
// rings as in the y pass below; the scalars' FP64 base window, which is the
// same for every lane and lives in scalar registers, shifts.
template <class P, int SO, int MO, int NSC, bool MOM, int RKM>
__global__ __launch_bounds__(256, P::kZWaves) void adv_z(Geo g, Prof pr, Fields fl, int kchunk) {
    using TS = typename P::TS;
    using TC = typename P::TC;
    using SFlux = typename P::SFlux;
    constexpr int RS = Interp<SO, TC>::R;
    constexpr int RM = Interp<MO, TC>::R;
    constexpr int NC = 2 * (RM - 1);
    constexpr int NS = NSC > 0 ? NSC : 1;
    constexpr int WS = 2 * RS, WM = 2 * RM;
    constexpr int WMAX = (NSC > 0 && WS > WM) || !MOM ? WS : WM;
    constexpr bool RING = WMAX + 1 <= 8;
    constexpr int L = RING ? WMAX + 1 : WMAX;
    const uint3 tile = xcd_tile<1>();  // XCD bands of 4-row blocks in y
    const int i = tile.x * 64 + threadIdx.x;
    const int j = tile.y * 4 + threadIdx.y;
    if (i >= g.nx || j >= g.ny) return;  // no barriers in this kernel
    const int k0 = g.gw + tile.z * kchunk;
    const int k1 = min(k0 + kchunk, g.nz + g.gw);
    const int ktop = g.nzg - 1;
    const unsigned plane = (unsigned)g.plane;
    const unsigned col = (unsigned)j * g.nx + i;
    const TS *ps[2] = {(const TS *)fl.q[0], (const TS *)fl.q[1]};
    const TS *pu = (const TS *)fl.q[VU];
    const TS *pv = (const TS *)fl.q[VV];
    const TS *pw = (const TS *)fl.q[VW];
    // horizontal neighbour column offsets for the centred advecting w
    unsigned cx[NC], cy[NC];
#pragma unroll
    for (int m = 0; m < NC; ++m) {
        cx[m] = (unsigned)j * g.nx + wrapi(i + m - (RM - 2), g.nx);
        cy[m] = (unsigned)wrapi(j + m - (RM - 2), g.ny) * g.nx + i;
    }
    auto base_at = [&](int sc, int lev) { return P::kBase ? (fl.base[sc] ? fl.base[sc][lev] : 0.0) : 0.0; };
    // windows of face c: level c - (R-1) + m in slot m (rings: phase 0)
    TC rs[NS][L];
    double bw[NS][2 * RS];
    TC ru[L], rv[L], rw[L];
    TC cwx[NC], cwy[NC];  // w along x and y at level c
    int c = k0 - 1;
#pragma unroll
    for (int sc = 0; sc < NSC; ++sc)
#pragma unroll
        for (int m = 0; m < 2 * RS; ++m) {
            const int lev = c - (RS - 1) + m;
            rs[sc][m] = (TC)ld(ps[sc], lev * plane + col);
            bw[sc][m] = base_at(sc, lev);
        }
    if (MOM) {
#pragma unroll
        for (int m = 0; m < 2 * RM; ++m) {
            const unsigned o = (unsigned)(c - (RM - 1) + m) * plane + col;
            ru[m] = (TC)ld(pu, o);
            rv[m] = (TC)ld(pv, o);
            rw[m] = (TC)ld(pw, o);
        }
    }
    auto load_centred = [&](int cc) {
        if (MOM) {
            const unsigned lo = (unsigned)cc * plane;
#pragma unroll
            for (int m = 0; m < NC; ++m) {
                cwx[m] = (TC)ld(pw, lo + cx[m]);
                cwy[m] = (TC)ld(pw, lo + cy[m]);
            }
        }
    };
    // next-face window loads, issued before the current face's arithmetic
    TC nws[NS], nu = 0, nv = 0, nw = 0;
    double nbw[NS];
    auto issue_next = [&](int cc) {
        const int ls = min(cc + 1 + RS, ktop);
#pragma unroll
        for (int sc = 0; sc < NSC; ++sc) {
            nws[sc] = (TC)ld(ps[sc], ls * plane + col);
            nbw[sc] = base_at(sc, ls);
        }
        if (MOM) {
            const unsigned o = (unsigned)min(cc + 1 + RM, ktop) * plane + col;
            nu = (TC)ld(pu, o);
            nv = (TC)ld(pv, o);
            nw = (TC)ld(pw, o);
        }
    };
    // the next levels into the rings' free slots (phase PH), or, without
    // rings, the windows shifted; the base window always shifts
    auto advance = [&](auto phc) {
        constexpr int PH = RING ? decltype(phc)::value : 0;
#pragma unroll
        for (int sc = 0; sc < NSC; ++sc) {
#pragma unroll
            for (int m = 0; m < WS - 1; ++m) bw[sc][m] = bw[sc][m + 1];
            bw[sc][WS - 1] = nbw[sc];
        }
        if (RING) {
#pragma unroll
            for (int sc = 0; sc < NSC; ++sc) rs[sc][(WS + PH) % L] = nws[sc];
            if (MOM) {
                ru[(WM + PH) % L] = nu;
                rv[(WM + PH) % L] = nv;
                rw[(WM + PH) % L] = nw;
            }
        } else {
#pragma unroll
            for (int sc = 0; sc < NSC; ++sc) {
#pragma unroll
                for (int m = 0; m < WS - 1; ++m) rs[sc][m] = rs[sc][m + 1];
                rs[sc][WS - 1] = nws[sc];
            }
            if (MOM) {
#pragma unroll
                for (int m = 0; m < WM - 1; ++m) {
                    ru[m] = ru[m + 1];
                    rv[m] = rv[m + 1];
                    rw[m] = rw[m + 1];
                }
                ru[WM - 1] = nu;
                rv[WM - 1] = nv;
                rw[WM - 1] = nw;
            }
        }
    };
    // the windows of the face at phase PH
    TC ws[NS][WS], wsu[WM], wsv[WM], wsw[WM];
    auto windows = [&](auto phc) {
        constexpr int PH = RING ? decltype(phc)::value : 0;
#pragma unroll
        for (int sc = 0; sc < NSC; ++sc)
#pragma unroll
            for (int m = 0; m < WS; ++m) ws[sc][m] = rs[sc][(m + PH) % L];
        if (MOM) {
#pragma unroll
            for (int m = 0; m < WM; ++m) {
                wsu[m] = ru[(m + PH) % L];
                wsv[m] = rv[(m + PH) % L];
                wsw[m] = rw[(m + PH) % L];
            }
        }
    };
    SFlux Fs[NS], Fs_prev[NS];
    TC Fu = 0, Fv = 0, Fw = 0, Fu_prev = 0, Fv_prev = 0, Fw_prev = 0;
    auto fluxes = [&](int cc, SFlux *F, TC &Fu_, TC &Fv_, TC &Fw_) {
        const double rho0c = pr.rho0[cc];
        const TC W = MOM ? wsw[RM - 1] : (TC)ld(pw, (unsigned)cc * plane + col);
#pragma unroll
        for (int sc = 0; sc < NSC; ++sc) F[sc] = P::template scal_flux<SO, true>(ws[sc], bw[sc], W, rho0c);
        if (MOM) {
            // u advected by w: U = centred(w along x), rho0[k]
            Fu_ = P::template mom_flux<MO>(wsu, Interp<MO, TC>::centred(cwx), rho0c);
            // v advected by w: U = centred(w along y), rho0[k]
            Fv_ = P::template mom_flux<MO>(wsv, Interp<MO, TC>::centred(cwy), rho0c);
            // w advected by w: U = centred(w along z) from the window, rho0_half[k+1]
            TC q[NC];
#pragma unroll
            for (int m = 0; m < NC; ++m) q[m] = wsw[m + 1];
            Fw_ = P::template mom_flux<MO>(wsw, Interp<MO, TC>::centred(q), pr.rho0_half[cc + 1]);
        }
    };
    const double dzi = g.dxi[2];
    issue_next(c);
    load_centred(c);
    windows(std::integral_constant<int, 0>{});
    fluxes(c, Fs_prev, Fu_prev, Fv_prev, Fw_prev);
    advance(std::integral_constant<int, 0>{});
    // faces at phases 1, 2, .., L-1, 0, 1, ...
    auto face = [&](auto phc) {
        // the level as an opaque scalar: offsets affine in the loop counter
        // would become one induction register per unrolled face and address
        const int cr = __builtin_amdgcn_readfirstlane(c);
        const unsigned o = (unsigned)cr * plane + col;
        issue_next(cr);
        load_centred(cr);
        windows(phc);
        fluxes(cr, Fs, Fu, Fv, Fw);
        // this cell's tendencies and stage copies, all read before any store
        // (the compiler must assume the output arrays alias them); read after
        // the flux arithmetic, which keeps them out of its register peak
        TS told[5] = {}, cpv[5] = {};
#pragma unroll
        for (int v = 0; v < 5; ++v) {
            if (v < NSC || (MOM && v >= VU)) {
                told[v] = lds_((const TS *)fl.t[v], o);
                if (RKM == 2) cpv[v] = lds_((const TS *)fl.cp[v], o);
            }
        }
        const double ah = pr.alpha0_half[cr];
        // the final tendencies, with an uncontracted subtraction so the
        // stored (separate RK3) and fused paths round identically
        TS tn[5] = {};
#pragma unroll
        for (int sc = 0; sc < NSC; ++sc)
            tn[sc] = add_rn(told[sc], -P::template scal_div<true>(Fs[sc], Fs_prev[sc], ah, dzi, 1.0));
        if (MOM) {
            tn[VU] = add_rn(told[VU], -P::mom_div(Fu, Fu_prev, ah, dzi));
            tn[VV] = add_rn(told[VV], -P::mom_div(Fv, Fv_prev, ah, dzi));
            tn[VW] = add_rn(told[VW], -P::mom_div(Fw, Fw_prev, pr.alpha0[cr], dzi));
        }
        if (RKM == 0) {
#pragma unroll
            for (int v = 0; v < 5; ++v)
                if (v < NSC || (MOM && v >= VU)) sts_((TS *)fl.t[v], o, tn[v]);
        } else {
#pragma unroll
            for (int sc = 0; sc < NSC; ++sc)
                rk_store<TS, RKM>(fl, sc, o, col, plane, cr, g.gw, g.nzg, (TS)ws[sc][RS - 1], tn[sc], cpv[sc], false);
            if (MOM) {
                rk_store<TS, RKM>(fl, VU, o, col, plane, cr, g.gw, g.nzg, (TS)wsu[RM - 1], tn[VU], cpv[VU], false);
                rk_store<TS, RKM>(fl, VV, o, col, plane, cr, g.gw, g.nzg, (TS)wsv[RM - 1], tn[VV], cpv[VV], false);
                rk_store<TS, RKM>(fl, VW, o, col, plane, cr, g.gw, g.nzg, (TS)wsw[RM - 1], tn[VW], cpv[VW], true);
            }
        }
#pragma unroll
        for (int sc = 0; sc < NSC; ++sc) Fs_prev[sc] = Fs[sc];
        Fu_prev = Fu;
        Fv_prev = Fv;
        Fw_prev = Fw;
        advance(phc);
    };
    c = k0;
    if (c < k1) {
        auto step = [&](auto q) {
            face(std::integral_constant<int, (1 + decltype(q)::value) % L>{});
            return ++c < k1;
        };
        while (phases<0, RING ? L : 1>(step)) {
        }
    }
}

// -------------------------------------------------------------------------
// y pass: one thread per (i, k) with k interior, marching faces c = j0-1 .. j1-1
// (periodic).  Block (64, 4); grid (ceil(nx/64), ceil(nz/4), nchunks).
// Pipelined like the z pass: the next row of every window and this face's
// centred v stencils are loaded at the top of each face, the tendencies
// together after the fluxes; the first face is peeled off the loop.
// The stencil windows are rings of L = max(2 RS, 2 RM) + 1 registers: at
// phase p (face count mod L) window slot m is ring slot (m + p) % L, and the
// next row loads into the slot that the oldest row has left, so the march,
// unrolled by L faces, moves no register (a shifted window costs one copy per
// entry and face).  Above L = 8 (WENO7 and higher) the unrolled march would
// outgrow the instruction cache, and the windows shift instead (L = W).
template <class P, int SO, int MO, int NSC, bool MOM>
__global__ __launch_bounds__(256) void adv_y(Geo g, Prof pr, Fields fl, int jchunk) {
    using TS = typename P::TS;
    using TC = typename P::TC;
    using SFlux = typename P::SFlux;
    constexpr int RS = Interp<SO, TC>::R;
    constexpr int RM = Interp<MO, TC>::R;
    constexpr int NC = 2 * (RM - 1);
    constexpr int NS = NSC > 0 ? NSC : 1;
    constexpr int WS = 2 * RS, WM = 2 * RM;
    constexpr int WMAX = (NSC > 0 && WS > WM) || !MOM ? WS : WM;
    constexpr bool RING = WMAX + 1 <= 8;
    constexpr int L = RING ? WMAX + 1 : WMAX;
    const uint3 tile = xcd_tile<1>();  // XCD bands of 4-level blocks in z
    const int i = tile.x * 64 + threadIdx.x;
    const int kk = tile.y * 4 + threadIdx.y;
    if (i >= g.nx || kk >= g.nz) return;
    const int k = kk + g.gw;
    const int j0 = tile.z * jchunk;
    const int j1 = min(j0 + jchunk, g.ny);
    const unsigned plane = (unsigned)g.plane;
    const unsigned nx = (unsigned)g.nx;
    const unsigned lev = (unsigned)k * plane + i;  // offset of (i, 0, k)
    const TS *ps[2] = {(const TS *)fl.q[0], (const TS *)fl.q[1]};
    const TS *pu = (const TS *)fl.q[VU];
    const TS *pv = (const TS *)fl.q[VV];
    const TS *pw = (const TS *)fl.q[VW];
    const double rh = pr.rho0_half[k];
    const double r0 = pr.rho0[k];
    const double ah = pr.alpha0_half[k];
    const double a0 = pr.alpha0[k];
    const double dyi = g.dxi[1];
    // offsets (relative to the row) of the centred stencils of v along x and z
    int dxo[NC], dzo[NC];
#pragma unroll
    for (int m = 0; m < NC; ++m) {
        dxo[m] = wrapi(i + m - (RM - 2), g.nx) - i;
        dzo[m] = (m - (RM - 2)) * (int)plane;
    }
    double bw[NS][1], brho[NS];
#pragma unroll
    for (int sc = 0; sc < NSC; ++sc) {
        bw[sc][0] = P::kBase ? (fl.base[sc] ? fl.base[sc][k] : 0.0) : 0.0;
        brho[sc] = bw[sc][0] * rh;
    }
    // rings (phase 0: window slot m in ring slot m)
    TC rs[NS][L];
    TC ru[L], rv[L], rw[L];
    TC cvx[NC], cvz[NC];  // v along x and z at row c
    int c = j0 - 1;
#pragma unroll
    for (int sc = 0; sc < NSC; ++sc)
#pragma unroll
        for (int m = 0; m < WS; ++m)
            rs[sc][m] = (TC)ld(ps[sc], lev + (unsigned)wrapi(c - (RS - 1) + m, g.ny) * nx);
    if (MOM) {
#pragma unroll
        for (int m = 0; m < WM; ++m) {
            const unsigned o = lev + (unsigned)wrapi(c - (RM - 1) + m, g.ny) * nx;
            ru[m] = (TC)ld(pu, o);
            rv[m] = (TC)ld(pv, o);
            rw[m] = (TC)ld(pw, o);
        }
    }
    // rows (wrapped) of the entries that enter the windows at face c + 1,
    // c + 1 + R, and of the current face
    int jn_s = wrapi(c + 1 + RS, g.ny), jn_m = wrapi(c + 1 + RM, g.ny);
    int cw = c < 0 ? c + g.ny : c;
    SFlux Fs[NS], Fs_prev[NS];
    TC Fu = 0, Fv = 0, Fw = 0, Fu_prev = 0, Fv_prev = 0, Fw_prev = 0;
    // one face at ring phase PH (a compile-time constant when RING): the next
    // rows into the free ring slots, the centred stencils, the fluxes and,
    // unless FIRST, the tendencies; then the row counters advance (and, when
    // the windows do not rotate, they shift)
    auto face = [&](auto phc, auto firstc) {
        constexpr int PH = RING ? decltype(phc)::value : 0;
        constexpr bool FIRST = decltype(firstc)::value;
        auto slot = [](int m) { return (m + PH) % L; };
        const unsigned orow = lev + (unsigned)cw * nx;  // (i, c, k)
        // next rows: window slot W of this face (slot W - 1 of the next)
        TC nws[NS], nu = 0, nv = 0, nw = 0;
#pragma unroll
        for (int sc = 0; sc < NSC; ++sc) nws[sc] = (TC)ld(ps[sc], lev + (unsigned)jn_s * nx);
        if (MOM) {
            const unsigned o = lev + (unsigned)jn_m * nx;
            nu = (TC)ld(pu, o);
            nv = (TC)ld(pv, o);
            nw = (TC)ld(pw, o);
        }
        if (RING) {
#pragma unroll
            for (int sc = 0; sc < NSC; ++sc) rs[sc][slot(WS)] = nws[sc];
            if (MOM) {
                ru[slot(WM)] = nu;
                rv[slot(WM)] = nv;
                rw[slot(WM)] = nw;
            }
        }
        if (MOM) {
#pragma unroll
            for (int m = 0; m < NC; ++m) {
                cvx[m] = (TC)ld(pv, orow + dxo[m]);
                cvz[m] = (TC)ld(pv, orow + dzo[m]);
            }
        }
        // keep the loads above the flux arithmetic (the scheduler would sink
        // them next to their first use)
        if (!FIRST) __builtin_amdgcn_sched_barrier(0);
        // this face's windows
        TC ws[NS][WS], wsu[WM], wsv[WM], wsw[WM];
#pragma unroll
        for (int sc = 0; sc < NSC; ++sc)
#pragma unroll
            for (int m = 0; m < WS; ++m) ws[sc][m] = rs[sc][slot(m)];
        if (MOM) {
#pragma unroll
            for (int m = 0; m < WM; ++m) {
                wsu[m] = ru[slot(m)];
                wsv[m] = rv[slot(m)];
                wsw[m] = rw[slot(m)];
            }
        }
        SFlux *F = FIRST ? Fs_prev : Fs;
        TC &Fu_ = FIRST ? Fu_prev : Fu;
        TC &Fv_ = FIRST ? Fv_prev : Fv;
        TC &Fw_ = FIRST ? Fw_prev : Fw;
        const TC V = MOM ? wsv[RM - 1] : (TC)ld(pv, orow);
#pragma unroll
        for (int sc = 0; sc < NSC; ++sc) F[sc] = P::template scal_flux<SO, false>(ws[sc], bw[sc], V, rh);
        // the scalar fluxes cover the latency of the centred stencil loads
        __builtin_amdgcn_sched_barrier(0);
        if (MOM) {
            Fu_ = P::template mom_flux<MO>(wsu, Interp<MO, TC>::centred(cvx), rh);  // v along x
            TC q[NC];
#pragma unroll
            for (int m = 0; m < NC; ++m) q[m] = wsv[m + 1];  // v along y
            Fv_ = P::template mom_flux<MO>(wsv, Interp<MO, TC>::centred(q), rh);
            Fw_ = P::template mom_flux<MO>(wsw, Interp<MO, TC>::centred(cvz), r0);  // v along z
        }
        if (!FIRST) {
            // tendencies, all read before any store (see the z pass)
            TS told[5] = {};
#pragma unroll
            for (int v = 0; v < 5; ++v)
                if (v < NSC || (MOM && v >= VU)) told[v] = lds_((const TS *)fl.t[v], orow);
#pragma unroll
            for (int sc = 0; sc < NSC; ++sc)
                sts_((TS *)fl.t[sc], orow,
                     told[sc] - P::template scal_div<false>(Fs[sc], Fs_prev[sc], ah, dyi, brho[sc]));
            if (MOM) {
                sts_((TS *)fl.t[VU], orow, (TS)(told[VU] - P::mom_div(Fu, Fu_prev, ah, dyi)));
                sts_((TS *)fl.t[VV], orow, (TS)(told[VV] - P::mom_div(Fv, Fv_prev, ah, dyi)));
                sts_((TS *)fl.t[VW], orow, (TS)(told[VW] - P::mom_div(Fw, Fw_prev, a0, dyi)));
            }
#pragma unroll
            for (int sc = 0; sc < NSC; ++sc) Fs_prev[sc] = Fs[sc];
            Fu_prev = Fu;
            Fv_prev = Fv;
            Fw_prev = Fw;
        }
        if (!RING) {
#pragma unroll
            for (int sc = 0; sc < NSC; ++sc) {
#pragma unroll
                for (int m = 0; m < WS - 1; ++m) rs[sc][m] = rs[sc][m + 1];
                rs[sc][WS - 1] = nws[sc];
            }
            if (MOM) {
#pragma unroll
                for (int m = 0; m < WM - 1; ++m) {
                    ru[m] = ru[m + 1];
                    rv[m] = rv[m + 1];
                    rw[m] = rw[m + 1];
                }
                ru[WM - 1] = nu;
                rv[WM - 1] = nv;
                rw[WM - 1] = nw;
            }
        }
        cw = cw + 1 == g.ny ? 0 : cw + 1;
        jn_s = jn_s + 1 == g.ny ? 0 : jn_s + 1;
        jn_m = jn_m + 1 == g.ny ? 0 : jn_m + 1;
    };
    face(std::integral_constant<int, 0>{}, std::true_type{});
    c = j0;
    if (c < j1) {
        // faces at phases 1, 2, .., L-1, 0, 1, ...: each returns whether
        // another face follows
        auto step = [&](auto q) {
            face(std::integral_constant<int, (1 + decltype(q)::value) % L>{}, std::false_type{});
            return ++c < j1;
        };
        while (phases<0, RING ? L : 1>(step)) {
        }
    }
}

// -------------------------------------------------------------------------
// y and z passes fused (WENO5 scalars and momentum): one thread per (i, j)
// column marching the faces c = k0-1 .. k1-1 of a z chunk as the z pass does,
// with the y face fluxes of each level computed from LDS.  A block of 64 x TY
// threads covers 64 columns of TY rows.  Per face c:
//  A. stage level c of every field in LDS for the block's rows (from the
//     march's windows) and R halo rows on each side, and v and w with the x
//     halo of the centred stencils (halo values loaded one face ahead); then
//     issue this cell's tendency and stage-copy loads, the next window level
//     and the next face's halo.
//  B. (barrier) the z face flux from the windows, with w centred along x and
//     y from the tile; the y flux of the face above each row from the tile,
//     into LDS; the block's lowest face (below row j0) one variable per wave.
//  C. (barrier) each cell's tendency takes the y flux difference (the lower
//     face from LDS), rounded as the y pass stores it, then the z flux
//     difference, in the per-variable order of the separate passes, so the
//     bits are theirs; then the RK3 stage as in the z pass.
// Against the y pass followed by the z pass this drops the y pass's tendency
// read and write and its re-read of the fields.  Columns beyond a partial
// tile (nx not a multiple of 64, ny of TY; PART) march wrapped copies of
// real columns and store nothing, so every thread reaches every barrier.
// A chunk holds at most ADV_YZ_KMAX levels (advyz_zsplit): its level
// profiles are staged in LDS (read from global memory they are vector loads,
// whose waits would drain the prefetches in flight).

template <class P, int SO, int MO, int NSC, bool MOM, int RKM, int TY, bool PART>
__global__ __launch_bounds__(64 * TY, P::kYZWaves) void adv_yz(Geo g, Prof pr, Fields fl, int kchunk) {
    using TS = typename P::TS;
    using TC = typename P::TC;
    using SFlux = typename P::SFlux;
    constexpr int R = Interp<SO, TC>::R;
    static_assert(SO == 5 && MO == 5 && MOM && NSC >= 1 && NSC <= 2, "fused y-z pass: WENO5 scalars and momentum");
    constexpr int NC = 2 * (R - 1);  // centred stencil (interp_4)
    constexpr int NR = NC + 1;       // ring of v levels: NC read per face, one staged ahead
    constexpr int WN = 2 * R;        // WENO window
    constexpr int XL = R - 2;        // centred x stencil: i - XL .. i + XL + NC - 1
    constexpr int TX = 64 + NC - 1;
    constexpr int TR = TY + 2 * R;  // tile rows j0 - R .. j0 + TY + R - 1
    constexpr int NV = NSC + 3;
    constexpr int FU = NSC, FV = NSC + 1, FW = NSC + 2;  // tile field index
    constexpr int NHR = (2 * R + TY - 1) / TY;           // halo rows per thread row
    constexpr int NXH = (2 * TY + 1) * (NC - 1);         // x halo items (v: TY + 1 rows, w: TY rows)
    constexpr int NPL = ADV_YZ_KMAX + 2 * R;             // profile levels k0 - R .. k0 + kchunk + R - 1
    constexpr int NPR = P::kBase ? 4 + NSC : 4;
    static_assert(NXH <= 64 * TY, "x halo items exceed the block");
    // tile and face fluxes double-buffered by level parity: level c + 1 is
    // staged while slower waves still read level c, so one barrier per face
    __shared__ TC tile[2][NV][TR][TX];
    __shared__ TC flx[2][NV][TY + 1][64];  // y flux of face j0 - 1 + f + 1/2 in slot f
    __shared__ TC vring[NR][64];           // v of row j0 - 1 at level lev in slot lev % NR
    __shared__ double lp[NPR][NPL];        // rho0, rho0_half, alpha0, alpha0_half (, scalar bases)
    const uint3 tl = xcd_tile<1>();        // XCD bands of TY-row blocks in y
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int nx = g.nx, ny = g.ny;
    const int ib = tl.x * 64, j0 = tl.y * TY;
    // a real column, which stores its cells (PART: the grid has partial
    // tiles; otherwise every thread owns its column, and the stores are not
    // under a branch, into which the compiler would sink the tendency loads)
    const bool own = !PART || (ib + tx < nx && j0 + ty < ny);
    const int iw = (ib + tx) % nx, jw = (j0 + ty) % ny;
    const int k0 = g.gw + tl.z * kchunk;
    const int k1 = min(k0 + kchunk, g.nz + g.gw);
    const int ktop = g.nzg - 1;
    const int kb = k0 - R;  // level of profile slot 0
    const unsigned plane = (unsigned)g.plane;
    const unsigned col = (unsigned)jw * nx + iw;
    auto wrap = [](int a, int n) { return ((a % n) + n) % n; };
    for (int l = ty * 64 + tx; l < NPL; l += 64 * TY) {
        const int lev = min(kb + l, ktop);
        lp[0][l] = pr.rho0[lev];
        lp[1][l] = pr.rho0_half[lev];
        lp[2][l] = pr.alpha0[lev];
        lp[3][l] = pr.alpha0_half[lev];
        if (P::kBase) {
#pragma unroll
            for (int sc = 0; sc < NSC; ++sc) lp[(4 + sc) % NPR][l] = fl.base[sc] ? fl.base[sc][lev] : 0.0;
        }
    }
    auto prof = [&](int q, int lev) { return lp[q][lev - kb]; };
    const TS *pf[NV];
#pragma unroll
    for (int sc = 0; sc < NSC; ++sc) pf[sc] = (const TS *)fl.q[sc];
    pf[FU] = (const TS *)fl.q[VU];
    pf[FV] = (const TS *)fl.q[VV];
    pf[FW] = (const TS *)fl.q[VW];
    // halo rows of this thread row: hr = ty + h TY < 2R, tile row hr (below)
    // or TY + hr (above); the others load row 0's cells and store nothing
    unsigned hoff[NHR];
    int hrow[NHR];
    bool hval[NHR];
#pragma unroll
    for (int h = 0; h < NHR; ++h) {
        const int hr = ty + h * TY;
        hval[h] = hr < 2 * R;
        hrow[h] = !hval[h] ? 0 : (hr < R ? hr : TY + hr);
        hoff[h] = (unsigned)wrap(j0 - R + hrow[h], ny) * nx + iw;
    }
    // x halo item of this thread: v of tile rows R-1 .. R+TY-1, then w of
    // tile rows R .. R+TY-1; tile columns 0 .. XL-1 and 64+XL .. TX-1
    const int tl1 = ty * 64 + tx;
    const bool xval = tl1 < NXH;
    const int rf = (xval ? tl1 : 0) / (NC - 1), cs = (xval ? tl1 : 0) % (NC - 1);
    const bool xw = rf > TY;
    const int xrow = xw ? R + (rf - TY - 1) : R - 1 + rf;
    const int xcol = cs < XL ? cs : 64 + cs;
    const unsigned xoff = (unsigned)wrap(j0 - R + xrow, ny) * nx + wrap(ib - XL + xcol, nx);
    const TS *xptr = xw ? pf[FW] : pf[FV];
    const unsigned voff = (unsigned)wrap(j0 - 1, ny) * nx + iw;  // v of row j0 - 1

    // windows of face c: level c - (R-1) + m in slot m
    TC rs[NSC][WN], ru[WN], rv[WN], rw[WN];
    int c = k0 - 1;
#pragma unroll
    for (int m = 0; m < WN; ++m) {
        const unsigned o = (unsigned)(c - (R - 1) + m) * plane + col;
#pragma unroll
        for (int sc = 0; sc < NSC; ++sc) rs[sc][m] = (TC)ld(pf[sc], o);
        ru[m] = (TC)ld(pf[FU], o);
        rv[m] = (TC)ld(pf[FV], o);
        rw[m] = (TC)ld(pf[FW], o);
    }
    // the halo of the tile at level lev and the ring's level lev + 2, for
    // staging level lev (every load unconditional; the ring and x halo first)
    TC hp[NHR][NV], xh, vr;
    auto prefetch = [&](int lev) {
        const unsigned lo = (unsigned)min(lev, ktop) * plane;
        vr = (TC)ld(pf[FV], (unsigned)min(lev + 2, ktop) * plane + voff);
        xh = (TC)ld(xptr, lo + xoff);
#pragma unroll
        for (int h = 0; h < NHR; ++h)
#pragma unroll
            for (int v = 0; v < NV; ++v) hp[h][v] = (TC)ld(pf[v], lo + hoff[h]);
    };
    // stage level cc into tile[cc & 1]; slot m of the windows holds it
    auto stage = [&](int cc, int m) {
        auto &t = tile[cc & 1];
        if (ty == TY - 1) vring[(cc + 2) % NR][tx] = vr;
        if (xval) t[xw ? FW : FV][xrow][xcol] = xh;
#pragma unroll
        for (int sc = 0; sc < NSC; ++sc) t[sc][R + ty][XL + tx] = rs[sc][m];
        t[FU][R + ty][XL + tx] = ru[m];
        t[FV][R + ty][XL + tx] = rv[m];
        t[FW][R + ty][XL + tx] = rw[m];
#pragma unroll
        for (int h = 0; h < NHR; ++h)
            if (hval[h]) {
#pragma unroll
                for (int v = 0; v < NV; ++v) t[v][hrow[h]][XL + tx] = hp[h][v];
            }
    };
    prefetch(c);
    if (ty == TY - 1) {
#pragma unroll
        for (int m = 0; m < NC - 1; ++m) {
            const int lev = c - 1 + m;
            vring[lev % NR][tx] = (TC)ld(pf[FV], (unsigned)lev * plane + voff);
        }
    }
    stage(c, R - 1);
    prefetch(c + 1);
    __syncthreads();

    SFlux Fs[NSC], Fs_prev[NSC];
    TC Fu = 0, Fv = 0, Fw = 0, Fu_prev = 0, Fv_prev = 0, Fw_prev = 0;
    const double dzi = g.dxi[2], dyi = g.dxi[1];
    // one face; FIRST: the chunk's lowest face (k0 - 1), z flux only.
    //  B. fluxes from tile[c & 1]; y fluxes into flx[c & 1]; this cell's
    //     tendency loads; level c + 1 staged into tile[(c + 1) & 1]; the
    //     next window level and the halo of level c + 2 loaded
    //  (barrier)
    //  C. flux differences, stage update, window shift
    auto face = [&](auto firstc) {
        constexpr bool FIRST = decltype(firstc)::value;
        const unsigned o = (unsigned)c * plane + col;
        const int b = c & 1;
        // the z face flux
        {
            TC cwx[NC], cwy[NC];
#pragma unroll
            for (int m = 0; m < NC; ++m) {
                cwx[m] = tile[b][FW][R + ty][tx + m];            // w along x
                cwy[m] = tile[b][FW][R + ty - XL + m][XL + tx];  // w along y
            }
            const double rho0c = prof(0, c);
#pragma unroll
            for (int sc = 0; sc < NSC; ++sc) {
                double bz[WN];
#pragma unroll
                for (int m = 0; m < WN; ++m) bz[m] = P::kBase ? prof((4 + sc) % NPR, c - (R - 1) + m) : 0.0;
                Fs[sc] = P::template scal_flux<SO, true>(rs[sc], bz, rw[R - 1], rho0c);
            }
            Fu = P::template mom_flux<MO>(ru, Interp<MO, TC>::centred(cwx), rho0c);
            Fv = P::template mom_flux<MO>(rv, Interp<MO, TC>::centred(cwy), rho0c);
            TC q[NC];
#pragma unroll
            for (int m = 0; m < NC; ++m) q[m] = rw[m + 1];  // w along z
            Fw = P::template mom_flux<MO>(rw, Interp<MO, TC>::centred(q), prof(1, c + 1));
        }
        TC Fy[NV] = {}, Vp = 0;
        const TC Vc = rv[R - 1];
        const double rh = prof(1, c), r0 = prof(0, c);
        TS told[5] = {}, cpv[5] = {};
        if constexpr (!FIRST) {
            // the face above this row: window rows j-R+1 .. j+R = tile rows ty+1 ..
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                TC win[WN];
#pragma unroll
                for (int m = 0; m < WN; ++m) win[m] = tile[b][v][ty + 1 + m][XL + tx];
                TC q[NC];
                if (v < NSC) {
                    const double bdum[WN] = {};
                    Fy[v] = P::x_keep(P::template scal_flux<SO, false>(win, bdum, Vc, rh));
                } else if (v == FU) {
#pragma unroll
                    for (int m = 0; m < NC; ++m) q[m] = tile[b][FV][R + ty][tx + m];  // v along x
                    Fy[v] = P::template mom_flux<MO>(win, Interp<MO, TC>::centred(q), rh);
                } else if (v == FV) {
#pragma unroll
                    for (int m = 0; m < NC; ++m) q[m] = win[m + 1];  // v along y
                    Fy[v] = P::template mom_flux<MO>(win, Interp<MO, TC>::centred(q), rh);
                } else {
#pragma unroll
                    for (int m = 0; m < NC; ++m) q[m] = rv[m + 1];  // v along z
                    Fy[v] = P::template mom_flux<MO>(win, Interp<MO, TC>::centred(q), r0);
                }
                flx[b][v][ty + 1][tx] = Fy[v];
            }
            Vp = tile[b][FV][R + ty - 1][XL + tx];
            // the face below row j0, one variable per wave (uniform branches)
            for (int ev = ty; ev < NV; ev += TY) {
                TC win[WN], q[NC];
#pragma unroll
                for (int m = 0; m < WN; ++m) win[m] = tile[b][ev][m][XL + tx];
                TC F;
                if (ev < NSC) {
                    const double bdum[WN] = {};
                    F = P::x_keep(P::template scal_flux<SO, false>(win, bdum, tile[b][FV][R - 1][XL + tx], rh));
                } else if (ev == FU) {
#pragma unroll
                    for (int m = 0; m < NC; ++m) q[m] = tile[b][FV][R - 1][tx + m];
                    F = P::template mom_flux<MO>(win, Interp<MO, TC>::centred(q), rh);
                } else if (ev == FV) {
#pragma unroll
                    for (int m = 0; m < NC; ++m) q[m] = win[m + 1];
                    F = P::template mom_flux<MO>(win, Interp<MO, TC>::centred(q), rh);
                } else {
#pragma unroll
                    for (int m = 0; m < NC; ++m) q[m] = vring[(c - 1 + m) % NR][tx];
                    F = P::template mom_flux<MO>(win, Interp<MO, TC>::centred(q), r0);
                }
                flx[b][ev][0][tx] = F;
            }
            // this cell's tendencies and stage copies, read before any store
            // (the outputs may alias the copies), after the flux arithmetic,
            // which keeps them out of its register peak
#pragma unroll
            for (int v = 0; v < 5; ++v)
                if (v < NSC || v >= VU) {
                    told[v] = lds_((const TS *)fl.t[v], o);
                    if (RKM == 2) cpv[v] = lds_((const TS *)fl.cp[v], o);
                }
        }
        // level c + 1 (window slot R) into the other tile, then the loads for
        // the next face
        stage(c + 1, R);
        TC nws[NSC], nu, nv, nw;
        {
            const unsigned on = (unsigned)min(c + 1 + R, ktop) * plane + col;
#pragma unroll
            for (int sc = 0; sc < NSC; ++sc) nws[sc] = (TC)ld(pf[sc], on);
            nu = (TC)ld(pf[FU], on);
            nv = (TC)ld(pf[FV], on);
            nw = (TC)ld(pf[FW], on);
        }
        prefetch(c + 2);
        __syncthreads();
        if constexpr (!FIRST) {
            // C: y then z flux differences, then the stage update
            const double ah = prof(3, c), a0 = prof(2, c);
            TS tn[5] = {};
#pragma unroll
            for (int sc = 0; sc < NSC; ++sc) {
                const double brho = P::kBase ? prof((4 + sc) % NPR, c) * rh : 0.0;
                const TS t1 = told[sc] - P::x_scal_div(Fy[sc], flx[b][sc][ty][tx], Vc, Vp, ah, dyi, brho);
                tn[sc] = add_rn(t1, -P::template scal_div<true>(Fs[sc], Fs_prev[sc], ah, dzi, 1.0));
            }
            const TS tu = (TS)(told[VU] - P::mom_div(Fy[FU], flx[b][FU][ty][tx], ah, dyi));
            const TS tv = (TS)(told[VV] - P::mom_div(Fy[FV], flx[b][FV][ty][tx], ah, dyi));
            const TS tw = (TS)(told[VW] - P::mom_div(Fy[FW], flx[b][FW][ty][tx], a0, dyi));
            tn[VU] = add_rn(tu, -P::mom_div(Fu, Fu_prev, ah, dzi));
            tn[VV] = add_rn(tv, -P::mom_div(Fv, Fv_prev, ah, dzi));
            tn[VW] = add_rn(tw, -P::mom_div(Fw, Fw_prev, a0, dzi));
            if (own) {
                if (RKM == 0) {
#pragma unroll
                    for (int v = 0; v < 5; ++v)
                        if (v < NSC || v >= VU) sts_((TS *)fl.t[v], o, tn[v]);
                } else {
#pragma unroll
                    for (int sc = 0; sc < NSC; ++sc)
                        rk_store<TS, RKM>(fl, sc, o, col, plane, c, g.gw, g.nzg, (TS)rs[sc][R - 1], tn[sc], cpv[sc],
                                          false);
                    rk_store<TS, RKM>(fl, VU, o, col, plane, c, g.gw, g.nzg, (TS)ru[R - 1], tn[VU], cpv[VU], false);
                    rk_store<TS, RKM>(fl, VV, o, col, plane, c, g.gw, g.nzg, (TS)rv[R - 1], tn[VV], cpv[VV], false);
                    rk_store<TS, RKM>(fl, VW, o, col, plane, c, g.gw, g.nzg, (TS)rw[R - 1], tn[VW], cpv[VW], true);
                }
            }
        }
#pragma unroll
        for (int sc = 0; sc < NSC; ++sc) Fs_prev[sc] = Fs[sc];
        Fu_prev = Fu;
        Fv_prev = Fv;
        Fw_prev = Fw;
        // shift the windows
#pragma unroll
        for (int m = 0; m < WN - 1; ++m) {
#pragma unroll
            for (int sc = 0; sc < NSC; ++sc) rs[sc][m] = rs[sc][m + 1];
            ru[m] = ru[m + 1];
            rv[m] = rv[m + 1];
            rw[m] = rw[m + 1];
        }
#pragma unroll
        for (int sc = 0; sc < NSC; ++sc) rs[sc][WN - 1] = nws[sc];
        ru[WN - 1] = nu;
        rv[WN - 1] = nv;
        rw[WN - 1] = nw;
    };
    face(std::true_type{});
    for (c = k0; c < k1; ++c) face(std::false_type{});
}

// -------------------------------------------------------------------------
// x pass: one block per (j, k) row, in three phases separated by barriers.
//  1. stage in LDS: the rows of every field (plus a periodic halo of H
//     cells), the advecting u centred along y and z for the v and w fluxes
//     (gathered from the neighbouring rows and levels), and the tendencies
//     the pass must read (w's, which thermodynamics wrote; all five unless
//     the others are known to be zero).  Every global load is issued here.
//  2. every face flux of every variable into LDS.
//  3. the flux divergences, stored once per cell and variable.
// LDS (TC words): rows NV*W | fluxes NF*nx | centred u 2*nx | old tendencies NT*nx.
template <class P, int SO, int MO, int NSC, bool MOM>
struct XLayout {
    using TC = typename P::TC;
    static constexpr int RS = Interp<SO, TC>::R;
    static constexpr int RM = Interp<MO, TC>::R;
    static constexpr int H = (MOM && RM > RS) ? RM : RS;
    static constexpr int NV = NSC + (MOM ? 3 : 1);  // u row always staged (advecting velocity)
    static constexpr int NF = NSC + (MOM ? 3 : 0);
    // tendencies staged: w's when the others are zero on entry, else all
    static __host__ __device__ int nt(bool tzero) { return tzero ? (MOM ? 1 : 0) : NF; }
    static __host__ __device__ size_t words(int nx, bool tzero) {
        return (size_t)NV * (nx + 2 * H) + (size_t)NF * nx + (MOM ? 2 * (size_t)nx : 0) + (size_t)nt(tzero) * nx;
    }
};

template <class P, int SO, int MO, int NSC, bool MOM>
__global__ __launch_bounds__(256) void adv_x(Geo g, Prof pr, Fields fl) {
    using TS = typename P::TS;
    using TC = typename P::TC;
    using L = XLayout<P, SO, MO, NSC, MOM>;
    constexpr int RS = L::RS, RM = L::RM, H = L::H, NV = L::NV, NF = L::NF;
    constexpr int NC = 2 * (RM - 1);
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int nx = g.nx;
    const int W = nx + 2 * H;
    const bool tzero = fl.tzero != 0;
    TC *rows = reinterpret_cast<TC *>(smem_raw);
    TC *flux = rows + NV * W;          // flux[f * nx + c]: face c+1/2 of variable f
    TC *ucen = flux + NF * nx;         // u centred along y (v flux) then along z (w flux)
    TC *told = ucen + (MOM ? 2 * nx : 0);
    const uint3 tile = xcd_tile<0>();  // XCD bands of rows in y
    const int j = tile.x;
    const int k = tile.y + g.gw;
    const unsigned plane = (unsigned)g.plane;
    const unsigned row = ((unsigned)k * g.ny + j) * nx;
    const int tid = threadIdx.x;
    const int B = blockDim.x;
    const TS *src[5];
#pragma unroll
    for (int sc = 0; sc < NSC; ++sc) src[sc] = (const TS *)fl.q[sc];
    src[NSC] = (const TS *)fl.q[VU];
    if (MOM) {
        src[NSC + 1] = (const TS *)fl.q[VV];
        src[NSC + 2] = (const TS *)fl.q[VW];
    }
    // variable f of the flux rows -> field index (scalars, then u, v, w)
    auto field = [](int f) { return f < NSC ? f : VU + (f - NSC); };

    // phase 1, in batches whose loads are all issued before the LDS writes
    // (a load written to LDS inside a plain loop waits once per element).
    // Even nx: the row's cells in pairs (one 2-vector load per field and
    // pair, no wrap arithmetic), then the 2H periodic halo cells; otherwise
    // element by element with the wrap.
    if ((nx & 1) == 0) {
        using TS2 = typename std::conditional<sizeof(TS) == 4, float2, double2>::type;
        for (int e = 2 * tid; e < nx; e += 2 * B) {
            TS2 tmp[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) tmp[v] = *reinterpret_cast<const TS2 *>(src[v] + row + e);
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                rows[v * W + H + e] = (TC)tmp[v].x;
                rows[v * W + H + e + 1] = (TC)tmp[v].y;
            }
        }
        for (int h = tid; h < 2 * H; h += B) {
            const int si = h < H ? nx - H + h : h - H;  // left halo: the row's end; right: its start
            const int di = h < H ? h : nx + h;
            TC tmp[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) tmp[v] = (TC)ld(src[v], row + si);
#pragma unroll
            for (int v = 0; v < NV; ++v) rows[v * W + di] = tmp[v];
        }
    } else {
        for (int t0 = tid; t0 < W; t0 += 4 * B) {
            TC tmp[NV][4];
#pragma unroll
            for (int v = 0; v < NV; ++v)
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int t = t0 + u * B;
                    tmp[v][u] = t < W ? (TC)ld(src[v], row + wrapi(t - H, nx)) : TC(0);
                }
#pragma unroll
            for (int v = 0; v < NV; ++v)
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    if (t0 + u * B < W) rows[v * W + t0 + u * B] = tmp[v][u];
        }
    }
    const int nt = L::nt(tzero);
    const int f0 = NF - nt;  // the staged tendencies are those of the last nt variables
    if (MOM || nt > 0) {
        const TS *pu = (const TS *)fl.q[VU];
        int yo[NC], zo[NC];
#pragma unroll
        for (int m = 0; m < NC; ++m) {
            yo[m] = (wrapi(j + m - (RM - 2), g.ny) - j) * nx;
            zo[m] = (m - (RM - 2)) * (int)plane;
        }
        for (int c0 = tid; c0 < nx; c0 += 2 * B) {
            TC uy[2] = {}, uz[2] = {}, tv[2][NF] = {};
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int c = c0 + u * B;
                if (c >= nx) continue;
#pragma unroll
                for (int f = 0; f < NF; ++f)
                    if (f >= f0) tv[u][f] = (TC)ld((const TS *)fl.t[field(f)], row + c);
                if (MOM) {
                    unsigned oy[NC], oz[NC];
#pragma unroll
                    for (int m = 0; m < NC; ++m) {
                        oy[m] = row + c + yo[m];
                        oz[m] = row + c + zo[m];
                    }
                    uy[u] = centred_at<P, MO>(pu, oy);
                    uz[u] = centred_at<P, MO>(pu, oz);
                }
            }
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int c = c0 + u * B;
                if (c >= nx) continue;
                if (MOM) {
                    ucen[c] = uy[u];
                    ucen[nx + c] = uz[u];
                }
#pragma unroll
                for (int f = 0; f < NF; ++f)
                    if (f >= f0) told[(f - f0) * nx + c] = tv[u][f];
            }
        }
    }
    __syncthreads();

    // phase 2
    const double rh = pr.rho0_half[k];
    const double r0 = pr.rho0[k];
    const double ah = pr.alpha0_half[k];
    const double a0 = pr.alpha0[k];
    const double dxi = g.dxi[0];
    const TC *urow = rows + NSC * W + H;  // urow[i] = u at cell i (wrapped halo)
    double bw[NSC > 0 ? NSC : 1][1];
#pragma unroll
    for (int sc = 0; sc < NSC; ++sc) bw[sc][0] = P::kBase ? (fl.base[sc] ? fl.base[sc][k] : 0.0) : 0.0;
    for (int c = tid; c < nx; c += B) {
#pragma unroll
        for (int sc = 0; sc < NSC; ++sc) {
            const TC *srow = rows + sc * W + H;
            TC win[2 * RS];
#pragma unroll
            for (int m = 0; m < 2 * RS; ++m) win[m] = srow[c - (RS - 1) + m];
            flux[sc * nx + c] = P::x_keep(P::template scal_flux<SO, false>(win, bw[sc], urow[c], rh));
        }
        if (MOM) {
#pragma unroll
            for (int a = 0; a < 3; ++a) {  // advected component u, v, w
                const TC *prow = rows + (NSC + a) * W + H;
                TC win[2 * RM];
#pragma unroll
                for (int m = 0; m < 2 * RM; ++m) win[m] = prow[c - (RM - 1) + m];
                TC U;
                if (a == 0) {
                    TC q[NC];
#pragma unroll
                    for (int m = 0; m < NC; ++m) q[m] = urow[c + m - (RM - 2)];
                    U = Interp<MO, TC>::centred(q);
                } else {
                    U = ucen[(a - 1) * nx + c];
                }
                flux[(NSC + a) * nx + c] = P::template mom_flux<MO>(win, U, a == 2 ? r0 : rh);
            }
        }
    }
    __syncthreads();

    // phase 3
    for (int c = tid; c < nx; c += B) {
        const int cm = c == 0 ? nx - 1 : c - 1;
#pragma unroll
        for (int f = 0; f < NF; ++f) {
            TS d;
            if (f < NSC) {
                d = P::x_scal_div(flux[f * nx + c], flux[f * nx + cm], urow[c], urow[cm], ah, dxi, bw[f][0] * rh);
            } else {
                d = P::mom_div(flux[f * nx + c], flux[f * nx + cm], f == NSC + 2 ? a0 : ah, dxi);
            }
            // 0 - d when the tendency is zero on entry (same bits as adding to 0)
            const int ti = f - (NF - nt);
            const TS old = ti >= 0 ? (TS)told[ti * nx + c] : TS(0);
            *at((TS *)fl.t[field(f)], row + c) = old - d;
        }
    }
}

template <class P, int SO, int MO, int NSC, bool MOM>
size_t x_smem_bytes(int nx, bool tzero) {
    return XLayout<P, SO, MO, NSC, MOM>::words(nx, tzero) * sizeof(typename P::TC);
}

// Launch the three passes in the reference's per-variable order (x, y, z).
template <class P, int SO, int MO, int NSC, bool MOM>
int launch(const Geo &g, const Prof &pr, const Fields &fl, hipStream_t st) {
    const size_t smem = x_smem_bytes<P, SO, MO, NSC, MOM>(g.nx, fl.tzero != 0);
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
    if (fl.passes & 1) {
        hipLaunchKernelGGL((adv_x<P, SO, MO, NSC, MOM>), dim3(g.ny, g.nz), dim3(bx), smem, st, g, pr, fl);
        PYCLES_CHECK_LAUNCH();
    }
    // y and z fused (WENO5 with momentum and scalars), z-chunked like the z pass
    if constexpr (SO == 5 && MO == 5 && MOM && NSC >= 1) {
        if ((fl.passes & 6) == 6 && (P::kFuseYZ || fl.fuse_yz)) {
            // 4 rows per block: 8 rows halve the lower faces' extra work but
            // need more registers than 4 waves per SIMD allow with the RK3
            // stage (spills), and at 3 waves a 512-thread block fits once per CU
            constexpr int TY = 4;
            const ZSplit zs = advyz_zsplit(g.nx, g.ny, g.nz);
            const dim3 grid((g.nx + 63) / 64, (g.ny + TY - 1) / TY, zs.nch), blk(64, TY);
            const bool part = g.nx % 64 != 0 || g.ny % TY != 0;
            const int rkm = fl.rk < 0 ? 0 : (fl.rk == 0 ? 1 : 2);
#define ADV_YZ(RKM_, PART_) hipLaunchKernelGGL((adv_yz<P, SO, MO, NSC, MOM, RKM_, TY, PART_>), grid, blk, 0, st, g, pr, fl, zs.len)
            if (part) {
                if (rkm == 0) ADV_YZ(0, true);
                else if (rkm == 1) ADV_YZ(1, true);
                else ADV_YZ(2, true);
            } else {
                if (rkm == 0) ADV_YZ(0, false);
                else if (rkm == 1) ADV_YZ(1, false);
                else ADV_YZ(2, false);
            }
#undef ADV_YZ
            PYCLES_CHECK_LAUNCH();
            return PYCLES_OK;
        }
    }
    // y: split the march into chunks so the grid holds enough waves.
    const ZSplit ys = advy_zsplit(g.nx, g.ny, g.nz);
    const int ych = ys.nch, jchunk = ys.len;
    if (fl.passes & 2) {
        hipLaunchKernelGGL((adv_y<P, SO, MO, NSC, MOM>), dim3((g.nx + 63) / 64, (g.nz + 3) / 4, ych),
                           dim3(64, 4), 0, st, g, pr, fl, jchunk);
        PYCLES_CHECK_LAUNCH();
    }
    const ZSplit zs = advz_zsplit(g.nx, g.ny, g.nz);
    const int zch = zs.nch, kchunk = zs.len;
    if (fl.passes & 4) {
        const dim3 zg((g.nx + 63) / 64, (g.ny + 3) / 4, zch);
        if (fl.rk < 0)
            hipLaunchKernelGGL((adv_z<P, SO, MO, NSC, MOM, 0>), zg, dim3(64, 4), 0, st, g, pr, fl, kchunk);
        else if (fl.rk == 0)
            hipLaunchKernelGGL((adv_z<P, SO, MO, NSC, MOM, 1>), zg, dim3(64, 4), 0, st, g, pr, fl, kchunk);
        else
            hipLaunchKernelGGL((adv_z<P, SO, MO, NSC, MOM, 2>), zg, dim3(64, 4), 0, st, g, pr, fl, kchunk);
        PYCLES_CHECK_LAUNCH();
    }
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
