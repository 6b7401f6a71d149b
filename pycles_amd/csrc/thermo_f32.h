// FP32 saturation adjustment with an FP64 residual check.
//
// The entropy equation s = s(T; p0, qt) of eos_c (thermodynamics_sa.h:17-66)
// is solved in the reduced form
//     s_red = cpm ln(T/T~) - (1-qt) Rd ln(pd/p~) - qt Rv ln(pv/p~) - L sigma / T,
//     s_red = s - (1-qt) sd~ - qt sv~,
// whose terms are O(100) J/kg/K instead of O(7000): s_red is formed in FP64
// from the FP64 level profiles of s and qt plus the FP32 deviations, and the
// remaining arithmetic is FP32 without catastrophic cancellation.  The
// iteration is the reference's secant (same start T_1, same first step T_2,
// same stopping test |dT| < 1e-3 K and sigma >= 0, same lagged qv).  The
// result is then checked in FP64: if |s - s(T)| exceeds SA_RESID_TOL the cell
// is recomputed with the FP64 secant of thermo_funcs.h.
#pragma once
#include "thermo_funcs.h"

// |s - s(T)| in J/kg/K.  ds/dT on the saturated branch is ~10 J/kg/K^2
// (cpm/T plus the latent-heat term), so the bound keeps T within ~3e-4 K
// (1e-6 relative) of the root, below the reference secant's own 1e-3 K
// stopping test.
#define SA_RESID_TOL 3e-3

namespace thermo32 {

// Hardware reciprocal, log2 and exp2 (about 1 ulp each) instead of the
// correctly rounded library sequences: every saturated FP32 result is
// checked against the FP64 residual, and the unsaturated inversion needs
// ~1e-7 relative, far inside the FP32 form's 1e-5 target.
__device__ __forceinline__ float fdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ float frcp(float b) { return __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ float flog(float x) { return __builtin_amdgcn_logf(x) * 0.69314718055994531f; }
__device__ __forceinline__ float fexp(float x) { return __builtin_amdgcn_exp2f(x * 1.4426950408889634f); }

struct TableF {
    const float *y;
    float x_min, dx, dxi;
    int n;
};

__device__ __forceinline__ float lookup(const TableF &t, float x) {
    float fi = floorf((x - t.x_min) * t.dxi);
    fi = fi >= 0.0f ? fi : 0.0f;
    fi = fi <= (float)(t.n - 2) ? fi : (float)(t.n - 2);
    const int indx = (int)fi;
    const float y1 = t.y[indx];
    const float y2 = t.y[indx + 1];
    const float x1 = t.x_min + t.dx * (float)indx;
    return y1 + (x - x1) * (y2 - y1) * t.dxi;
}

__device__ __forceinline__ float latent_heat(float T) {
    const float TC = T - 273.15f;
    return (2500.8f - 2.36f * TC + 0.0016f * TC * TC - 0.00006f * TC * TC * TC) * 1000.0f;
}

__device__ __forceinline__ float pv_c(float p0, float qt, float qv) {
    const float e = (float)PC_EPS_VI;
    return fdiv(p0 * e * qv, 1.0f - qt + e * qv);
}

__device__ __forceinline__ float qv_star_c(float p0, float qt, float pv) {
    return fdiv((float)PC_EPS_V * (1.0f - qt) * pv, p0 - pv);
}

// reduced entropy of the saturated state at T (the terms of s(T) that are
// not the constants (1-qt) sd~ + qt sv~)
__device__ __forceinline__ float g_red(float T, float qt, float pd, float pv, float L, float sigma, float cpm) {
    const float lnT = flog(T * (float)(1.0 / PC_T_TILDE));
    return cpm * lnT - (1.0f - qt) * (float)PC_RD * flog(pd * (float)(1.0 / PC_P_TILDE)) -
           qt * (float)PC_RV * flog(pv * (float)(1.0 / PC_P_TILDE)) - fdiv(L * sigma, T);
}

// FP64 arithmetic of the residual check, without the library sequences'
// handling of special cases (every argument here is a positive normal
// number).  div64 (common.h): the hardware reciprocal, two Newton steps and
// one correction of the quotient.  log64: x = m 2^e with m in [sqrt(1/2),
// sqrt(2)), ln x = e ln 2 + 2 atanh(t), t = (m - 1)/(m + 1), the series to
// t^17; within 1e-15 absolute of the C library's log on [6e-6, 20].  The
// check resolves 3e-3 J/kg/K on terms of O(100) J/kg/K.
__device__ __forceinline__ double log64(double x) {
    double m = __builtin_amdgcn_frexp_mant(x);  // [0.5, 1)
    int e = __builtin_amdgcn_frexp_exp(x);
    if (m < 0.70710678118654752) {
        m *= 2.0;
        e -= 1;
    }
    const double t = div64(m - 1.0, m + 1.0);
    const double t2 = t * t;
    double p = 1.0 / 17;
    p = __builtin_fma(p, t2, 1.0 / 15);
    p = __builtin_fma(p, t2, 1.0 / 13);
    p = __builtin_fma(p, t2, 1.0 / 11);
    p = __builtin_fma(p, t2, 1.0 / 9);
    p = __builtin_fma(p, t2, 1.0 / 7);
    p = __builtin_fma(p, t2, 1.0 / 5);
    p = __builtin_fma(p, t2, 1.0 / 3);
    p = __builtin_fma(p, t2, 1.0);
    return __builtin_fma((double)e, 0.69314718055994531, 2.0 * t * p);
}

// FP64 residual s - s(T) of the saturated state reached by the iteration.
__device__ __forceinline__ double resid64(double s_red, double p0, double qt, double T, double qv_star) {
    const double pv = div64(p0 * PC_EPS_VI * qv_star, 1.0 - qt + PC_EPS_VI * qv_star);
    const double pd = p0 - pv;
    const double sigma = qt - qv_star;
    const double L = thermo::latent_heat_variable(T);
    const double cpm = (1.0 - qt) * PC_CPD + qt * PC_CPV;
    const double g = cpm * log64(T * (1.0 / PC_T_TILDE)) - (1.0 - qt) * PC_RD * log64(pd * (1.0 / PC_P_TILDE)) -
                     qt * PC_RV * log64(pv * (1.0 / PC_P_TILDE)) - div64(L * sigma, T);
    return s_red - g;
}

// The FP64 secant for the rare rejected cell, kept out of line so its
// registers do not weigh on the FP32 path.  Arguments and results pass by
// value: a pointer to the caller's T, qv, ql would keep those in scratch
// memory for the whole level loop, with a full memory wait at every reload.
struct TQ {
    float T, qv, ql;
};
__device__ __noinline__ TQ fallback64(thermo::Table lt64, double p0, double s, double qt, int max_iter) {
    double T64, qv64, ql64, qi64;
    thermo::eos_sa(lt64, p0, s, qt, &T64, &qv64, &ql64, &qi64, max_iter);
    return TQ{(float)T64, (float)qv64, (float)ql64};
}

// State after the unsaturated test of eos_c (thermodynamics_sa.h:23-33).
struct SAStart {
    float s_red, cpm, T_1, qv_star_1, pd_1, pv_1;
    bool saturated;
};

__device__ __forceinline__ SAStart eos_sa_f32_start(const TableF &lt, float p0f, double s_red64, float qt) {
    SAStart st;
    st.s_red = (float)s_red64;
    st.cpm = (1.0f - qt) * (float)PC_CPD + qt * (float)PC_CPV;
    const float pv_1 = pv_c(p0f, qt, qt);
    const float pd_1 = p0f - pv_1;
    const float g_dry = -(1.0f - qt) * (float)PC_RD * flog(pd_1 * (float)(1.0 / PC_P_TILDE)) -
                        qt * (float)PC_RV * flog(pv_1 * (float)(1.0 / PC_P_TILDE));
    st.T_1 = (float)PC_T_TILDE * fexp(fdiv(st.s_red - g_dry, st.cpm));
    st.qv_star_1 = qv_star_c(p0f, qt, lookup(lt, st.T_1));
    st.saturated = qt > st.qv_star_1;
    st.pd_1 = pd_1;
    st.pv_1 = pv_1;
    return st;
}

// Newton correction of a first guess T on the saturated branch:
// f(T) = s_red - g_red(T) with qv = qv*(T), f'(T) = -dg_red/dT from the
// table slope of pv*(T) and the derivative of the latent-heat polynomial.
__device__ __forceinline__ float newton_step(const TableF &lt, float p0f, float qt, float s_red, float cpm, float T) {
    float fi = floorf((T - lt.x_min) * lt.dxi);
    fi = fi >= 0.0f ? fi : 0.0f;
    fi = fi <= (float)(lt.n - 2) ? fi : (float)(lt.n - 2);
    const int indx = (int)fi;
    const float y1 = lt.y[indx], y2 = lt.y[indx + 1];
    const float x1 = lt.x_min + lt.dx * (float)indx;
    const float dpv = (y2 - y1) * lt.dxi;
    const float pv_star = y1 + (T - x1) * dpv;
    const float qv_star = qv_star_c(p0f, qt, pv_star);
    const float pv = pv_c(p0f, qt, qv_star);
    const float pd = p0f - pv;
    const float sigma = qt - qv_star;
    const float L = latent_heat(T);
    const float f = s_red - g_red(T, qt, pd, pv, L, sigma, cpm);
    const float TC = T - 273.15f;
    const float dL = (-2.36f + 0.0032f * TC - 0.00018f * TC * TC) * 1000.0f;
    const float iT = frcp(T);
    const float dqv = fdiv((float)PC_EPS_V * (1.0f - qt) * p0f * dpv, (p0f - pv_star) * (p0f - pv_star));
    const float dg = cpm * iT + (fdiv((1.0f - qt) * (float)PC_RD, pd) - fdiv(qt * (float)PC_RV, pv)) * dpv -
                     (dL * sigma - L * dqv) * iT + L * sigma * iT * iT;
    return T + fdiv(f, dg);
}

// FP64 residual |s - s(T)| of a saturated FP32 result is within SA_RESID_TOL.
__device__ __forceinline__ bool passes_check(const thermo::Table &lt64, double p0, double s_red64, double qt64,
                                             float T) {
    const double Td = (double)T;
    const double pvs = thermo::lookup(lt64, Td);
    const double qvs = div64(PC_EPS_V * (1.0 - qt64) * pvs, p0 - pvs);
    return fabs(resid64(s_red64, p0, qt64, Td, qvs)) <= SA_RESID_TOL;
}

// The reference's secant in FP32 (same first step, same stopping test
// |dT| < 1e-3 K and sigma >= 0, same lagged qv).  Returns its iterations.
__device__ __forceinline__ int secant_f32(const TableF &lt, const SAStart &st, float p0f, float qt, int max_iter,
                                          float *T, float *qv, float *ql) {
    // the secant's first point: f at T_1 with the unsaturated partial
    // pressures, and the first step, as the reference does
    const float sigma_1 = qt - st.qv_star_1;
    const float L_1 = latent_heat(st.T_1);
    float f_1 = st.s_red - g_red(st.T_1, qt, st.pd_1, st.pv_1, L_1, sigma_1, st.cpm);
    float T_1 = st.T_1;
    float T_2 = st.T_1 + fdiv(sigma_1 * L_1, (1.0f - qt) * (float)PC_CPD + st.qv_star_1 * (float)PC_CPV);
    float delta_T, qv_star_2, sigma_2;
    int it = 0;
    do {
        qv_star_2 = qv_star_c(p0f, qt, lookup(lt, T_2));
        const float pv_2 = pv_c(p0f, qt, qv_star_2);
        const float pd_2 = p0f - pv_2;
        sigma_2 = qt - qv_star_2;
        const float L_2 = latent_heat(T_2);
        const float f_2 = st.s_red - g_red(T_2, qt, pd_2, pv_2, L_2, sigma_2, st.cpm);
        const float den = f_2 - f_1;
        const float T_n = den != 0.0f ? T_2 - fdiv(f_2 * (T_2 - T_1), den) : T_2;
        T_1 = T_2;
        T_2 = T_n;
        f_1 = f_2;
        delta_T = fabsf(T_2 - T_1);
        ++it;
        if (max_iter > 0 && it >= max_iter) break;
    } while (delta_T >= 1.0e-3f || sigma_2 < 0.0f);
    *T = T_2;
    *qv = qv_star_2;
    *ql = sigma_2;
    return it;
}

// Saturated branch.  T_guess > 0: `newton` Newton corrections of the guess,
// then qv = qv*(T); a result that fails the FP64 residual check is redone by
// the FP32 secant.  T_guess <= 0: the FP32 secant.  A secant result that
// fails the check is recomputed by the FP64 secant.  Returns the iteration
// count (Newton corrections plus secant steps); *fallback is set when the
// first result was not kept.
__device__ __forceinline__ int eos_sa_f32_finish(const TableF &lt, const thermo::Table &lt64, const SAStart &st,
                                                 float p0f, double p0, double s, double s_red64, float qt,
                                                 double qt64, float T_guess, int newton, float *T, float *qv,
                                                 float *ql, int max_iter, int *fallback) {
    int it = 0;
    *fallback = 0;
    if (T_guess > 0.0f) {
        float Tn = T_guess;
        for (; it < newton; ++it) Tn = newton_step(lt, p0f, qt, st.s_red, st.cpm, Tn);
        if (passes_check(lt64, p0, s_red64, qt64, Tn)) {
            const float qvs = qv_star_c(p0f, qt, lookup(lt, Tn));
            *T = Tn;
            *qv = qvs;
            *ql = qt - qvs;
            return it;
        }
        *fallback = 1;
    }
    it += secant_f32(lt, st, p0f, qt, max_iter, T, qv, ql);
    if (!passes_check(lt64, p0, s_red64, qt64, *T)) {
        const TQ r = fallback64(lt64, p0, s, qt64, max_iter);
        *T = r.T;
        *qv = r.qv;
        *ql = r.ql;
        *fallback = 1;
    }
    return it;
}

}  // namespace thermo32
