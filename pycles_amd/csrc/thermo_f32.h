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

#define SA_RESID_TOL 2e-3

namespace thermo32 {

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
    return p0 * e * qv / (1.0f - qt + e * qv);
}

__device__ __forceinline__ float qv_star_c(float p0, float qt, float pv) {
    return (float)PC_EPS_V * (1.0f - qt) * pv / (p0 - pv);
}

// reduced entropy of the saturated state at T (the terms of s(T) that are
// not the constants (1-qt) sd~ + qt sv~)
__device__ __forceinline__ float g_red(float T, float qt, float pd, float pv, float L, float sigma, float cpm) {
    const float lnT = logf(T * (float)(1.0 / PC_T_TILDE));
    return cpm * lnT - (1.0f - qt) * (float)PC_RD * logf(pd * (float)(1.0 / PC_P_TILDE)) -
           qt * (float)PC_RV * logf(pv * (float)(1.0 / PC_P_TILDE)) - L * sigma / T;
}

// FP64 residual s - s(T) of the saturated state reached by the iteration.
__device__ __forceinline__ double resid64(double s_red, double p0, double qt, double T, double qv_star) {
    const double pv = thermo::pv_c(p0, qt, qv_star);
    const double pd = p0 - pv;
    const double sigma = qt - qv_star;
    const double L = thermo::latent_heat_variable(T);
    const double cpm = (1.0 - qt) * PC_CPD + qt * PC_CPV;
    const double g = cpm * log(T / PC_T_TILDE) - (1.0 - qt) * PC_RD * log(pd / PC_P_TILDE) -
                     qt * PC_RV * log(pv / PC_P_TILDE) - L * sigma / T;
    return s_red - g;
}

// The FP64 secant for the rare rejected cell, kept out of line so its
// registers do not weigh on the FP32 path.
__device__ __noinline__ void fallback64(const thermo::Table &lt64, double p0, double s, double qt, int max_iter,
                                        float *T, float *qv, float *ql) {
    double T64, qv64, ql64, qi64;
    thermo::eos_sa(lt64, p0, s, qt, &T64, &qv64, &ql64, &qi64, max_iter);
    *T = (float)T64;
    *qv = (float)qv64;
    *ql = (float)ql64;
}

// Returns secant iterations (0 = unsaturated); *T, *qv, *ql outputs.
// *fallback set when the FP64 check rejected the FP32 result.
__device__ __forceinline__ int eos_sa_f32(const TableF &lt, const thermo::Table &lt64, float p0f, double p0,
                                          double s, double s_red64, float qt, double qt64, float T_guess,
                                          float *T, float *qv, float *ql, int max_iter, int *fallback) {
    *fallback = 0;
    const float s_red = (float)s_red64;
    const float cpm = (1.0f - qt) * (float)PC_CPD + qt * (float)PC_CPV;
    const float pv_1 = pv_c(p0f, qt, qt);
    const float pd_1 = p0f - pv_1;
    const float g_dry = -(1.0f - qt) * (float)PC_RD * logf(pd_1 * (float)(1.0 / PC_P_TILDE)) -
                        qt * (float)PC_RV * logf(pv_1 * (float)(1.0 / PC_P_TILDE));
    float T_1 = (float)PC_T_TILDE * expf((s_red - g_dry) / cpm);
    const float qv_star_1 = qv_star_c(p0f, qt, lookup(lt, T_1));
    if (qt <= qv_star_1) {
        *T = T_1;
        *qv = qt;
        *ql = 0.0f;
        return 0;
    }
    const float sigma_1 = qt - qv_star_1;
    const float L_1 = latent_heat(T_1);
    // f at T_1 with the unsaturated partial pressures, as the reference does
    float f_1 = s_red - g_red(T_1, qt, pd_1, pv_1, L_1, sigma_1, cpm);
    float T_2 = T_1 + sigma_1 * L_1 / ((1.0f - qt) * (float)PC_CPD + qv_star_1 * (float)PC_CPV);
    if (T_guess > 0.0f) T_2 = T_guess;
    float delta_T, qv_star_2, sigma_2;
    int it = 0;
    do {
        qv_star_2 = qv_star_c(p0f, qt, lookup(lt, T_2));
        const float pv_2 = pv_c(p0f, qt, qv_star_2);
        const float pd_2 = p0f - pv_2;
        sigma_2 = qt - qv_star_2;
        const float L_2 = latent_heat(T_2);
        const float f_2 = s_red - g_red(T_2, qt, pd_2, pv_2, L_2, sigma_2, cpm);
        const float den = f_2 - f_1;
        const float T_n = den != 0.0f ? T_2 - f_2 * (T_2 - T_1) / den : T_2;
        T_1 = T_2;
        T_2 = T_n;
        f_1 = f_2;
        delta_T = fabsf(T_2 - T_1);
        ++it;
        if (max_iter > 0 && it >= max_iter) break;
    } while (delta_T >= 1.0e-3f || sigma_2 < 0.0f);
    // FP64 check of the returned state
    const double Td = (double)T_2;
    const double qvs = thermo::qv_star_c(p0, qt64, thermo::lookup(lt64, Td));
    const double r = resid64(s_red64, p0, qt64, Td, qvs);
    if (!(fabs(r) <= SA_RESID_TOL)) {
        fallback64(lt64, p0, s, qt64, max_iter, T, qv, ql);
        *fallback = 1;
        return it;
    }
    *T = T_2;
    *qv = qv_star_2;
    *ql = sigma_2;
    return it;
}

}  // namespace thermo32
