// FP64 device thermodynamic functions.  Expression trees follow the cited
// reference lines (thermodynamic_functions.h, entropies.h,
// thermodynamics_sa.h, thermodynamics_dry.h, lookup.h, Microphysics.pxd) so
// that, without contraction, only the transcendental functions (device libm
// instead of glibc) differ from a CPU build.
#pragma once
#include "common.h"

namespace thermo {

struct Table {
    const double *y;  // LDS copy
    double x_min, dx, dxi;
    int n;
};

// lookup.h:48-54.  The index is clamped to the table (the reference reads
// out of bounds for temperatures outside it); in-range values are unchanged.
__device__ __forceinline__ double lookup(const Table &t, double x) {
    double fi = floor((x - t.x_min) * t.dxi);
    fi = fi >= 0.0 ? fi : 0.0;
    fi = fi <= (double)(t.n - 2) ? fi : (double)(t.n - 2);
    const long long indx = (long long)fi;
    const double y1 = t.y[indx];
    const double y2 = t.y[indx + 1];
    const double x1 = t.x_min + t.dx * indx;
    return y1 + (x - x1) * (y2 - y1) * t.dxi;
}

// Microphysics.pxd:71-75
__device__ __forceinline__ double latent_heat_variable(double T) {
    const double TC = T - 273.15;
    return (2500.8 - 2.36 * TC + 0.0016 * TC * TC - 0.00006 * TC * TC * TC) * 1000.0;
}

// The functions with divisions take a division policy D: DivIeee (the IEEE
// division, the default) or DivFast (div64 in common.h, the same quotient for
// these operands: pressures, temperatures, 1 - qt and sums of positive terms
// as divisors, in fewer instructions).  The FP64 thermodynamics kernel uses
// DivFast; the FP32 kernels' out-of-line FP64 fallback keeps DivIeee, whose
// register use their allocation was tuned with.  The secant step's
// f_2 - f_1 is an IEEE division under both.
struct DivIeee {
    static __device__ __forceinline__ double div(double a, double b) { return a / b; }
};
struct DivFast {
    static __device__ __forceinline__ double div(double a, double b) { return div64(a, b); }
};

// thermodynamic_functions.h
__device__ __forceinline__ double exner_c(double p0) { return pow((p0 / PC_P_TILDE), PC_KAPPA); }
template <class D = DivIeee> __device__ __forceinline__ double pd_c(double p0, double qt, double qv) {
    return D::div(p0 * (1.0 - qt), 1.0 - qt + PC_EPS_VI * qv);
}
template <class D = DivIeee> __device__ __forceinline__ double pv_c(double p0, double qt, double qv) {
    return D::div(p0 * PC_EPS_VI * qv, 1.0 - qt + PC_EPS_VI * qv);
}
__device__ __forceinline__ double cpm_c(double qt) { return (1.0 - qt) * PC_CPD + qt * PC_CPV; }
template <class D = DivIeee> __device__ __forceinline__ double qv_star_c(double p0, double qt, double pv) {
    return D::div(PC_EPS_V * (1.0 - qt) * pv, p0 - pv);
}
template <class D = DivIeee> __device__ __forceinline__ double alpha_c(double p0, double T, double qt, double qv) {
    return D::div(PC_RD * T, p0) * (1.0 - qt + PC_EPS_VI * qv);
}
template <class D = DivIeee> __device__ __forceinline__ double buoyancy_c(double alpha0, double alpha) {
    return D::div(PC_G * (alpha - alpha0), alpha0);
}
__device__ __forceinline__ double theta_rho_c(double p0, double T, double qt, double qv) {
    return T * (1.0 - qt + PC_EPS_VI * qv) / exner_c(p0);
}
__device__ __forceinline__ double thetali_c(double p0, double T, double qt, double ql, double qi, double L) {
    return T / exner_c(p0) * exp(-L * (ql / (1.0 - qt) + qi / (1.0 - qt)) / (T * PC_CPD));
}

// entropies.h:5-15
template <class D = DivIeee> __device__ __forceinline__ double sd_c(double pd, double T) {
    return PC_SD_TILDE + PC_CPD * log(D::div(T, PC_T_TILDE)) - PC_RD * log(D::div(pd, PC_P_TILDE));
}
template <class D = DivIeee> __device__ __forceinline__ double sv_c(double pv, double T) {
    return PC_SV_TILDE + PC_CPV * log(D::div(T, PC_T_TILDE)) - PC_RV * log(D::div(pv, PC_P_TILDE));
}
template <class D = DivIeee> __device__ __forceinline__ double sc_c(double L, double T) { return D::div(-L, T); }

// thermodynamics_sa.h:10-15
template <class D = DivIeee>
__device__ __forceinline__ double temperature_no_ql(double pd, double pv, double s, double qt) {
    return PC_T_TILDE * exp(D::div(s - (1.0 - qt) * (PC_SD_TILDE - PC_RD * log(D::div(pd, PC_P_TILDE))) -
                                       qt * (PC_SV_TILDE - PC_RV * log(D::div(pv, PC_P_TILDE))),
                                   (1.0 - qt) * PC_CPD + qt * PC_CPV));
}

// eos_c, thermodynamics_sa.h:17-66, lambda == 1; capped at max_iter secant
// iterations (max_iter <= 0: uncapped as in the reference); the stopping test
// is |dT| < tol (the reference's 1e-3 K by default).  Returns the number of
// secant iterations (0 = unsaturated closed form).
template <class D = DivIeee>
__device__ __forceinline__ int eos_sa(const Table &lt, double p0, double s, double qt, double *T, double *qv,
                                      double *ql, double *qi, int max_iter, double tol = 1.0e-3) {
    *qv = qt;
    *ql = 0.0;
    *qi = 0.0;
    const double pv_1 = pv_c<D>(p0, qt, qt);
    const double pd_1 = p0 - pv_1;
    double T_1 = temperature_no_ql<D>(pd_1, pv_1, s, qt);
    const double pv_star_1 = lookup(lt, T_1);
    const double qv_star_1 = qv_star_c<D>(p0, qt, pv_star_1);
    if (qt <= qv_star_1) {
        *T = T_1;
        return 0;
    }
    const double sigma_1 = qt - qv_star_1;
    const double L_1 = latent_heat_variable(T_1);
    const double s_1 = sd_c<D>(pd_1, T_1) * (1.0 - qt) + sv_c<D>(pv_1, T_1) * qt + sc_c<D>(L_1, T_1) * sigma_1;
    double f_1 = s - s_1;
    double T_2 = T_1 + D::div(sigma_1 * L_1, (1.0 - qt) * PC_CPD + qv_star_1 * PC_CPV);
    double delta_T, qv_star_2, sigma_2;
    int it = 0;
    do {
        const double pv_star_2 = lookup(lt, T_2);
        qv_star_2 = qv_star_c<D>(p0, qt, pv_star_2);
        const double pv_2 = pv_c<D>(p0, qt, qv_star_2);
        const double pd_2 = p0 - pv_2;
        sigma_2 = qt - qv_star_2;
        const double L_2 = latent_heat_variable(T_2);
        const double s_2 = sd_c<D>(pd_2, T_2) * (1.0 - qt) + sv_c<D>(pv_2, T_2) * qt + sc_c<D>(L_2, T_2) * sigma_2;
        const double f_2 = s - s_2;
        const double T_n = T_2 - f_2 * (T_2 - T_1) / (f_2 - f_1);
        T_1 = T_2;
        T_2 = T_n;
        f_1 = f_2;
        delta_T = fabs(T_2 - T_1);
        ++it;
        if (max_iter > 0 && it >= max_iter) break;
    } while (delta_T >= tol || sigma_2 < 0.0);
    *T = T_2;
    *qv = qv_star_2;
    *ql = 1.0 * sigma_2;
    *qi = (1.0 - 1.0) * sigma_2;
    return it;
}

// bvf_sa inner expression (thermodynamics_sa.h:155-178) for the middle level
// of the three-level windows w[0] = k-1, w[1] = k, w[2] = k+1.
template <class D = DivIeee>
__device__ __forceinline__ double bvf_sa(const Table &lt, double p0, const double *T, const double *qt,
                                         const double *qv, const double *thr, double dzi) {
    if (qv[1] < qt[1]) {
        const double Tk = T[1];
        const double Lv = latent_heat_variable(Tk);
        const double pv_star = lookup(lt, Tk);
        const double rsl = D::div(PC_EPS_V * pv_star, p0 - pv_star);
        const double gamma_w = D::div(PC_G / PC_CPD * D::div(1.0, 1.0 - qt[1]) * (1.0 + D::div(Lv * rsl, PC_RD * Tk)),
                                      D::div(cpm_c(qt[1]), PC_CPD) +
                                          D::div(Lv * Lv * (PC_EPS_V + rsl) * rsl, PC_CPD * PC_RD * Tk * Tk));
        const double dTdz = (0.5 * (T[1] + T[2]) - 0.5 * (T[0] + T[1])) * dzi;
        const double dqtdz = (0.5 * (qt[1] + qt[2]) - 0.5 * (qt[0] + qt[1])) * dzi;
        return D::div(PC_G, Tk) * (dTdz + gamma_w) * (1.0 + D::div(Lv * rsl, PC_RD * Tk)) - D::div(dqtdz, 1.0 - qt[1]);
    }
    return D::div(PC_G, thr[1]) * (0.5 * (thr[1] + thr[2]) - 0.5 * (thr[0] + thr[1])) * dzi;
}

// thermodynamics_dry.h:8-10
template <class D = DivIeee> __device__ __forceinline__ double eos_dry(double pd, double s) {
    return PC_T_TILDE * (exp(D::div(s - PC_SD_TILDE + PC_RD * log(D::div(pd, PC_P_TILDE)), PC_CPD)));
}

}  // namespace thermo
