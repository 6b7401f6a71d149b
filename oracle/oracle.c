/*
 * oracle.c -- FP64 CPU restatement of the PyCLES dynamical core hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker): see oracle.h.  Built with
 * -O2 -ffp-contract=off so that, like the reference built with -std=c99,
 * no multiply-add is contracted.  Loops over i are OpenMP-parallel; every
 * cell is computed independently, so results do not depend on the thread
 * count.  Reductions (horizontal mean, CFL) run serially in reference order.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Physical constants restated from generate_parameters.py:20-39.  Derived
 * constants are evaluated with the same IEEE double operations Python does. */
#define C_PI 3.141592653589793
#define C_G 9.80665
#define C_RD 287.1
#define C_RV 461.5
#define C_EPS_V (C_RD / C_RV)
#define C_EPS_VI (1.0 / C_EPS_V)
#define C_CPD 1004.0
#define C_CPV 1859.0
#define C_KAPPA (C_RD / C_CPD)
#define C_T_TILDE 298.15
#define C_P_TILDE 100000.0
#define C_SD_TILDE 6864.8
#define C_SV_TILDE 10513.6

void orc_dims_init(orc_dims *d, long nx, long ny, long nz, long gw, double dx, double dy,
                   double dz) {
    d->n[0] = nx; d->n[1] = ny; d->n[2] = nz;
    d->gw = gw;
    for (int i = 0; i < 3; ++i) d->nlg[i] = d->n[i] + 2 * gw;
    d->npg = d->nlg[0] * d->nlg[1] * d->nlg[2];
    d->dx[0] = dx; d->dx[1] = dy; d->dx[2] = dz;
    /* Grid.pyx:28-31 */
    for (int i = 0; i < 3; ++i) d->dxi[i] = 1.0 / d->dx[i];
}

/* ------------------------------------------------------------------------ */
/* Interpolation: advection_interpolation.h.  Coefficient products and sums
 * are written in the reference's evaluation order.                          */

double orc_interp_2(double a, double b) { return 0.5 * (a + b); } /* :4-6 */

double orc_interp_4(const double *q) { /* :12-14 */
    return (7.0 / 12.0) * (q[1] + q[2]) - (1.0 / 12.0) * (q[0] + q[3]);
}

double orc_interp_6(const double *q) { /* :20-24 */
    return ((37.0 / 60.0) * (q[2] + q[3]) - (2.0 / 15.0) * (q[1] + q[4]) +
            (1.0 / 60.0) * (q[0] + q[5]));
}

double orc_interp_8(const double *q) { /* :32-36 */
    return (533. / 840. * (q[3] + q[4]) - 139.0 / 840.0 * (q[2] + q[5]) +
            29.0 / 840.0 * (q[1] + q[6]) - 1.0 / 280.0 * (q[0] + q[7]));
}

static inline double sq_eps(double b) { return (b + 1e-10) * (b + 1e-10); }

double orc_interp_weno5(const double *q) { /* :79-104 */
    const double m2 = q[0], m1 = q[1], c = q[2], p1 = q[3], p2 = q[4];
    const double s0 = (1.0 / 3.0) * m2 - (7.0 / 6.0) * m1 + (11.0 / 6.0) * c;
    const double s1 = (-1.0 / 6.0) * m1 + (5.0 / 6.0) * c + (1.0 / 3.0) * p1;
    const double s2 = (1.0 / 3.0) * c + (5.0 / 6.0) * p1 - (1.0 / 6.0) * p2;
    const double b2 = (13.0 / 12.0 * (c - 2.0 * p1 + p2) * (c - 2.0 * p1 + p2) +
                       0.25 * (3.0 * c - 4.0 * p1 + p2) * (3.0 * c - 4.0 * p1 + p2));
    const double b1 = (13.0 / 12.0 * (m1 - 2.0 * c + p1) * (m1 - 2.0 * c + p1) +
                       0.25 * (m1 - p1) * (m1 - p1));
    const double b0 = (13.0 / 12.0 * (m2 - 2.0 * m1 + c) * (m2 - 2.0 * m1 + c) +
                       0.25 * (m2 - 4.0 * m1 + 3.0 * c) * (m2 - 4.0 * m1 + 3.0 * c));
    const double a0 = 0.1 / sq_eps(b0);
    const double a1 = 0.6 / sq_eps(b1);
    const double a2 = 0.3 / sq_eps(b2);
    const double inv = 1.0 / (a0 + a1 + a2);
    const double w0 = a0 * inv, w1 = a1 * inv, w2 = a2 * inv;
    return w0 * s0 + w1 * s1 + w2 * s2;
}

double orc_interp_weno7(const double *q) { /* :106-145 */
    const double m3 = q[0], m2 = q[1], m1 = q[2], c = q[3], p1 = q[4], p2 = q[5], p3 = q[6];
    const double s0 = (-1.0 / 4.0) * m3 + (13.0 / 12.0) * m2 + (-23.0 / 12.0) * m1 + (25.0 / 12.0) * c;
    const double s1 = (1.0 / 12.0) * m2 + (-5.0 / 12.0) * m1 + (13.0 / 12.0) * c + (1.0 / 4.0) * p1;
    const double s2 = (-1.0 / 12.0) * m1 + (7.0 / 12.0) * c + (7.0 / 12.0) * p1 + (-1.0 / 12.0) * p2;
    const double s3 = (1.0 / 4.0) * c + (13.0 / 12.0) * p1 + (-5.0 / 12.0) * p2 + (1.0 / 12.0) * p3;
    const double b0 = (m3 * (547.0 * m3 - 3882.0 * m2 + 4642.0 * m1 - 1854.0 * c) +
                       m2 * (7043.0 * m2 - 17246.0 * m1 + 7042.0 * c) +
                       m1 * (11003.0 * m1 - 9402.0 * c) + 2107.0 * c * c);
    const double b1 = (m2 * (267.0 * m2 - 1642.0 * m1 + 1602.0 * c - 494.0 * p1) +
                       m1 * (2843.0 * m1 - 5966.0 * c + 1922.0 * p1) +
                       c * (3443.0 * c - 2522.0 * p1) + 547.0 * p1 * p1);
    const double b2 = (m1 * (547.0 * m1 - 2522.0 * c + 1922.0 * p1 - 494.0 * p2) +
                       c * (3443.0 * c - 5966.0 * p1 + 1602.0 * p2) +
                       p1 * (2843.0 * p1 - 1642.0 * p2) + 267.0 * p2 * p2);
    const double b3 = (c * (2107.0 * c - 9402.0 * p1 + 7042.0 * p2 - 1854.0 * p3) +
                       p1 * (11003.0 * p1 - 17246.0 * p2 + 4642.0 * p3) +
                       p2 * (7043.0 * p2 - 3882.0 * p3) + 547.0 * p3 * p3);
    const double a0 = (1.0 / 35.0) / sq_eps(b0);
    const double a1 = (12.0 / 35.0) / sq_eps(b1);
    const double a2 = (18.0 / 35.0) / sq_eps(b2);
    const double a3 = (4.0 / 35.0) / sq_eps(b3);
    const double inv = 1.0 / (a0 + a1 + a2 + a3);
    const double w0 = a0 * inv, w1 = a1 * inv, w2 = a2 * inv, w3 = a3 * inv;
    return w0 * s0 + w1 * s1 + w2 * s2 + w3 * s3;
}

double orc_interp_weno9(const double *q) { /* :148-199 */
    const double m4 = q[0], m3 = q[1], m2 = q[2], m1 = q[3], c = q[4];
    const double p1 = q[5], p2 = q[6], p3 = q[7], p4 = q[8];
    const double s0 = (1.0 / 5.0) * m4 + (-21.0 / 20.0) * m3 + (137.0 / 60.0) * m2 + (-163.0 / 60.0) * m1 + (137.0 / 60.0) * c;
    const double s1 = (-1.0 / 20.0) * m3 + (17.0 / 60.0) * m2 + (-43.0 / 60.0) * m1 + (77.0 / 60.0) * c + (1.0 / 5.0) * p1;
    const double s2 = (1.0 / 30.0) * m2 + (-13.0 / 60.0) * m1 + (47.0 / 60.0) * c + (9.0 / 20.0) * p1 + (-1.0 / 20.0) * p2;
    const double s3 = (-1.0 / 20.0) * m1 + (9.0 / 20.0) * c + (47.0 / 60.0) * p1 + (-13.0 / 60.0) * p2 + (1.0 / 30.0) * p3;
    const double s4 = (1.0 / 5.0) * c + (77.0 / 60.0) * p1 + (-43.0 / 60.0) * p2 + (17.0 / 60.0) * p3 + (-1.0 / 20.0) * p4;
    const double b0 = (m4 * (22658.0 * m4 - 208501.0 * m3 + 364863.0 * m2 - 288007.0 * m1 + 86329.0 * c) +
                       m3 * (482963.0 * m3 - 1704396.0 * m2 + 1358458.0 * m1 - 411487.0 * c) +
                       m2 * (1521393.0 * m2 - 2462076.0 * m1 + 758823.0 * c) +
                       m1 * (1020563.0 * m1 - 649501.0 * c) + 107918.0 * c * c);
    const double b1 = (m3 * (6908.0 * m3 - 60871.0 * m2 + 99213.0 * m1 - 70237.0 * c + 18079.0 * p1) +
                       m2 * (138563.0 * m2 - 464976.0 * m1 + 337018.0 * c - 88297.0 * p1) +
                       m1 * (406293.0 * m1 - 611976.0 * c + 165153.0 * p1) +
                       c * (242723.0 * c - 140251.0 * p1) + 22658.0 * p1 * p1);
    const double b2 = (m2 * (6908.0 * m2 - 51001.0 * m1 + 67923.0 * c - 38947.0 * p1 + 8209.0 * p2) +
                       m1 * (104963.0 * m1 - 299076.0 * c + 179098.0 * p1 - 38947.0 * p2) +
                       c * (231153.0 * c - 299076.0 * p1 + 67923.0 * p2) +
                       p1 * (104963.0 * p1 - 51001.0 * p2) + 6908.0 * p2 * p2);
    const double b3 = (m1 * (22658.0 * m1 - 140251.0 * c + 165153.0 * p1 - 88297.0 * p2 + 18079.0 * p3) +
                       c * (242723.0 * c - 611976.0 * p1 + 337018.0 * p2 - 70237.0 * p3) +
                       p1 * (406293.0 * p1 - 464976.0 * p2 + 99213.0 * p3) +
                       p2 * (138563.0 * p2 - 60871.0 * p3) + 6908.0 * p3 * p3);
    const double b4 = (c * (107918.0 * c - 649501.0 * p1 + 758823.0 * p2 - 411487.0 * p3 + 86329.0 * p4) +
                       p1 * (1020563.0 * p1 - 2462076.0 * p2 + 1358458.0 * p3 - 288007.0 * p4) +
                       p2 * (1521393.0 * p2 - 1704396.0 * p3 + 364863.0 * p4) +
                       p3 * (482963.0 * p3 - 208501.0 * p4) + 22658.0 * p4 * p4);
    const double a0 = (1.0 / 126.0) / sq_eps(b0);
    const double a1 = (10.0 / 63.0) / sq_eps(b1);
    const double a2 = (10.0 / 21.0) / sq_eps(b2);
    const double a3 = (20.0 / 63.0) / sq_eps(b3);
    const double a4 = (5.0 / 126.0) / sq_eps(b4);
    const double inv = 1.0 / (a0 + a1 + a2 + a3 + a4);
    const double w0 = a0 * inv, w1 = a1 * inv, w2 = a2 * inv, w3 = a3 * inv, w4 = a4 * inv;
    return w0 * s0 + w1 * s1 + w2 * s2 + w3 * s3 + w4 * s4;
}

/* Reconstruction of order 2r-1 from q[0..2r-2]. */
static double weno(int order, const double *q) {
    switch (order) {
    case 7: return orc_interp_weno7(q);
    case 9: return orc_interp_weno9(q);
    default: return orc_interp_weno5(q);
    }
}

/* Centred interpolation one order below the WENO order
 * (momentum_advection_weno.h:167,308,471). */
static double centred(int order, const double *q) {
    switch (order) {
    case 7: return orc_interp_6(q);
    case 9: return orc_interp_8(q);
    default: return orc_interp_4(q);
    }
}

static int order_r(int order) { return order == 9 ? 5 : (order == 7 ? 4 : 3); }

static void strides(const orc_dims *d, long st[3]) {
    st[0] = d->nlg[1] * d->nlg[2];
    st[1] = d->nlg[2];
    st[2] = 1;
}

/* Face value pair (phi+ upwind for U>0, phi- for U<0) at face ijk+1/2 along
 * stride s; stencils as in scalar_advection.h:346-357. */
static void face_weno(int order, const double *phi, long ijk, long s, double *php, double *phm) {
    const int r = order_r(order);
    double qp[9], qm[9];
    for (int m = 0; m < 2 * r - 1; ++m) {
        qp[m] = phi[ijk + (m - (r - 1)) * s];       /* ijk-(r-1) .. ijk+(r-1) */
        qm[m] = phi[ijk + (r - m) * s];             /* ijk+r .. ijk-(r-2) */
    }
    *php = weno(order, qp);
    *phm = weno(order, qm);
}

/* weno_{fifth,seventh,ninth}_order_a, scalar_advection.h:319-578. */
void orc_scalar_flux(const orc_dims *d, const double *rho0, const double *rho0_half,
                     const double *vel, const double *phi, double *flux, int dir, int order) {
    long st[3];
    strides(d, st);
    const int r = order_r(order);
    const long lo = r - 1;
    const long hi0 = d->nlg[0] - r, hi1 = d->nlg[1] - r, hi2 = d->nlg[2] - r;
    const long s = st[dir];
    const double *rho = (dir == 2) ? rho0 : rho0_half;
#pragma omp parallel for schedule(static)
    for (long i = lo; i < hi0; ++i) {
        for (long j = lo; j < hi1; ++j) {
            for (long k = lo; k < hi2; ++k) {
                const long ijk = i * st[0] + j * st[1] + k;
                double php, phm;
                face_weno(order, phi, ijk, s, &php, &phm);
                const double U = vel[ijk];
                flux[ijk] = 0.5 * ((U + fabs(U)) * php + (U - fabs(U)) * phm) * rho[k];
            }
        }
    }
}

/* scalar_flux_divergence, flux_divergence.h:4-32 */
void orc_scalar_flux_divergence(const orc_dims *d, const double *alpha0,
                                const double *alpha0_half, const double *flux, double *tend,
                                int dir) {
    (void)alpha0;
    long st[3];
    strides(d, st);
    const long gw = d->gw;
    const double dxi = d->dxi[dir];
    const long sm1 = -st[dir];
#pragma omp parallel for schedule(static)
    for (long i = gw; i < d->nlg[0] - gw; ++i)
        for (long j = gw; j < d->nlg[1] - gw; ++j)
            for (long k = gw; k < d->nlg[2] - gw; ++k) {
                const long ijk = i * st[0] + j * st[1] + k;
                tend[ijk] -= alpha0_half[k] * (flux[ijk] - flux[ijk + sm1]) * dxi;
            }
}

/* ScalarAdvection.pyx:69-123 (no sedimentation velocities). */
void orc_scalar_advection_update(const orc_dims *d, const orc_ref *rs, const double *const *vel,
                                 int nsc, const double *const *phi, double *const *tend,
                                 double *flux_scratch, int order) {
    for (int sc = 0; sc < nsc; ++sc) {
        for (int dir = 0; dir < 3; ++dir) {
            double *flux = flux_scratch + (long)dir * d->npg;
            orc_scalar_flux(d, rs->rho0, rs->rho0_half, vel[dir], phi[sc], flux, dir, order);
            orc_scalar_flux_divergence(d, rs->alpha0, rs->alpha0_half, flux, tend[sc], dir);
        }
    }
}

/* momentum_flux_divergence, flux_divergence.h:34-80 */
static void momentum_flux_divergence(const orc_dims *d, const orc_ref *rs, const double *flux,
                                     double *tend, int d_advected, int d_advecting) {
    long st[3];
    strides(d, st);
    const long gw = d->gw;
    const double dxi = d->dxi[d_advecting];
    const long sm1 = -st[d_advecting];
    const double *alpha = (d_advected != 2) ? rs->alpha0_half : rs->alpha0;
#pragma omp parallel for schedule(static)
    for (long i = gw; i < d->nlg[0] - gw; ++i)
        for (long j = gw; j < d->nlg[1] - gw; ++j)
            for (long k = gw; k < d->nlg[2] - gw; ++k) {
                const long ijk = i * st[0] + j * st[1] + k;
                tend[ijk] -= alpha[k] * (flux[ijk] - flux[ijk + sm1]) * dxi;
            }
}

/* weno_{fifth,seventh,ninth}_order_m, momentum_advection_weno.h:113-581. */
void orc_momentum_pair(const orc_dims *d, const orc_ref *rs, const double *vel_advected,
                       const double *vel_advecting, double *tend, int d_advected,
                       int d_advecting, int order, double *flux) {
    long st[3];
    strides(d, st);
    const int r = order_r(order);
    const long lo = r - 1;
    const long hi0 = d->nlg[0] - r, hi1 = d->nlg[1] - r, hi2 = d->nlg[2] - r;
    const long s_ed = st[d_advecting];
    const long s_ing = st[d_advected];
    const int nc = 2 * (r - 1); /* centred points: interp_4/6/8 */
#pragma omp parallel for schedule(static)
    for (long i = lo; i < hi0; ++i) {
        for (long j = lo; j < hi1; ++j) {
            for (long k = lo; k < hi2; ++k) {
                const long ijk = i * st[0] + j * st[1] + k;
                double php, phm, qc[8];
                face_weno(order, vel_advected, ijk, s_ed, &php, &phm);
                for (int m = 0; m < nc; ++m) qc[m] = vel_advecting[ijk + (m - (r - 2)) * s_ing];
                const double U = centred(order, qc);
                double rho;
                if (d_advected != 2 && d_advecting != 2) rho = rs->rho0_half[k];
                else if (d_advected == 2 && d_advecting == 2) rho = rs->rho0_half[k + 1];
                else rho = rs->rho0[k];
                flux[ijk] = 0.5 * ((U + fabs(U)) * php + (U - fabs(U)) * phm) * rho;
            }
        }
    }
    momentum_flux_divergence(d, rs, flux, tend, d_advected, d_advecting);
}

/* MomentumAdvection.pyx:41-69 */
void orc_momentum_advection_update(const orc_dims *d, const orc_ref *rs, const double *const *vel,
                                   double *const *tend, int order, double *flux_scratch) {
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
            orc_momentum_pair(d, rs, vel[a], vel[b], tend[a], a, b, order, flux_scratch);
}

/* ------------------------------------------------------------------------ */
/* Thermodynamics                                                            */

/* latent_heat_variable, Microphysics.pxd:71-75 */
double orc_latent_heat_variable(double T) {
    const double TC = T - 273.15;
    return (2500.8 - 2.36 * TC + 0.0016 * TC * TC - 0.00006 * TC * TC * TC) * 1000.0;
}

/* lookup, lookup.h:48-54 */
double orc_lookup_eval(const orc_lookup *lt, double x) {
    /* The reference does no bounds check; out-of-table temperatures read
     * outside the array there.  Clamping keeps those reads inside the table
     * and leaves every in-range value unchanged. */
    double fi = floor((x - lt->x_min) * lt->dxi);
    if (!(fi >= 0.0)) fi = 0.0;
    if (fi > (double)(lt->n - 2)) fi = (double)(lt->n - 2);
    const size_t indx = (size_t)fi;
    const double y1 = lt->y[indx];
    const double y2 = lt->y[indx + 1];
    const double x1 = lt->x_min + lt->dx * indx;
    return y1 + (x - x1) * (y2 - y1) * lt->dxi;
}

/* thermodynamic_functions.h */
double orc_exner(double p0) { return pow((p0 / C_P_TILDE), C_KAPPA); }                /* :5-7 */
static double pd_c(double p0, double qt, double qv) {                                     /* :19-21 */
    return p0 * (1.0 - qt) / (1.0 - qt + C_EPS_VI * qv);
}
static double pv_c(double p0, double qt, double qv) {                                     /* :23-25 */
    return p0 * C_EPS_VI * qv / (1.0 - qt + C_EPS_VI * qv);
}
static double cpm_c(double qt) { return (1.0 - qt) * C_CPD + qt * C_CPV; }                /* :36-38 */
static double qv_star_c(double p0, double qt, double pv) {                                 /* :57-59 */
    return C_EPS_V * (1.0 - qt) * pv / (p0 - pv);
}
double orc_alpha(double p0, double T, double qt, double qv) {                             /* :61-63 */
    return (C_RD * T) / p0 * (1.0 - qt + C_EPS_VI * qv);
}
static double buoyancy_c(double alpha0, double alpha) { return C_G * (alpha - alpha0) / alpha0; }
static double theta_rho_c(double p0, double T, double qt, double qv) {                    /* :31-34 */
    return T * (1.0 - qt + C_EPS_VI * qv) / orc_exner(p0);
}
static double thetali_c(double p0, double T, double qt, double ql, double qi, double L) {  /* :14-17 */
    return T / orc_exner(p0) * exp(-L * (ql / (1.0 - qt) + qi / (1.0 - qt)) / (T * C_CPD));
}

/* entropies.h:5-15 */
static double sd_c(double pd, double T) {
    return C_SD_TILDE + C_CPD * log(T / C_T_TILDE) - C_RD * log(pd / C_P_TILDE);
}
static double sv_c(double pv, double T) {
    return C_SV_TILDE + C_CPV * log(T / C_T_TILDE) - C_RV * log(pv / C_P_TILDE);
}
static double sc_c(double L, double T) { return -L / T; }

/* temperature_no_ql, thermodynamics_sa.h:10-15 */
static double temperature_no_ql(double pd, double pv, double s, double qt) {
    return C_T_TILDE * exp((s - (1.0 - qt) * (C_SD_TILDE - C_RD * log(pd / C_P_TILDE)) -
                            qt * (C_SV_TILDE - C_RV * log(pv / C_P_TILDE))) /
                           ((1.0 - qt) * C_CPD + qt * C_CPV));
}

/* eos_c, thermodynamics_sa.h:17-66 (lambda == 1, L = latent_heat_variable). */
int orc_eos_sa(const orc_lookup *lt, double p0, double s, double qt, double *T, double *qv,
               double *ql, double *qi, int max_iter) {
    *qv = qt;
    *ql = 0.0;
    *qi = 0.0;
    const double pv_1 = pv_c(p0, qt, qt);
    const double pd_1 = p0 - pv_1;
    double T_1 = temperature_no_ql(pd_1, pv_1, s, qt);
    const double pv_star_1 = orc_lookup_eval(lt, T_1);
    const double qv_star_1 = qv_star_c(p0, qt, pv_star_1);
    if (qt <= qv_star_1) {
        *T = T_1;
        return 0;
    }
    const double sigma_1 = qt - qv_star_1;
    const double lam_1 = 1.0;
    const double L_1 = orc_latent_heat_variable(T_1);
    (void)lam_1;
    const double s_1 = sd_c(pd_1, T_1) * (1.0 - qt) + sv_c(pv_1, T_1) * qt + sc_c(L_1, T_1) * sigma_1;
    double f_1 = s - s_1;
    double T_2 = T_1 + sigma_1 * L_1 / ((1.0 - qt) * C_CPD + qv_star_1 * C_CPV);
    double delta_T;
    double qv_star_2, sigma_2, lam_2;
    int it = 0;
    do {
        const double pv_star_2 = orc_lookup_eval(lt, T_2);
        qv_star_2 = qv_star_c(p0, qt, pv_star_2);
        const double pv_2 = pv_c(p0, qt, qv_star_2);
        const double pd_2 = p0 - pv_2;
        sigma_2 = qt - qv_star_2;
        lam_2 = 1.0;
        const double L_2 = orc_latent_heat_variable(T_2);
        const double s_2 = sd_c(pd_2, T_2) * (1.0 - qt) + sv_c(pv_2, T_2) * qt + sc_c(L_2, T_2) * sigma_2;
        const double f_2 = s - s_2;
        const double T_n = T_2 - f_2 * (T_2 - T_1) / (f_2 - f_1);
        T_1 = T_2;
        T_2 = T_n;
        f_1 = f_2;
        delta_T = fabs(T_2 - T_1);
        ++it;
        if (max_iter > 0 && it >= max_iter) break;
    } while (delta_T >= 1.0e-3 || sigma_2 < 0.0);
    *T = T_2;
    *qv = qv_star_2;
    *ql = lam_2 * sigma_2;
    *qi = (1.0 - lam_2) * sigma_2;
    return it;
}

/* ThermodynamicsSA.entropy, ThermodynamicsSA.pyx:138-157 */
double orc_entropy_sa(double p0, double T, double qt, double ql, double qi) {
    const double qv = qt - ql - qi;
    const double pd = pd_c(p0, qt, qv);
    const double pv = pv_c(p0, qt, qv);
    const double L = orc_latent_heat_variable(T);
    return sd_c(pd, T) * (1.0 - qt) + sv_c(pv, T) * qt + sc_c(L, T) * (ql + qi);
}

/* buoyancy_update_sa, thermodynamics_sa.h:98-132 */
void orc_buoyancy_update_sa(const orc_dims *d, const double *alpha0_half, const double *alpha,
                            double *buoyancy, double *wt) {
    long st[3];
    strides(d, st);
    const long n0 = d->nlg[0], n1 = d->nlg[1], n2 = d->nlg[2];
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n0; ++i)
        for (long j = 0; j < n1; ++j)
            for (long k = 0; k < n2 - 1; ++k) {
                const long ijk = i * st[0] + j * st[1] + k;
                buoyancy[ijk] = buoyancy_c(alpha0_half[k], alpha[ijk]);
            }
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n0; ++i)
        for (long j = 0; j < n1; ++j)
            for (long k = 1; k < n2 - 3; ++k) {
                const long ijk = i * st[0] + j * st[1] + k;
                wt[ijk] = wt[ijk] + orc_interp_2(buoyancy[ijk], buoyancy[ijk + 1]);
            }
}

/* bvf_sa, thermodynamics_sa.h:134-182 */
void orc_bvf_sa(const orc_dims *d, const orc_lookup *lt, const double *p0, const double *T,
                const double *qt, const double *qv, double *theta_rho, double *bvf) {
    long st[3];
    strides(d, st);
    const long n0 = d->nlg[0], n1 = d->nlg[1], n2 = d->nlg[2];
    const double dzi = 1.0 / d->dx[2];
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n0; ++i)
        for (long j = 0; j < n1; ++j)
            for (long k = 1; k < n2 - 1; ++k) {
                const long ijk = i * st[0] + j * st[1] + k;
                theta_rho[ijk] = theta_rho_c(p0[k], T[ijk], qt[ijk], qv[ijk]);
            }
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n0; ++i)
        for (long j = 0; j < n1; ++j)
            for (long k = 2; k < n2 - 2; ++k) {
                const long ijk = i * st[0] + j * st[1] + k;
                if (qv[ijk] < qt[ijk]) {
                    const double Tk = T[ijk];
                    const double Lv = orc_latent_heat_variable(Tk);
                    const double pv_star = orc_lookup_eval(lt, Tk);
                    const double rsl = C_EPS_V * pv_star / (p0[k] - pv_star);
                    const double gamma_w =
                        C_G / C_CPD * (1.0 / (1.0 - qt[ijk])) * (1.0 + Lv * rsl / (C_RD * Tk)) /
                        (cpm_c(qt[ijk]) / C_CPD + Lv * Lv * (C_EPS_V + rsl) * rsl / (C_CPD * C_RD * Tk * Tk));
                    const double dTdz = (orc_interp_2(T[ijk], T[ijk + 1]) - orc_interp_2(T[ijk - 1], T[ijk])) * dzi;
                    const double dqtdz = (orc_interp_2(qt[ijk], qt[ijk + 1]) - orc_interp_2(qt[ijk - 1], qt[ijk])) * dzi;
                    bvf[ijk] = C_G / Tk * (dTdz + gamma_w) * (1.0 + Lv * rsl / (C_RD * Tk)) - dqtdz / (1.0 - qt[ijk]);
                } else {
                    bvf[ijk] = C_G / theta_rho[ijk] *
                               (orc_interp_2(theta_rho[ijk], theta_rho[ijk + 1]) -
                                orc_interp_2(theta_rho[ijk - 1], theta_rho[ijk])) * dzi;
                }
            }
}

/* thetali_update, thermodynamics_sa.h:185-211 */
void orc_thetali_update(const orc_dims *d, const double *p0, const double *T, const double *qt,
                        const double *ql, const double *qi, double *thetali) {
    long st[3];
    strides(d, st);
    const long n0 = d->nlg[0], n1 = d->nlg[1], n2 = d->nlg[2];
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n0; ++i)
        for (long j = 0; j < n1; ++j)
            for (long k = 0; k < n2; ++k) {
                const long ijk = i * st[0] + j * st[1] + k;
                const double Lv = orc_latent_heat_variable(T[ijk]);
                thetali[ijk] = thetali_c(p0[k], T[ijk], qt[ijk], ql[ijk], qi[ijk], Lv);
            }
}

/* ThermodynamicsSA.update, ThermodynamicsSA.pyx:176-213 */
void orc_thermo_sa_update(const orc_dims *d, const orc_ref *rs, const orc_lookup *lt,
                          double *s, double *qt, double *wt, double *T, double *qv, double *ql,
                          double *qi, double *alpha, double *buoyancy, double *theta_rho,
                          double *bvf, double *thetali, int do_clip, int max_iter,
                          orc_sa_stats *stats) {
    long st[3];
    strides(d, st);
    const long n0 = d->nlg[0], n1 = d->nlg[1], n2 = d->nlg[2];
    const double *p0 = rs->p0_half;
    /* clip_qt, thermodynamics_sa.h:213-220 */
    if (do_clip) {
#pragma omp parallel for schedule(static)
        for (long i = 0; i < d->npg; ++i) qt[i] = fmax(qt[i], 1e-11);
    }
    /* eos_update, thermodynamics_sa.h:68-96 */
    long n_sat = 0, n_it = 0, n_max = 0, n_cap = 0;
#pragma omp parallel for schedule(static) reduction(+ : n_sat, n_it, n_cap) reduction(max : n_max)
    for (long i = 0; i < n0; ++i)
        for (long j = 0; j < n1; ++j)
            for (long k = 0; k < n2; ++k) {
                const long ijk = i * st[0] + j * st[1] + k;
                const int it = orc_eos_sa(lt, p0[k], s[ijk], qt[ijk], &T[ijk], &qv[ijk], &ql[ijk],
                                          &qi[ijk], max_iter);
                alpha[ijk] = orc_alpha(p0[k], T[ijk], qt[ijk], qv[ijk]);
                if (it > 0) {
                    n_sat += 1;
                    n_it += it;
                    if (it > n_max) n_max = it;
                    if (max_iter > 0 && it >= max_iter) n_cap += 1;
                }
            }
    if (stats) {
        stats->saturated = n_sat;
        stats->iterations = n_it;
        stats->max_iterations = n_max;
        stats->capped = n_cap;
    }
    orc_buoyancy_update_sa(d, rs->alpha0_half, alpha, buoyancy, wt);
    orc_bvf_sa(d, lt, p0, T, qt, qv, theta_rho, bvf);
    orc_thetali_update(d, p0, T, qt, ql, qi, thetali);
}

/* thermodynamics_dry.h:8-10 */
double orc_eos_dry(double pd, double s) {
    return C_T_TILDE * (exp((s - C_SD_TILDE + C_RD * log(pd / C_P_TILDE)) / C_CPD));
}
/* ThermodynamicsDry.entropy -> sd_c (ThermodynamicsDry.pyx:93-97) */
double orc_entropy_dry(double pd, double T) { return sd_c(pd, T); }

/* buoyancy_update, thermodynamics_dry.h:39-73 */
void orc_buoyancy_update_dry(const orc_dims *d, const double *alpha0_half, const double *alpha,
                             double *buoyancy, double *wt) {
    long st[3];
    strides(d, st);
    const long n0 = d->nlg[0], n1 = d->nlg[1], n2 = d->nlg[2];
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n0 - 1; ++i)
        for (long j = 0; j < n1 - 1; ++j)
            for (long k = 0; k < n2 - 1; ++k) {
                const long ijk = i * st[0] + j * st[1] + k;
                buoyancy[ijk] = buoyancy_c(alpha0_half[k], alpha[ijk]);
            }
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n0 - 1; ++i)
        for (long j = 0; j < n1 - 1; ++j)
            for (long k = 0; k < n2 - 1; ++k) {
                const long ijk = i * st[0] + j * st[1] + k;
                wt[ijk] += orc_interp_2(buoyancy[ijk], buoyancy[ijk + 1]);
            }
}

/* bvf_dry, thermodynamics_dry.h:75-110 */
void orc_bvf_dry(const orc_dims *d, const double *p0, const double *T, double *theta, double *bvf) {
    long st[3];
    strides(d, st);
    const long n0 = d->nlg[0], n1 = d->nlg[1], n2 = d->nlg[2];
    const double dzi = 1.0 / d->dx[2];
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n0; ++i)
        for (long j = 0; j < n1; ++j)
            for (long k = 1; k < n2; ++k) {
                const long ijk = i * st[0] + j * st[1] + k;
                theta[ijk] = T[ijk] / orc_exner(p0[k]);
            }
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n0; ++i)
        for (long j = 0; j < n1; ++j)
            for (long k = 2; k < n2 - 1; ++k) {
                const long ijk = i * st[0] + j * st[1] + k;
                bvf[ijk] = C_G / theta[ijk] *
                           (orc_interp_2(theta[ijk], theta[ijk + 1]) - orc_interp_2(theta[ijk - 1], theta[ijk])) * dzi;
            }
}

/* ThermodynamicsDry.update, ThermodynamicsDry.pyx:105-120 */
void orc_thermo_dry_update(const orc_dims *d, const orc_ref *rs, double *s, double *wt,
                           double *T, double *alpha, double *buoyancy, double *theta,
                           double *bvf) {
    long st[3];
    strides(d, st);
    const long n0 = d->nlg[0], n1 = d->nlg[1], n2 = d->nlg[2];
    const double *p0 = rs->p0_half;
    /* eos_update, thermodynamics_dry.h:12-37 */
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n0; ++i)
        for (long j = 0; j < n1; ++j)
            for (long k = 0; k < n2; ++k) {
                const long ijk = i * st[0] + j * st[1] + k;
                T[ijk] = orc_eos_dry(p0[k], s[ijk]);
                alpha[ijk] = orc_alpha(p0[k], T[ijk], 0.0, 0.0);
            }
    orc_buoyancy_update_dry(d, rs->alpha0_half, alpha, buoyancy, wt);
    orc_bvf_dry(d, p0, T, theta, bvf);
}

/* ------------------------------------------------------------------------ */
/* Kinematics, SGS and diffusion                                              */

/* Kinematics.update (Kinematics.pyx:41-58): compute_velocity_gradient,
 * compute_strain_rate, compute_strain_rate_mag (kinematics.h:6-138).
 * vgrad and strain hold 9 npg arrays, component (vi, d) at (3 vi + d) npg. */
void orc_kinematics_update(const orc_dims *d, const double *const *vel, double *vgrad, double *strain,
                           double *smag) {
    long st[3];
    strides(d, st);
    const long n0 = d->nlg[0], n1 = d->nlg[1], n2 = d->nlg[2], npg = d->npg;
    long count = 0;
    for (int vi = 0; vi < 3; ++vi)
        for (int dd = 0; dd < 3; ++dd) {
            const double *v = vel[vi];
            double *g = vgrad + count * npg;
            const long sp1 = st[dd];
            const double dxi = d->dxi[dd];
#pragma omp parallel for schedule(static)
            for (long i = 0; i < n0 - 1; ++i)
                for (long j = 0; j < n1 - 1; ++j)
                    for (long k = 0; k < n2 - 1; ++k) {
                        const long ijk = i * st[0] + j * st[1] + k;
                        g[ijk] = (v[ijk + sp1] - v[ijk]) * dxi;
                    }
            ++count;
        }
    count = 0;
    for (int v1 = 0; v1 < 3; ++v1)
        for (int dd = 0; dd < 3; ++dd) {
            const long s1 = 3 * npg * v1 + npg * dd, s2 = 3 * npg * dd + npg * v1, sh = count * npg;
#pragma omp parallel for schedule(static)
            for (long i = 0; i < n0 - 1; ++i)
                for (long j = 0; j < n1 - 1; ++j)
                    for (long k = 0; k < n2 - 1; ++k) {
                        const long ijk = i * st[0] + j * st[1] + k;
                        strain[sh + ijk] = 0.5 * (vgrad[s1 + ijk] + vgrad[s2 + ijk]);
                    }
            ++count;
        }
    for (long i = 0; i < npg; ++i) smag[i] = 0.0;
    for (int dd = 0; dd < 3; ++dd) {
        const long shs = 4 * dd * npg;
#pragma omp parallel for schedule(static)
        for (long i = 1; i < n0 - 1; ++i)
            for (long j = 1; j < n1 - 1; ++j)
                for (long k = 1; k < n2 - 1; ++k) {
                    const long ijk = i * st[0] + j * st[1] + k;
                    const long t = shs + ijk - st[dd];
                    smag[ijk] += strain[t] * strain[t];
                }
    }
    for (int v1 = 0; v1 < 2; ++v1)
        for (int dd = v1; dd < 3; ++dd) {
            const long shs = 3 * npg * v1 + npg * dd;
#pragma omp parallel for schedule(static)
            for (long i = 1; i < n0 - 1; ++i)
                for (long j = 1; j < n1 - 1; ++j)
                    for (long k = 1; k < n2 - 1; ++k) {
                        const long ijk = i * st[0] + j * st[1] + k;
                        const long p1 = shs + ijk, p2 = shs + ijk - st[v1], p3 = shs + ijk - st[dd],
                                   p4 = shs + ijk - st[dd] - st[v1];
                        /* the loop over dd starts at v1, so the (0,0) and (1,1)
                         * pairs are counted again here; at i = 1 the (0,0)
                         * pair's last point lies before the array, where the
                         * reference reads whatever precedes it: read as zero
                         * (no consumer uses |S| there when gw >= 3) */
                        const double s4 = p4 >= 0 ? strain[p4] : 0.0;
                        const double si = 0.25 * (strain[p1] + strain[p2] + strain[p3] + s4);
                        smag[ijk] += 2.0 * si * si;
                    }
        }
    for (long i = 0; i < npg; ++i) smag[i] = sqrt(2.0 * smag[i]);
}

/* smagorinsky_update, sgs.h:4-19 (Smagorinsky.update without the wall or
 * iles adjustments, SGS.pyx:146-152). */
void orc_smagorinsky_update(const orc_dims *d, double *visc, double *diff, const double *bvf, const double *smag,
                            double cs, double prt) {
    const double delta = cbrt(d->dx[0] * d->dx[1] * d->dx[2]);
    for (long i = 0; i < d->npg; ++i) {
        visc[i] = cs * cs * delta * delta * smag[i];
        if (bvf[i] > 0.0) {
            const double fb = sqrt(fmax(1.0 - bvf[i] / (prt * smag[i] * smag[i]), 0.0));
            visc[i] = visc[i] * fb;
        }
        diff[i] = visc[i] / prt;
    }
}

/* second_order_diffusion, scalar_diffusion.h:6-58 (factor 1), into flux. */
static void second_order_diffusion(const orc_dims *d, const double *rho0, const double *rho0_half,
                                   const double *diffusivity, const double *phi, double *flux, double dx, int dd) {
    long st[3];
    strides(d, st);
    const long gw = d->gw;
    const double dxi = 1.0 / dx;
    const double *rho = dd == 2 ? rho0 : rho0_half;
    const long s = st[dd];
#pragma omp parallel for schedule(static)
    for (long i = gw - 1; i < d->nlg[0] - gw; ++i)
        for (long j = gw - 1; j < d->nlg[1] - gw; ++j)
            for (long k = gw - 1; k < d->nlg[2] - gw; ++k) {
                const long ijk = i * st[0] + j * st[1] + k;
                flux[ijk] = -1.0 * orc_interp_2(diffusivity[ijk], diffusivity[ijk + s]) * (phi[ijk + s] - phi[ijk]) *
                            rho[k] * dxi;
            }
    /* the surface flux is exactly zero (one rank owns the bottom) */
    for (long i = gw - 1; i < d->nlg[0] - gw; ++i)
        for (long j = gw - 1; j < d->nlg[1] - gw; ++j) flux[i * st[0] + j * st[1] + gw - 1] = 0.0;
}

/* ScalarDiffusion.update (ScalarDiffusion.pyx:76-119) for nsc scalars
 * without the qt entropy source (its default, :31-36). */
void orc_scalar_diffusion_update(const orc_dims *d, const orc_ref *rs, const double *diffusivity, int nsc,
                                 const double *const *phi, double *const *tend, double *flux) {
    for (int sc = 0; sc < nsc; ++sc)
        for (int dd = 0; dd < 3; ++dd) {
            second_order_diffusion(d, rs->rho0, rs->rho0_half, diffusivity, phi[sc], flux, d->dx[dd], dd);
            orc_scalar_flux_divergence(d, rs->alpha0, rs->alpha0_half, flux, tend[sc], dd);
        }
}

/* compute_diffusive_flux_m, momentum_diffusion.h:5-88. */
static void diffusive_flux_m(const orc_dims *d, const double *strain, const double *visc, double *flux,
                             const double *rho0, const double *rho0_half, int i1, int i2) {
    long st[3];
    strides(d, st);
    const long gw = d->gw;
    const long s1 = st[i1], s2 = st[i2];
#pragma omp parallel for schedule(static)
    for (long i = gw - 1; i < d->nlg[0] - gw; ++i)
        for (long j = gw - 1; j < d->nlg[1] - gw; ++j)
            for (long k = gw - 1; k < d->nlg[2] - gw; ++k) {
                const long ijk = i * st[0] + j * st[1] + k;
                if (i1 != 2 && i2 != 2) {
                    if (i1 == i2) {
                        flux[ijk] = -2.0 * strain[ijk] * visc[ijk + s1] * rho0_half[k];
                    } else {
                        const double vi = 0.25 * (visc[ijk] + visc[ijk + s1] + visc[ijk + s2] + visc[ijk + s1 + s2]);
                        flux[ijk] = -2.0 * strain[ijk] * vi * rho0_half[k];
                    }
                } else if (i1 == 2 && i2 == 2) {
                    flux[ijk] = -2.0 * strain[ijk] * visc[ijk + s1] * rho0_half[k + 1];
                } else {
                    const double vi = 0.25 * (visc[ijk] + visc[ijk + s1] + visc[ijk + s2] + visc[ijk + s1 + s2]);
                    flux[ijk] = -2.0 * strain[ijk] * vi * rho0[k];
                }
            }
    for (long i = gw - 1; i < d->nlg[0] - gw; ++i)
        for (long j = gw - 1; j < d->nlg[1] - gw; ++j) flux[i * st[0] + j * st[1] + gw - 1] = 0.0;
}

/* MomentumDiffusion.update (MomentumDiffusion.pyx:44-66): the nine stress
 * fluxes and their divergence, then the entropy source of the dissipation,
 * compute_entropy_source (momentum_diffusion.h:90-118), into s_tend. */
void orc_momentum_diffusion_update(const orc_dims *d, const orc_ref *rs, const double *strain,
                                   const double *visc, double *const *tend, double *flux, const double *smag,
                                   const double *T, double *s_tend) {
    long st[3];
    strides(d, st);
    for (int i1 = 0; i1 < 3; ++i1)
        for (int i2 = 0; i2 < 3; ++i2) {
            diffusive_flux_m(d, strain + (long)(3 * i1 + i2) * d->npg, visc, flux, rs->rho0, rs->rho0_half, i1, i2);
            momentum_flux_divergence(d, rs, flux, tend[i1], i1, i2);
        }
    const long gw = d->gw;
#pragma omp parallel for schedule(static)
    for (long i = gw; i < d->nlg[0] - gw; ++i)
        for (long j = gw; j < d->nlg[1] - gw; ++j)
            for (long k = gw; k < d->nlg[2] - gw; ++k) {
                const long ijk = i * st[0] + j * st[1] + k;
                s_tend[ijk] += smag[ijk] * smag[ijk] * visc[ijk] / T[ijk];
            }
}

/* ------------------------------------------------------------------------ */
/* BOMEX case physics: surface fluxes, large-scale forcing, Rayleigh damping  */

/* entropyflux_from_thetaflux_qtflux, surface.h:95-106 */
double orc_entropyflux_from_thetaflux_qtflux(double thetaflux, double qtflux, double p0_b, double T_b, double qt_b,
                                             double qv_b) {
    const double exner_b = orc_exner(p0_b);
    const double pd_b = pd_c(p0_b, qt_b, qv_b);
    const double pv_b = pv_c(p0_b, qt_b, qv_b);
    const double sd_b = sd_c(pd_b, T_b);
    const double sv_b = sv_c(pv_b, T_b);
    const double cp_b = cpm_c(qt_b);
    return cp_b * thetaflux * exner_b / T_b + qtflux * (sv_b - sd_b);
}

/* SurfaceBomex.update (Surface.pyx:326-373) and the moist branch of
 * SurfaceBase.update (Surface.pyx:150-183): the surface fluxes of s, qt, u
 * and v, added to the tendencies at the first interior level, and the
 * column diagnostics.  gustiness = 0 (SurfaceBomex does not set it).
 * flux2d (may be NULL): [8][nlg0*nlg1] = u_flux, v_flux, s_flux, qt_flux,
 * shf, lhf, b_flux, obukhov_length; t_mean_gw = horizontal mean of T at gw;
 * L is the latent heat of the run (Microphysics.pxd:65-75). */
void orc_surface_bomex(const orc_dims *d, const orc_ref *rs, const double *u, const double *v, const double *qt,
                       const double *T, const double *qv, const double *ql, double *ut, double *vt, double *st,
                       double *qtt, double u0, double v0, double t_mean_gw, int lh_constant, double *flux2d) {
    long stt[3];
    strides(d, stt);
    const long gw = d->gw, n0 = d->nlg[0], n1 = d->nlg[1], n2d = n0 * n1;
    const double theta_flux = 8.0e-3, ustar = 0.28, qt_flux = 5.2e-5, gustiness = 0.0;
    double *uf = calloc(n2d, sizeof(double)), *vf = calloc(n2d, sizeof(double)), *sf = calloc(n2d, sizeof(double));
    double *ws = calloc(n2d, sizeof(double));
    for (long i = 0; i < n0; ++i)
        for (long j = 0; j < n1; ++j) {
            const long ijk = i * stt[0] + j * stt[1] + gw, ij = i * n1 + j;
            sf[ij] = orc_entropyflux_from_thetaflux_qtflux(theta_flux, qt_flux, rs->p0_half[gw], T[ijk], qt[ijk], qv[ijk]);
        }
    /* compute_windspeed, surface.h:157-185 */
    for (long i = 1; i < n0; ++i)
        for (long j = 1; j < n1; ++j) {
            const long ijk = i * stt[0] + j * stt[1] + gw, ij = i * n1 + j;
            const double ui = orc_interp_2(u[ijk - stt[0]], u[ijk]) + u0;
            const double vi = orc_interp_2(v[ijk - stt[1]], v[ijk]) + v0;
            ws[ij] = fmax(sqrt(ui * ui + vi * vi), gustiness);
        }
    /* -ustar_**2 as the square of ustar */
    for (long i = 1; i < n0 - 1; ++i)
        for (long j = 1; j < n1 - 1; ++j) {
            const long ijk = i * stt[0] + j * stt[1] + gw, ij = i * n1 + j;
            uf[ij] = -(ustar * ustar) / orc_interp_2(ws[ij], ws[ij + n1]) * (u[ijk] + u0);
            vf[ij] = -(ustar * ustar) / orc_interp_2(ws[ij], ws[ij + 1]) * (v[ijk] + v0);
        }
    const double dzi = 1.0 / d->dx[2];
    const double tf = rs->alpha0_half[gw] / rs->alpha0[gw - 1] * dzi;
    for (long i = gw; i < n0 - gw; ++i)
        for (long j = gw; j < n1 - gw; ++j) {
            const long ijk = i * stt[0] + j * stt[1] + gw, ij = i * n1 + j;
            const double lv = lh_constant ? 2.501e6 : orc_latent_heat_variable(T[ijk]);
            const double lhf = qt_flux * rs->rho0_half[gw] * lv;
            const double pv = pv_c(rs->p0_half[gw], qt[ijk], qt[ijk] - ql[ijk]);
            const double pd = pd_c(rs->p0_half[gw], qt[ijk], qt[ijk] - ql[ijk]);
            const double sv = sv_c(pv, T[ijk]);
            const double sd = sd_c(pd, T[ijk]);
            const double shf = (sf[ij] * rs->rho0_half[gw] - lhf / lv * (sv - sd)) * T[ijk];
            const double cp_ = cpm_c(qt[ijk]);
            const double bf = C_G * rs->alpha0_half[gw] / cp_ / t_mean_gw *
                              (shf + (C_EPS_VI - 1.0) * cp_ * t_mean_gw * lhf / lv);
            const double lmo = -ustar * ustar * ustar / bf / 0.35; /* vkb, generate_parameters.py:44 */
            ut[ijk] += uf[ij] * tf;
            vt[ijk] += vf[ij] * tf;
            st[ijk] += sf[ij] * tf;
            qtt[ijk] += qt_flux * tf;
            if (flux2d) {
                flux2d[0 * n2d + ij] = uf[ij];
                flux2d[1 * n2d + ij] = vf[ij];
                flux2d[2 * n2d + ij] = sf[ij];
                flux2d[3 * n2d + ij] = qt_flux;
                flux2d[4 * n2d + ij] = shf;
                flux2d[5 * n2d + ij] = lhf;
                flux2d[6 * n2d + ij] = bf;
                flux2d[7 * n2d + ij] = lmo;
            }
        }
    free(uf);
    free(vf);
    free(sf);
    free(ws);
}

/* ForcingBomex.initialize (Forcing.pyx:105-145): the profiles on the
 * ghosted levels, prof = [5][nlg2] = ug, vg, dtdt, dqtdt, subsidence. */
void orc_forcing_bomex_profiles(const orc_dims *d, const double *zl_half, const double *p0_half, double *prof) {
    const long n2 = d->nlg[2];
    double *ug = prof, *vg = prof + n2, *dtdt = prof + 2 * n2, *dqtdt = prof + 3 * n2, *sub = prof + 4 * n2;
    for (long k = 0; k < n2; ++k) {
        const double z = zl_half[k];
        ug[k] = -10.0 + (1.8e-3) * z;
        vg[k] = 0.0;
        dtdt[k] = 0.0;
        dqtdt[k] = 0.0;
        sub[k] = 0.0;
        if (z <= 1500.0) dtdt[k] = (-2.0 / (3600 * 24.0)) * orc_exner(p0_half[k]);
        if (z > 1500.0)
            dtdt[k] = (-2.0 / (3600 * 24.0) + (z - 1500.0) * (0.0 - -2.0 / (3600 * 24.0)) / (3000.0 - 1500.0)) *
                      orc_exner(p0_half[k]);
        if (z <= 300.0) dqtdt[k] = -1.2e-8;
        if (z > 300.0 && z <= 500.0) dqtdt[k] = -1.2e-8 + (z - 300.0) * (0.0 - -1.2e-8) / (500.0 - 300.0);
        if (z <= 1500.0) sub[k] = 0.0 + z * (-0.65 / 100.0 - 0.0) / (1500.0 - 0.0);
        if (z > 1500.0 && z <= 2100.0) sub[k] = -0.65 / 100 + (z - 1500.0) * (0.0 - -0.65 / 100.0) / (2100.0 - 1500.0);
    }
}

/* apply_subsidence, Forcing.pyx:2013-2040: the last interior level and the
 * ghost levels above it repeat the tendency of the level below. */
static void apply_subsidence(const orc_dims *d, const double *sub, const double *values, double *tend) {
    long st[3];
    strides(d, st);
    const long gw = d->gw, kmax = d->nlg[2] - gw - 1;
    const double dxi = d->dxi[2];
#pragma omp parallel for schedule(static)
    for (long i = gw; i < d->nlg[0] - gw; ++i)
        for (long j = gw; j < d->nlg[1] - gw; ++j) {
            double t = 0.0;
            long k = gw;
            for (; k < kmax; ++k) {
                const long ijk = i * st[0] + j * st[1] + k;
                t = (values[ijk + 1] - values[ijk]) * dxi * sub[k];
                tend[ijk] -= t;
            }
            for (; k < d->nlg[2]; ++k) tend[i * st[0] + j * st[1] + k] -= t;
        }
}

/* ForcingBomex.update (Forcing.pyx:147-203): large_scale_p_gradient
 * (:1982-2011) with the horizontal means of u and v, the large-scale
 * entropy and moisture tendencies (s_tendency_c, entropies.h:17-21), then
 * subsidence of s, qt, u, v. */
void orc_forcing_bomex(const orc_dims *d, const orc_ref *rs, const double *prof, const double *u, const double *v,
                       const double *s, const double *qt, const double *T, const double *ql, double *ut, double *vt,
                       double *stend, double *qtt, double coriolis, double u0, double v0) {
    long st[3];
    strides(d, st);
    const long gw = d->gw, n2 = d->nlg[2];
    const double *ug = prof, *vg = prof + n2, *dtdt = prof + 2 * n2, *dqtdt = prof + 3 * n2, *sub = prof + 4 * n2;
    double *umean = calloc(n2, sizeof(double)), *vmean = calloc(n2, sizeof(double));
    orc_horizontal_mean(d, u, umean);
    orc_horizontal_mean(d, v, vmean);
#pragma omp parallel for schedule(static)
    for (long i = gw; i < d->nlg[0] - gw; ++i)
        for (long j = gw; j < d->nlg[1] - gw; ++j)
            for (long k = gw; k < n2 - gw; ++k) {
                const long ijk = i * st[0] + j * st[1] + k;
                const double u_at_v = umean[k] + u0, v_at_u = vmean[k] + v0;
                ut[ijk] = ut[ijk] - coriolis * (vg[k] - v_at_u);
                vt[ijk] = vt[ijk] + coriolis * (ug[k] - u_at_v);
            }
#pragma omp parallel for schedule(static)
    for (long i = gw; i < d->nlg[0] - gw; ++i)
        for (long j = gw; j < d->nlg[1] - gw; ++j)
            for (long k = gw; k < n2 - gw; ++k) {
                const long ijk = i * st[0] + j * st[1] + k;
                const double p0 = rs->p0_half[k], q = qt[ijk], qv = q - ql[ijk], t = T[ijk];
                const double pv = pv_c(p0, q, qv), pd = pd_c(p0, q, qv);
                stend[ijk] += cpm_c(q) * dtdt[k] / t + (sv_c(pv, t) - sd_c(pd, t)) * dqtdt[k];
                qtt[ijk] += dqtdt[k];
            }
    apply_subsidence(d, sub, s, stend);
    apply_subsidence(d, sub, qt, qtt);
    apply_subsidence(d, sub, u, ut);
    apply_subsidence(d, sub, v, vt);
    free(umean);
    free(vmean);
}

/* Rayleigh.initialize (Damping.pyx:71-93): gamma on half and full levels. */
void orc_rayleigh_profiles(const orc_dims *d, const double *zl, const double *zl_half, double gamma_r, double z_d,
                           double *gamma_z, double *gamma_zhalf) {
    const double z_top = d->dx[2] * d->n[2];
    for (long k = 0; k < d->nlg[2]; ++k) {
        gamma_zhalf[k] = 0.0;
        gamma_z[k] = 0.0;
        if (zl_half[k] >= z_top - z_d)
            gamma_zhalf[k] = gamma_r * pow(sin((C_PI / 2.0) * (1.0 - (z_top - zl_half[k]) / z_d)), 2.0);
        if (zl[k] >= z_top - z_d) gamma_z[k] = gamma_r * pow(sin((C_PI / 2.0) * (1.0 - (z_top - zl[k]) / z_d)), 2.0);
    }
}

/* Rayleigh.update (Damping.pyx:95-133) of one variable: w uses the
 * half-level profile, the others the full-level one, as the reference. */
void orc_rayleigh_update(const orc_dims *d, const double *gamma, const double *values, double *tend) {
    long st[3];
    strides(d, st);
    const long gw = d->gw, n2 = d->nlg[2];
    double *mean = calloc(n2, sizeof(double));
    orc_horizontal_mean(d, values, mean);
#pragma omp parallel for schedule(static)
    for (long i = gw; i < d->nlg[0] - gw; ++i)
        for (long j = gw; j < d->nlg[1] - gw; ++j)
            for (long k = gw; k < n2 - gw; ++k) {
                const long ijk = i * st[0] + j * st[1] + k;
                tend[ijk] -= (values[ijk] - mean[k]) * gamma[k];
            }
    free(mean);
}

/* ------------------------------------------------------------------------ */
/* State, boundaries, time stepping                                          */

/* One variable: periodic x then periodic y (build_buffer/buffer_to_values
 * through a self Sendrecv, prognostic_variables.h:12-190), then set_bcs in z
 * (prognostic_variables.h:205-270). */
void orc_update_bcs(const orc_dims *d, double *v, double bc) {
    long st[3];
    strides(d, st);
    const long gw = d->gw, n0 = d->nlg[0], n1 = d->nlg[1], n2 = d->nlg[2];
    /* x: ghost planes [0,gw) <- [n0-2gw, n0-gw), [n0-gw, n0) <- [gw, 2gw) */
    for (long i = 0; i < gw; ++i) {
        memcpy(v + i * st[0], v + (i + n0 - 2 * gw) * st[0], sizeof(double) * st[0]);
        memcpy(v + (i + n0 - gw) * st[0], v + (i + gw) * st[0], sizeof(double) * st[0]);
    }
    /* y */
    for (long i = 0; i < n0; ++i)
        for (long j = 0; j < gw; ++j) {
            memcpy(v + i * st[0] + j * st[1], v + i * st[0] + (j + n1 - 2 * gw) * st[1], sizeof(double) * n2);
            memcpy(v + i * st[0] + (j + n1 - gw) * st[1], v + i * st[0] + (j + gw) * st[1], sizeof(double) * n2);
        }
    /* z: s = +1 (top) then s = -1 (bottom) in the order of the shift loop
     * [-1, 1]: PrognosticVariables.pyx:192; the two ghost regions are disjoint. */
    for (long i = 0; i < n0; ++i)
        for (long j = 0; j < n1; ++j) {
            double *c = v + i * st[0] + j * st[1];
            /* s == -1 branch: bc_start = nlg-gw (top) */
            {
                const long b = n2 - gw;
                if (bc == 1.0) {
                    for (long k = 0; k < gw; ++k) c[b + k] = c[b - k - 1];
                } else if (bc == 2.0) {
                    for (long k = 0; k < gw; ++k) c[b + k] = c[b - 1];
                } else {
                    c[b] = 0.0;
                    for (long k = 1; k < gw; ++k) c[b + k] = bc * c[b - k];
                }
            }
            /* s == +1 branch: bc_start = gw-1 (bottom) */
            {
                const long b = gw - 1;
                if (bc == 1.0) {
                    for (long k = 0; k < gw; ++k) c[b - k] = c[b + k + 1];
                } else if (bc == 2.0) {
                    for (long k = 0; k < gw; ++k) c[b - k] = c[b + 1];
                } else {
                    c[b] = 0.0;
                    for (long k = 1; k < gw; ++k) c[b - k] = bc * c[b + k];
                }
            }
        }
}

/* TimeStepping.update_third, TimeStepping.pyx:131-151 */
void orc_rk3_update(long n, int rk_step, double dt, double *values, double *tend, double *copy) {
    if (rk_step == 0) {
#pragma omp parallel for schedule(static)
        for (long i = 0; i < n; ++i) {
            copy[i] = values[i];
            values[i] += tend[i] * dt;
            tend[i] = 0.0;
        }
    } else if (rk_step == 1) {
#pragma omp parallel for schedule(static)
        for (long i = 0; i < n; ++i) {
            values[i] = 0.75 * copy[i] + 0.25 * (values[i] + tend[i] * dt);
            tend[i] = 0.0;
        }
    } else {
#pragma omp parallel for schedule(static)
        for (long i = 0; i < n; ++i) {
            values[i] = (1.0 / 3.0) * copy[i] + (2.0 / 3.0) * (values[i] + tend[i] * dt);
            tend[i] = 0.0;
        }
    }
}

/* HorizontalMean: interior i, j; all k; sum in (i, j) loop order, times 1/(n0*n1). */
void orc_horizontal_mean(const orc_dims *d, const double *values, double *mean) {
    long st[3];
    strides(d, st);
    const long gw = d->gw;
    for (long k = 0; k < d->nlg[2]; ++k) mean[k] = 0.0;
    for (long i = gw; i < d->nlg[0] - gw; ++i)
        for (long j = gw; j < d->nlg[1] - gw; ++j)
            for (long k = 0; k < d->nlg[2]; ++k) mean[k] += values[i * st[0] + j * st[1] + k];
    const double inv = 1.0 / (double)(d->n[1] * d->n[0]);
    for (long k = 0; k < d->nlg[2]; ++k) mean[k] = mean[k] * inv;
}

/* compute_cfl_max, TimeStepping.pyx:222-263 (no sedimentation velocities) */
double orc_cfl_max(const orc_dims *d, const double *u, const double *v, const double *w,
                   double dt) {
    long st[3];
    strides(d, st);
    const long gw = d->gw;
    double cfl = -9999.0;
    for (long i = gw; i < d->nlg[0] - gw; ++i)
        for (long j = gw; j < d->nlg[1] - gw; ++j)
            for (long k = gw; k < d->nlg[2] - gw; ++k) {
                const long ijk = i * st[0] + j * st[1] + k;
                const double wa = fabs(w[ijk]);
                cfl = fmax(cfl, dt * (fabs(u[ijk]) * d->dxi[0] + fabs(v[ijk]) * d->dxi[1] + wa * d->dxi[2]));
            }
    return cfl + 1e-11;
}

/* ------------------------------------------------------------------------ */
/* Pressure                                                                  */

/* SparseSolvers.pyx:24-43 */
void orc_tdma_solve(long n, double *x, const double *a, const double *b, const double *c,
                    double *scratch) {
    scratch[0] = c[0] / b[0];
    x[0] = x[0] / b[0];
    for (long i = 1; i < n; ++i) {
        const double m = 1.0 / (b[i] - a[i] * scratch[i - 1]);
        scratch[i] = c[i] * m;
        x[i] = (x[i] - a[i] * x[i - 1]) * m;
    }
    for (long i = n - 2; i >= 0; --i) x[i] = x[i] - scratch[i] * x[i + 1];
}

static int is_pow2(long n) { return n > 0 && (n & (n - 1)) == 0; }

/* Iterative radix-2 decimation-in-time FFT (powers of two) or a direct DFT
 * otherwise; both evaluate sum_n x[n] exp(sign*2*pi*i*m*n/N). */
void orc_fft(long n, double *re, double *im, int inverse) {
    const double sign = inverse ? 1.0 : -1.0;
    if (!is_pow2(n)) {
        double *tr = (double *)malloc(sizeof(double) * n);
        double *ti = (double *)malloc(sizeof(double) * n);
        for (long m = 0; m < n; ++m) {
            double sr = 0.0, si = 0.0;
            for (long q = 0; q < n; ++q) {
                const long e = (m * q) % n;
                const double ang = sign * 2.0 * C_PI * (double)e / (double)n;
                sr += re[q] * cos(ang) - im[q] * sin(ang);
                si += re[q] * sin(ang) + im[q] * cos(ang);
            }
            tr[m] = sr;
            ti[m] = si;
        }
        memcpy(re, tr, sizeof(double) * n);
        memcpy(im, ti, sizeof(double) * n);
        free(tr);
        free(ti);
    } else {
        /* bit reversal */
        for (long i = 1, j = 0; i < n; ++i) {
            long bit = n >> 1;
            for (; j & bit; bit >>= 1) j ^= bit;
            j ^= bit;
            if (i < j) {
                double t = re[i]; re[i] = re[j]; re[j] = t;
                t = im[i]; im[i] = im[j]; im[j] = t;
            }
        }
        for (long len = 2; len <= n; len <<= 1) {
            const long half = len >> 1;
            for (long m = 0; m < half; ++m) {
                const double ang = sign * 2.0 * C_PI * (double)m / (double)len;
                const double wr = cos(ang), wi = sin(ang);
                for (long s = 0; s < n; s += len) {
                    const long a = s + m, b = s + m + half;
                    const double xr = re[b] * wr - im[b] * wi;
                    const double xi = re[b] * wi + im[b] * wr;
                    re[b] = re[a] - xr;
                    im[b] = im[a] - xi;
                    re[a] = re[a] + xr;
                    im[a] = im[a] + xi;
                }
            }
        }
    }
    if (inverse) {
        const double inv = 1.0 / (double)n;
        for (long i = 0; i < n; ++i) {
            re[i] *= inv;
            im[i] *= inv;
        }
    }
}

/* second_order_divergence x3 after zeroing (PressureSolver.pyx:56-65,138-175) */
void orc_divergence(const orc_dims *d, const orc_ref *rs, const double *u, const double *v,
                    const double *w, double *div) {
    long st[3];
    strides(d, st);
    const long gw = d->gw;
    memset(div, 0, sizeof(double) * d->npg);
    const double *vel[3] = {u, v, w};
    for (int dir = 0; dir < 3; ++dir) {
        const long sm1 = -st[dir];
        const double dxi = 1.0 / d->dx[dir];
        const double *q = vel[dir];
#pragma omp parallel for schedule(static)
        for (long i = gw; i < d->nlg[0] - gw; ++i)
            for (long j = gw; j < d->nlg[1] - gw; ++j)
                for (long k = gw; k < d->nlg[2] - gw; ++k) {
                    const long ijk = i * st[0] + j * st[1] + k;
                    if (dir != 2)
                        div[ijk] += (q[ijk] / rs->alpha0_half[k] - q[ijk + sm1] / rs->alpha0_half[k]) * dxi;
                    else
                        div[ijk] += ((q[ijk]) / rs->alpha0[k] - (q[ijk + sm1]) / rs->alpha0[k - 1]) * dxi;
                }
    }
}

/* compute_modified_wave_numbers, PressureFFTParallel.pyx:64-98 */
static void wave_numbers(long n, double dx, double *k2) {
    for (long i = 0; i < n; ++i) {
        double xi;
        if (i <= n / 2) xi = (double)i;
        else xi = (double)(i - n);
        k2[i] = (2.0 * cos((2.0 * C_PI / n) * xi) - 2.0) / dx / dx;
    }
    k2[0] = 0.0;
}

/* PressureSolver.update with PressureFFTParallel.solve on one rank. */
void orc_pressure_update(const orc_dims *d, const orc_ref *rs, double *u, double *v, double *w,
                         double *p, double *div) {
    long st[3];
    strides(d, st);
    const long gw = d->gw, n0 = d->n[0], n1 = d->n[1], n2 = d->n[2];
    const long L0 = d->nlg[0], L1 = d->nlg[1], L2 = d->nlg[2];
    /* remove_mean_u3 (PressureSolver.pyx:113-136) */
    double *mean = (double *)malloc(sizeof(double) * L2);
    orc_horizontal_mean(d, w, mean);
    for (long i = 0; i < L0; ++i)
        for (long j = 0; j < L1; ++j)
            for (long k = 0; k < L2; ++k) w[i * st[0] + j * st[1] + k] = w[i * st[0] + j * st[1] + k] - mean[k];
    free(mean);
    orc_divergence(d, rs, u, v, w, div);

    /* solve: FFT along x then y on the interior, TDMA per column, inverse. */
    const long npl = n0 * n1 * n2;
    double *re = (double *)calloc(npl, sizeof(double));
    double *im = (double *)calloc(npl, sizeof(double));
    /* interior layout here: [i][j][k] */
#define IX(i, j, k) (((i) * n1 + (j)) * n2 + (k))
    for (long i = 0; i < n0; ++i)
        for (long j = 0; j < n1; ++j)
            for (long k = 0; k < n2; ++k) re[IX(i, j, k)] = div[(i + gw) * st[0] + (j + gw) * st[1] + k + gw];
    /* x transforms (X pencils) */
#pragma omp parallel
    {
        double *tr = (double *)malloc(sizeof(double) * (n0 > n1 ? n0 : n1));
        double *ti = (double *)malloc(sizeof(double) * (n0 > n1 ? n0 : n1));
#pragma omp for schedule(static)
        for (long jk = 0; jk < n1 * n2; ++jk) {
            const long j = jk / n2, k = jk % n2;
            for (long i = 0; i < n0; ++i) { tr[i] = re[IX(i, j, k)]; ti[i] = im[IX(i, j, k)]; }
            orc_fft(n0, tr, ti, 0);
            for (long i = 0; i < n0; ++i) { re[IX(i, j, k)] = tr[i]; im[IX(i, j, k)] = ti[i]; }
        }
#pragma omp for schedule(static)
        for (long ik = 0; ik < n0 * n2; ++ik) {
            const long i = ik / n2, k = ik % n2;
            for (long j = 0; j < n1; ++j) { tr[j] = re[IX(i, j, k)]; ti[j] = im[IX(i, j, k)]; }
            orc_fft(n1, tr, ti, 0);
            for (long j = 0; j < n1; ++j) { re[IX(i, j, k)] = tr[j]; im[IX(i, j, k)] = ti[j]; }
        }
        free(tr);
        free(ti);
    }
    /* TDMA per (i, j) column: compute_off_diagonals / compute_diagonal
     * (PressureFFTParallel.pyx:100-148), then 2x TDMA.solve and the (0,0) zero. */
    double *kx2 = (double *)malloc(sizeof(double) * n0);
    double *ky2 = (double *)malloc(sizeof(double) * n1);
    wave_numbers(n0, d->dx[0], kx2);
    wave_numbers(n1, d->dx[1], ky2);
    double *a = (double *)malloc(sizeof(double) * n2);
    double *c = (double *)malloc(sizeof(double) * n2);
    const double dzi2 = d->dxi[2] * d->dxi[2];
    a[0] = 0.0;
    c[0] = d->dxi[2] * d->dxi[2] * rs->rho0[gw];
    for (long k = 1; k < n2 - 1; ++k) {
        a[k] = d->dxi[2] * d->dxi[2] * rs->rho0[k + gw - 1];
        c[k] = d->dxi[2] * d->dxi[2] * rs->rho0[k + gw];
    }
    a[n2 - 1] = d->dxi[2] * d->dxi[2] * rs->rho0[n2 - 1 + gw - 1];
    c[n2 - 1] = 0.0;
    (void)dzi2;
#pragma omp parallel
    {
        double *b = (double *)malloc(sizeof(double) * n2);
        double *dr = (double *)malloc(sizeof(double) * n2);
        double *di = (double *)malloc(sizeof(double) * n2);
        double *scr = (double *)malloc(sizeof(double) * n2);
#pragma omp for schedule(static)
        for (long ij = 0; ij < n0 * n1; ++ij) {
            const long i = ij / n1, j = ij % n1;
            const double kk = kx2[i] + ky2[j];
            b[0] = (rs->rho0_half[gw] * kk - (rs->rho0[gw]) * d->dxi[2] * d->dxi[2]);
            for (long k = 1; k < n2 - 1; ++k)
                b[k] = (rs->rho0_half[k + gw] * kk - (rs->rho0[k + gw] + rs->rho0[k + gw - 1]) * d->dxi[2] * d->dxi[2]);
            b[n2 - 1] = (rs->rho0_half[n2 - 1 + gw] * kk - (rs->rho0[n2 - 1 + gw - 1]) * d->dxi[2] * d->dxi[2]);
            for (long k = 0; k < n2; ++k) { dr[k] = re[IX(i, j, k)]; di[k] = im[IX(i, j, k)]; }
            if (i != 0 || j != 0) {
                orc_tdma_solve(n2, dr, a, b, c, scr);
                orc_tdma_solve(n2, di, a, b, c, scr);
                for (long k = 0; k < n2; ++k) { re[IX(i, j, k)] = dr[k]; im[IX(i, j, k)] = di[k]; }
            } else {
                for (long k = 0; k < n2; ++k) { re[IX(i, j, k)] = 0.0; im[IX(i, j, k)] = 0.0; }
            }
        }
        free(b); free(dr); free(di); free(scr);
    }
    free(a); free(c); free(kx2); free(ky2);
    /* inverse: y then x */
#pragma omp parallel
    {
        double *tr = (double *)malloc(sizeof(double) * (n0 > n1 ? n0 : n1));
        double *ti = (double *)malloc(sizeof(double) * (n0 > n1 ? n0 : n1));
#pragma omp for schedule(static)
        for (long ik = 0; ik < n0 * n2; ++ik) {
            const long i = ik / n2, k = ik % n2;
            for (long j = 0; j < n1; ++j) { tr[j] = re[IX(i, j, k)]; ti[j] = im[IX(i, j, k)]; }
            orc_fft(n1, tr, ti, 1);
            for (long j = 0; j < n1; ++j) { re[IX(i, j, k)] = tr[j]; im[IX(i, j, k)] = ti[j]; }
        }
#pragma omp for schedule(static)
        for (long jk = 0; jk < n1 * n2; ++jk) {
            const long j = jk / n2, k = jk % n2;
            for (long i = 0; i < n0; ++i) { tr[i] = re[IX(i, j, k)]; ti[i] = im[IX(i, j, k)]; }
            orc_fft(n0, tr, ti, 1);
            for (long i = 0; i < n0; ++i) { re[IX(i, j, k)] = tr[i]; im[IX(i, j, k)] = ti[i]; }
        }
        free(tr);
        free(ti);
    }
    /* DV.p = Re(pres): ghosts 0, interior from the solve (:215-218) */
    memset(p, 0, sizeof(double) * d->npg);
    for (long i = 0; i < n0; ++i)
        for (long j = 0; j < n1; ++j)
            for (long k = 0; k < n2; ++k) p[(i + gw) * st[0] + (j + gw) * st[1] + k + gw] = re[IX(i, j, k)];
#undef IX
    free(re);
    free(im);
    /* communicate_variable(p): periodic x/y, sym z (DiagnosticVariables.pyx:65-112) */
    orc_update_bcs(d, p, 1.0);
    /* second_order_pressure_correction over [0, nlg-1)^3 (PressureSolver.pyx:83-110) */
#pragma omp parallel for schedule(static)
    for (long i = 0; i < L0 - 1; ++i)
        for (long j = 0; j < L1 - 1; ++j)
            for (long k = 0; k < L2 - 1; ++k) {
                const long ijk = i * st[0] + j * st[1] + k;
                u[ijk] -= (p[ijk + st[0]] - p[ijk]) * d->dxi[0];
                v[ijk] -= (p[ijk + st[1]] - p[ijk]) * d->dxi[1];
                w[ijk] -= (p[ijk + 1] - p[ijk]) * d->dxi[2];
            }
    /* PV.update_all_bcs: u, v sym, w asym */
    orc_update_bcs(d, u, 1.0);
    orc_update_bcs(d, v, 1.0);
    orc_update_bcs(d, w, -1.0);
}
