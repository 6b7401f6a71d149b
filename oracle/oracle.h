/*
 * oracle.h -- FP64 CPU restatement of the PyCLES per-timestep dynamical core.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker for the HIP kernels in
 * pycles_amd/.  Only tests/, __graft_entry__.smoke() and the cpu_baseline leg
 * of bench.py may load it.  The product path never calls it.
 *
 * Layout (same as the reference): one variable is a flat C-order array
 * [i][j][k] over the ghosted local extents nlg = n + 2*gw, z fastest
 * (reference Csrc/flux_divergence.h:15-16, PrognosticVariables.pxd:45-46).
 * Single-rank semantics: x and y are periodic (reference
 * ParallelMPI.pyx:54-59 with one rank), z uses set_bcs.
 *
 * Every loop range and every arithmetic expression follows the cited
 * reference line so that results are bit-comparable with the reference's
 * own C compiled with -ffp-contract=off.
 */
#ifndef PYCLES_ORACLE_H
#define PYCLES_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Mirror of reference Csrc/grid.h:2-24 restricted to what the path reads. */
typedef struct orc_dims {
    long n[3];     /* interior points */
    long nlg[3];   /* n + 2*gw */
    long gw;
    long npg;      /* nlg0*nlg1*nlg2 */
    double dx[3];
    double dxi[3];
} orc_dims;

/* Reference Csrc/lookup.h:4-14 (fields used by lookup()). */
typedef struct orc_lookup {
    long n;
    double x_min;
    double dx;
    double dxi;
    const double *y;
} orc_lookup;

/* Reference-state profiles, each of length nlg[2] (ReferenceState.pyx:20-25). */
typedef struct orc_ref {
    const double *p0, *p0_half, *alpha0, *alpha0_half, *rho0, *rho0_half;
} orc_ref;

/* Saturation-adjustment telemetry (the reference has none; the secant loop
 * of thermodynamics_sa.h:44-59 is uncapped). */
typedef struct orc_sa_stats {
    long saturated;      /* cells that entered the secant branch */
    long iterations;     /* total secant iterations */
    long max_iterations; /* largest per-cell count */
    long capped;         /* cells that hit the iteration cap */
} orc_sa_stats;

void orc_dims_init(orc_dims *d, long nx, long ny, long nz, long gw,
                   double dx, double dy, double dz);

/* ---- interpolation (advection_interpolation.h) ---- */
double orc_interp_2(double a, double b);
double orc_interp_4(const double *q);              /* q[0..3] = phim1..phip2 */
double orc_interp_6(const double *q);              /* q[0..5] */
double orc_interp_8(const double *q);              /* q[0..7] */
double orc_interp_weno5(const double *q);          /* q[0..4] = phim2..phip2 */
double orc_interp_weno7(const double *q);          /* q[0..6] */
double orc_interp_weno9(const double *q);          /* q[0..8] */

/* ---- advection ---- */
void orc_scalar_flux(const orc_dims *d, const double *rho0, const double *rho0_half,
                     const double *vel, const double *phi, double *flux, int dir, int order);
void orc_scalar_flux_divergence(const orc_dims *d, const double *alpha0,
                                const double *alpha0_half, const double *flux,
                                double *tend, int dir);
/* ScalarAdvection.update for nsc scalars (ScalarAdvection.pyx:69-123).
 * vel[3] = u, v, w; flux_scratch holds 3*npg doubles (reused per scalar). */
void orc_scalar_advection_update(const orc_dims *d, const orc_ref *rs, const double *const *vel,
                                 int nsc, const double *const *phi, double *const *tend,
                                 double *flux_scratch, int order);
/* One (advected, advecting) pair of weno_*_order_m + momentum_flux_divergence. */
void orc_momentum_pair(const orc_dims *d, const orc_ref *rs, const double *vel_advected,
                       const double *vel_advecting, double *tend, int d_advected,
                       int d_advecting, int order, double *flux_scratch);
/* MomentumAdvection.update (MomentumAdvection.pyx:41-69). */
void orc_momentum_advection_update(const orc_dims *d, const orc_ref *rs, const double *const *vel,
                                   double *const *tend, int order, double *flux_scratch);

/* ---- thermodynamics ---- */
double orc_latent_heat_variable(double T);
double orc_lookup_eval(const orc_lookup *lt, double x);
/* eos_c (thermodynamics_sa.h:17-66) with an iteration cap (max_iter<=0: no cap).
 * Returns the number of secant iterations (0 = unsaturated branch). */
int orc_eos_sa(const orc_lookup *lt, double p0, double s, double qt, double *T, double *qv,
               double *ql, double *qi, int max_iter);
double orc_entropy_sa(double p0, double T, double qt, double ql, double qi);
double orc_alpha(double p0, double T, double qt, double qv);
double orc_exner(double p0);
/* ThermodynamicsSA.update (ThermodynamicsSA.pyx:176-213). Arrays are npg long. */
void orc_thermo_sa_update(const orc_dims *d, const orc_ref *rs, const orc_lookup *lt,
                          double *s, double *qt, double *wt, double *T, double *qv, double *ql,
                          double *qi, double *alpha, double *buoyancy, double *theta_rho,
                          double *bvf, double *thetali, int do_clip, int max_iter,
                          orc_sa_stats *stats);
/* The parts of ThermodynamicsSA.update (thermodynamics_sa.h:98-211) and
 * ThermodynamicsDry.update (thermodynamics_dry.h:39-110), over the ghosted
 * reference layout. */
void orc_buoyancy_update_sa(const orc_dims *d, const double *alpha0_half, const double *alpha,
                            double *buoyancy, double *wt);
void orc_bvf_sa(const orc_dims *d, const orc_lookup *lt, const double *p0, const double *T,
                const double *qt, const double *qv, double *theta_rho, double *bvf);
void orc_thetali_update(const orc_dims *d, const double *p0, const double *T, const double *qt,
                        const double *ql, const double *qi, double *thetali);
void orc_buoyancy_update_dry(const orc_dims *d, const double *alpha0_half, const double *alpha,
                             double *buoyancy, double *wt);
void orc_bvf_dry(const orc_dims *d, const double *p0, const double *T, double *theta, double *bvf);
double orc_eos_dry(double pd, double s);
double orc_entropy_dry(double pd, double T);
/* ThermodynamicsDry.update (ThermodynamicsDry.pyx:105-120). */
void orc_thermo_dry_update(const orc_dims *d, const orc_ref *rs, double *s, double *wt,
                           double *T, double *alpha, double *buoyancy, double *theta,
                           double *bvf);

/* ---- state, boundaries, time stepping ---- */
/* PrognosticVariables.update_all_bcs for one variable on one rank
 * (PrognosticVariables.pyx:183-233 + prognostic_variables.h:205-270). */
/* ---- kinematics, SGS and diffusion ---- */
void orc_kinematics_update(const orc_dims *d, const double *const *vel, double *vgrad, double *strain,
                           double *smag);
void orc_smagorinsky_update(const orc_dims *d, double *visc, double *diff, const double *bvf, const double *smag,
                            double cs, double prt);
void orc_scalar_diffusion_update(const orc_dims *d, const orc_ref *rs, const double *diffusivity, int nsc,
                                 const double *const *phi, double *const *tend, double *flux);
void orc_momentum_diffusion_update(const orc_dims *d, const orc_ref *rs, const double *strain,
                                   const double *visc, double *const *tend, double *flux, const double *smag,
                                   const double *T, double *s_tend);
/* ---- BOMEX case physics ---- */
double orc_entropyflux_from_thetaflux_qtflux(double thetaflux, double qtflux, double p0_b, double T_b, double qt_b,
                                             double qv_b);
void orc_surface_bomex(const orc_dims *d, const orc_ref *rs, const double *u, const double *v, const double *qt,
                       const double *T, const double *qv, const double *ql, double *ut, double *vt, double *st,
                       double *qtt, double u0, double v0, double t_mean_gw, int lh_constant, double *flux2d);
void orc_forcing_bomex_profiles(const orc_dims *d, const double *zl_half, const double *p0_half, double *prof);
void orc_forcing_bomex(const orc_dims *d, const orc_ref *rs, const double *prof, const double *u, const double *v,
                       const double *s, const double *qt, const double *T, const double *ql, double *ut, double *vt,
                       double *stend, double *qtt, double coriolis, double u0, double v0);
void orc_rayleigh_profiles(const orc_dims *d, const double *zl, const double *zl_half, double gamma_r, double z_d,
                           double *gamma_z, double *gamma_zhalf);
void orc_rayleigh_update(const orc_dims *d, const double *gamma, const double *values, double *tend);
void orc_update_bcs(const orc_dims *d, double *values, double bc_factor);
/* TimeStepping.update_third (TimeStepping.pyx:131-151) over n = npg*nv points. */
void orc_rk3_update(long n, int rk_step, double dt, double *values, double *tend,
                    double *copy);
/* ParallelMPI.HorizontalMean (ParallelMPI.pyx:196-241), mean has nlg[2] entries. */
void orc_horizontal_mean(const orc_dims *d, const double *values, double *mean);
/* TimeStepping.compute_cfl_max (TimeStepping.pyx:222-263), without sedimentation
 * velocities; returns cfl_max including the +1e-11. */
double orc_cfl_max(const orc_dims *d, const double *u, const double *v, const double *w,
                   double dt);

/* ---- pressure ---- */
/* Thomas algorithm, SparseSolvers.pyx:24-43. scratch has n entries. */
void orc_tdma_solve(long n, double *x, const double *a, const double *b, const double *c,
                    double *scratch);
/* Complex DFT along a contiguous row of length n, sign -1 forward (unnormalised),
 * +1 inverse (normalised by 1/n), matching scipy.fftpack.fft / ifft. */
void orc_fft(long n, double *re, double *im, int inverse);
/* PressureSolver.update (PressureSolver.pyx:38-81) with PressureFFTParallel.solve
 * (PressureFFTParallel.pyx:150-220) on one rank. bc[3] are the u,v,w bc factors.
 * div and p are npg arrays (DV divergence and dynamic_pressure). */
void orc_pressure_update(const orc_dims *d, const orc_ref *rs, double *u, double *v, double *w,
                         double *p, double *div);
/* Projected divergence (second_order_divergence of PressureSolver.pyx:138-175
 * applied to the current velocities) into div (interior only). */
void orc_divergence(const orc_dims *d, const orc_ref *rs, const double *u, const double *v,
                    const double *w, double *div);

#ifdef __cplusplus
}
#endif
#endif
