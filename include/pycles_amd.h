/*
 * pycles_amd.h -- C ABI of the MI355X (gfx950) dynamical core of PyCLES.
 *
 * Plain C: pointers, sizes and a hipStream_t passed as void*.  All field
 * pointers are DEVICE pointers.  Every entry point returns 0 on success or a
 * negative PYCLES_ERR_* code; pycles_last_error() gives the message.  No entry
 * point allocates device memory except the *_create functions, which run once
 * at set-up (the reference allocates a scratch flux per momentum call,
 * momentum_advection_weno.h:119/243; nothing here does).
 *
 * Device layout of one field ("device layout"): [k][j][i], x fastest, over
 * nzg = n[2] + 2*gw levels and the n[0]*n[1] interior columns.  x and y are
 * periodic by index wrap (the reference keeps gw ghost columns there, filled
 * by a self Sendrecv on one rank, PrognosticVariables.pyx:183-233); the gw
 * ghost levels in z are kept and filled as the reference's set_bcs does
 * (prognostic_variables.h:205-270).  pycles_layout_ingress/egress convert to
 * and from the reference's ghosted [i][j][k] array.
 *
 * Precision tags: PYCLES_F64 -- fields are double and all arithmetic is
 * FP64, reproducing the reference's operation order; PYCLES_F32 -- fields
 * are float, bulk arithmetic FP32 (profiles and reductions stay FP64).
 *
 * Reference profile arrays (rho0, rho0_half, alpha0, alpha0_half, p0,
 * p0_half) are FP64 device arrays of length nzg, exactly the reference's
 * ReferenceState arrays (ReferenceState.pyx:20-25).
 */
#ifndef PYCLES_AMD_H
#define PYCLES_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PYCLES_OK 0
#define PYCLES_ERR_ARG -1
#define PYCLES_ERR_HIP -2
#define PYCLES_ERR_UNSUPPORTED -3
#define PYCLES_ERR_NONFINITE -4

#define PYCLES_F64 0
#define PYCLES_F32 1

/* Latent heat models replacing the reference's C function pointers
 * double (*L_fp)(double, double) (Microphysics.pxd:65-75). */
#define PYCLES_LH_VARIABLE 0
#define PYCLES_LH_CONSTANT 1

/* Mirror of the reference struct DimStruct (Csrc/grid.h:2-24) for one GPU
 * owning the whole domain: n = interior points, gw = ghost width. */
typedef struct pycles_dims {
    int64_t n[3];
    int64_t gw;
    double dx[3];
    double dxi[3];
} pycles_dims;

/* Reference-state profiles (FP64 device arrays of length nzg). */
typedef struct pycles_ref {
    const double *rho0, *rho0_half, *alpha0, *alpha0_half, *p0, *p0_half;
} pycles_ref;

/* FP32 deviation storage of the scalars.  In PYCLES_F32 the prognostic
 * scalars are stored as phi' = phi - base(k) with base an FP64 profile of
 * length nzg (mirrored in the ghost levels); NULL base = stored as is.
 * Ignored in PYCLES_F64. */
typedef struct pycles_scalar_base {
    const double *base[2];
} pycles_scalar_base;

const char *pycles_last_error(void);
const char *pycles_version(void);
int pycles_device_count(void);

/* ---- advection --------------------------------------------------------- */
/* Replaces ScalarAdvection.update (ScalarAdvection.pyx:69-123) and
 * MomentumAdvection.update (MomentumAdvection.pyx:41-69), i.e.
 * compute_advective_fluxes_a + scalar_flux_divergence
 * (scalar_advection.h:689-727, flux_divergence.h:4-32) and
 * compute_advective_tendencies_m (momentum_advection.h:13-93).
 * vel[3] = u, v, w; tend_vel[3] = their tendencies (NULL = no momentum
 * advection); scal/tend_scal hold nscal (0, 1 or 2) scalars.
 * Tendencies are accumulated (tend -= div(flux)), per variable in the
 * reference order x, y, z.  Orders 5, 7, 9 (WENO5/7/9 with centred
 * interpolation of order 4/6/8 for the advecting velocity); gw must be at
 * least (order+1)/2.
 * flags: PYCLES_ADV_TEND_ZERO declares that the scalar, u and v tendencies
 * are zero on entry (the w tendency is still read: it carries buoyancy); they
 * are then written instead of read-modified-written, with identical bits.
 * PYCLES_ADV_PASS_X/Y/Z: when any is set, only the selected directional
 * passes run (for timing one kernel; a partial call is not a full update).
 * The y and z passes run as one fused kernel when both are selected, the
 * orders are 5 and 5 and the call has momentum and at least one scalar: by
 * default for FP32 storage; PYCLES_ADV_FUSE_YZ asks for it in FP64 as well
 * (same bits as the separate passes; for tests, it is slower there). */
#define PYCLES_ADV_TEND_ZERO 1
#define PYCLES_ADV_PASS_X 2
#define PYCLES_ADV_PASS_Y 4
#define PYCLES_ADV_PASS_Z 8
#define PYCLES_ADV_FUSE_YZ 16
int pycles_advection(const pycles_dims *dims, const pycles_ref *ref, const void *const vel[3],
                     void *const tend_vel[3], int nscal, const void *const *scal,
                     void *const *tend_scal, const pycles_scalar_base *base, int scalar_order,
                     int momentum_order, int precision, int flags, void *stream);
/* pycles_advection followed by the RK3 stage in the same sweep: TimeStepping.
 * update_third (TimeStepping.pyx:131-151) and the z boundary fill of
 * PrognosticVariables.update_all_bcs (PrognosticVariables.pyx:183-233;
 * sym for the scalars, u and v, asym for w).  The z pass applies the stage to
 * each cell as soon as its tendency is final and writes the new value, and
 * the ghost levels it feeds, to out_* (arrays distinct from the inputs; the
 * caller swaps them in).  copy_* are the stage copies (written at rk_step 0,
 * read at 1 and 2).  At rk_step 0 they may be NULL: no copy is written, and
 * the caller keeps that stage's input unchanged and passes it as the copy of
 * stages 1 and 2; at rk_step 2 out_* may be that same array (each cell's copy
 * is read before its output is written).  The tendency arrays keep the x and
 * y contributions.
 * Same bits as pycles_advection then pycles_rk3_stage.  Needs the momentum. */
int pycles_advection_rk3(const pycles_dims *dims, const pycles_ref *ref, const void *const vel[3],
                         void *const tend_vel[3], int nscal, const void *const *scal,
                         void *const *tend_scal, const pycles_scalar_base *base, int scalar_order,
                         int momentum_order, int precision, int flags, void *const out_vel[3],
                         void *const *out_scal, void *const copy_vel[3], void *const *copy_scal,
                         int rk_step, double dt, void *stream);

/* ---- kinematics, SGS and diffusion ------------------------------------- */
/* Replaces Kinematics.update (Kinematics.pyx:41-58: compute_velocity_gradient,
 * compute_strain_rate, compute_strain_rate_mag, kinematics.h:6-138) and
 * Smagorinsky.update (SGS.pyx:146-152: smagorinsky_update, sgs.h:4-19) in one
 * pass: the strain-rate magnitude |S| (0 on the bottom and top ghost
 * levels, as the reference leaves it), the eddy viscosity
 * nu = (cs Delta)^2 |S| (times sqrt(max(1 - N^2/(prt |S|^2), 0)) where
 * N^2 > 0, Delta = (dx dy dz)^(1/3)) and the diffusivity nu/prt, on every
 * stored point.  vel = u, v, w; bvf = N^2 (the thermodynamics output);
 * all fields in the given precision. */
int pycles_sgs_smagorinsky(const pycles_dims *dims, const void *const vel[3], const void *bvf, double cs,
                           double prt, void *strain_mag, void *visc, void *diff, int precision, void *stream);
/* flags of pycles_diffusion: PYCLES_DIFF_SCALARS replaces
 * ScalarDiffusion.update (ScalarDiffusion.pyx:76-119: second_order_diffusion,
 * scalar_diffusion.h:6-58, and scalar_flux_divergence) for the nscal (1..2)
 * scalars with the diffusivity diff; PYCLES_DIFF_MOMENTUM the stress part of
 * MomentumDiffusion.update (MomentumDiffusion.pyx:44-66:
 * compute_diffusive_flux_m, momentum_diffusion.h:5-88, and
 * momentum_flux_divergence) with the viscosity visc, and
 * PYCLES_DIFF_ENTROPY_SOURCE its compute_entropy_source (:90-118) into the
 * first scalar's tendency from |S|, visc and the temperature.  One pass does
 * the selected parts in that order; tendencies are accumulated after what
 * they hold, per variable in the reference's order.  The surface flux is
 * zero (one GPU owns the bottom); base as in pycles_advection.  gw >= 2. */
#define PYCLES_DIFF_SCALARS 1
#define PYCLES_DIFF_MOMENTUM 2
#define PYCLES_DIFF_ENTROPY_SOURCE 4
int pycles_diffusion(const pycles_dims *dims, const pycles_ref *ref, const void *const vel[3],
                     void *const tend_vel[3], int nscal, const void *const *scal, void *const *tend_scal,
                     const pycles_scalar_base *base, const void *visc, const void *diff, const void *strain_mag,
                     const void *temperature, int flags, int precision, void *stream);

/* ---- BOMEX case physics ------------------------------------------------ */
/* Replaces SurfaceBomex.update with the moist branch of SurfaceBase.update
 * (Surface.pyx:297-373, 100-183; entropyflux_from_thetaflux_qtflux and
 * compute_windspeed, surface.h:95-106, 157-185): theta flux 8e-3 K m/s, qt
 * flux 5.2e-5 m/s, u* = 0.28 m/s, no gustiness; the fluxes of u, v, s, qt are
 * added to tend[0..3] at the first interior level times
 * alpha0_half[gw]/alpha0[gw-1]/dz.  qt, T, qv, ql at the first level;
 * base: the FP32 profile of qt (base->base[1]).  flux2d (device FP64
 * [8][nx*ny], may be NULL) receives the column fluxes u, v, s, qt and the
 * diagnostics shf, lhf, buoyancy flux and Obukhov length, for which t_mean
 * (the level means of T, device FP64 [nzg]) is read. */
int pycles_surface_bomex(const pycles_dims *dims, const pycles_ref *ref, const void *u, const void *v,
                         const void *qt, const void *T, const void *qv, const void *ql,
                         const pycles_scalar_base *base, void *const tend[4], double u0, double v0,
                         const double *t_mean, double *flux2d, int precision, void *stream);
/* Replaces ForcingBomex.update (Forcing.pyx:147-203): the geostrophic
 * pressure gradient with Coriolis (large_scale_p_gradient, :1982-2011) from
 * the level means uv_mean = [2][nzg] of u and v, the large-scale entropy and
 * moisture tendencies (s_tendency_c, entropies.h:17-21) and the subsidence of
 * s, qt, u, v (apply_subsidence, Forcing.pyx:2013-2040); profiles = [5][nzg]
 * ug, vg, dT/dt, dqt/dt, subsidence velocity as ForcingBomex.initialize
 * builds them (Forcing.pyx:105-145).  tend = u, v, s, qt tendencies. */
int pycles_forcing_bomex(const pycles_dims *dims, const pycles_ref *ref, const double *profiles,
                         const double *uv_mean, const void *u, const void *v, const void *s, const void *qt,
                         const void *T, const void *ql, const pycles_scalar_base *base, void *const tend[4],
                         double coriolis, double u0, double v0, int precision, void *stream);
/* Replaces Rayleigh.update (Damping.pyx:95-133) for nv (<= 8) fields:
 * tend -= (phi - mean_k(phi)) gamma[v][k] over the interior, means = [nv][nzg]
 * the level means of the stored fields (pycles_level_mean), gamma[v] the
 * profile of field v (the reference uses the half-level profile for w and
 * the full-level one for the others). */
int pycles_rayleigh_damping(const pycles_dims *dims, int nv, const void *const *fields, void *const *tend,
                            const double *const *gamma, const double *means, int precision, void *stream);

/* ---- thermodynamics ---------------------------------------------------- */
/* pv*(T) lookup table (reference Csrc/lookup.h:4-54). */
typedef struct pycles_lookup {
    const double *y; /* device, n entries */
    int64_t n;
    double x_min, dx, dxi;
} pycles_lookup;

/* Saturation-adjustment counters (device int64[4]): saturated cells,
 * iterations (Newton corrections after the learned guess plus secant steps),
 * cells that hit the secant cap, and cells whose first FP32 result failed the
 * FP64 residual check (a failed guess is redone by the FP32 secant, a failed
 * FP32 secant by the FP64 secant).  May be NULL. */
typedef struct pycles_sa_options {
    int do_qt_clipping;  /* ThermodynamicsSA.pyx:60-67 */
    int max_iter;        /* secant cap (the reference loop is uncapped) */
    int latent_heat;     /* PYCLES_LH_* */
    const void *mlp;     /* packed learned first guess (pycles_mlp_pack) or NULL; PYCLES_F32 only */
    int newton_steps;    /* Newton corrections after the learned guess, 1..4 */
    long long *counters; /* device int64[4] or NULL */
    int wt_overwrite;    /* 1: the w tendency is zero on entry, write instead of add */
    double secant_tol;   /* the FP64 secant's stopping test |dT| < secant_tol (K); 0: the
                            reference's 1e-3 K.  Smaller values measure how far the
                            reference's own stopping moves a run (no reference counterpart) */
} pycles_sa_options;

/* Replaces ThermodynamicsSA.update (ThermodynamicsSA.pyx:176-213):
 * clip_qt, eos_update, buoyancy_update_sa, bvf_sa, thetali_update
 * (thermodynamics_sa.h:68-220).  s, qt are PV fields (qt clipped in place),
 * wt the w tendency (buoyancy added); the rest are DV outputs (any may be
 * NULL except T, qv, ql, alpha). */
int pycles_thermo_sa(const pycles_dims *dims, const pycles_ref *ref, const pycles_lookup *lt,
                     void *s, void *qt, void *wt, void *T, void *qv, void *ql, void *alpha,
                     void *buoyancy, void *theta_rho, void *bvf, void *thetali,
                     const pycles_scalar_base *base, const pycles_sa_options *opt, int precision,
                     void *stream);

/* Replaces ThermodynamicsDry.update (ThermodynamicsDry.pyx:105-120).
 * wt_overwrite as in pycles_sa_options. */
int pycles_thermo_dry(const pycles_dims *dims, const pycles_ref *ref, void *s, void *wt, void *T,
                      void *alpha, void *buoyancy, void *theta, void *bvf,
                      const pycles_scalar_base *base, int precision, int wt_overwrite, void *stream);

/* ---- time stepping and boundaries -------------------------------------- */
/* TimeStepping.update_third (TimeStepping.pyx:131-151) for nv fields,
 * followed by the z boundary fill of PrognosticVariables.update_all_bcs
 * (bc_type[v] = 1.0 sym, -1.0 asym).  zero_tendencies = 1 zeroes the
 * tendencies as the reference does; 0 leaves them for a caller whose next
 * stage writes them (thermodynamics with wt_overwrite, advection with
 * PYCLES_ADV_TEND_ZERO) before reading. */
int pycles_rk3_stage(const pycles_dims *dims, int nv, void *const *values, void *const *tend,
                     void *const *copies, const double *bc_type, int rk_step, double dt,
                     int precision, int zero_tendencies, void *stream);
/* z boundary fill only (set_bcs, prognostic_variables.h:205-270). */
int pycles_apply_z_bcs(const pycles_dims *dims, int nv, void *const *values,
                       const double *bc_type, int precision, void *stream);
/* HorizontalMean (ParallelMPI.pyx:196-241) of one field into an FP64
 * device array of nzg entries (every level, ghosts included). */
int pycles_level_mean(const pycles_dims *dims, const void *field, double *mean, int precision,
                      void *stream);
/* compute_cfl_max (TimeStepping.pyx:222-263) into one FP64 device value
 * (max over interior of dt*(|u|/dx+|v|/dy+|w|/dz), plus 1e-11).  A NaN or
 * infinite velocity makes the value NaN (the reference's fmax would drop a
 * NaN). */
int pycles_cfl_max(const pycles_dims *dims, const void *u, const void *v, const void *w,
                   double dt, double *out, int precision, void *stream);
/* Waits for the stream and copies a device CFL value (pycles_cfl_max or the
 * cfl_out of pycles_pressure_update) to *host_cfl: the host read of
 * TimeStepping.adjust_timestep (TimeStepping.pyx:94-98, 255-266).  Returns
 * PYCLES_ERR_NONFINITE when the value is NaN or infinite, i.e. when a
 * velocity is not finite (the reference lets NaNs run on silently). */
int pycles_cfl_read(const double *dev_cfl, double *host_cfl, void *stream);

/* ---- pressure projection ----------------------------------------------- */
typedef struct pycles_pressure_plan pycles_pressure_plan;
/* Set-up: modified wavenumbers, tridiagonal coefficients and the spectral
 * workspace (PressureFFTParallel.initialize, PressureFFTParallel.pyx:27-148).
 * nx and ny are powers of two in [8, 4096]; nz is in [2, 2048]. */
int pycles_pressure_plan_create(const pycles_dims *dims, const double *rho0_host,
                                const double *rho0_half_host, int precision,
                                pycles_pressure_plan **plan);
void pycles_pressure_plan_destroy(pycles_pressure_plan *plan);
/* Replaces PressureSolver.update (PressureSolver.pyx:38-81): remove the
 * level mean of w, divergence of rho0*u, horizontal real FFT, per-wavenumber
 * tridiagonal solve in z with the (0,0) mode set to zero, inverse FFT,
 * pressure z boundary fill, u -= grad p, boundary fill of u, v, w.
 * p (dynamic pressure) and div (may be NULL) use the field precision.
 * If cfl_out is non-NULL the CFL maximum of the corrected velocities
 * (as pycles_cfl_max with dt) is written there too. */
int pycles_pressure_update(pycles_pressure_plan *plan, const pycles_ref *ref, void *u, void *v,
                           void *w, void *p, void *div, double dt, double *cfl_out, void *stream);
/* second_order_divergence of the current velocities (interior, field precision). */
int pycles_divergence(const pycles_dims *dims, const pycles_ref *ref, const void *u,
                      const void *v, const void *w, void *div, int precision, void *stream);

/* ---- layout conversion ------------------------------------------------- */
/* The entry points a caller holding the reference's arrays uses at ingress
 * and egress (the reference indexes them as Csrc/grid.h:2-24 lays out a
 * variable of PV.values/DV.values, PrognosticVariables.pyx:77-79).
 * Reference ghosted [i][j][k] (nlg = n + 2gw in every direction, FP64) to
 * device layout in the given precision, minus base (may be NULL); the x/y
 * ghost columns of ref_layout are not read.  Both arrays are device arrays. */
int pycles_layout_ingress(const pycles_dims *dims, const double *ref_layout, void *dev,
                          const double *base, int precision, void *stream);
/* Device layout to the reference layout (x/y ghosts = periodic images),
 * plus base (may be NULL). */
int pycles_layout_egress(const pycles_dims *dims, const void *dev, double *ref_layout,
                         const double *base, int precision, void *stream);

/* ---- statistics --------------------------------------------------------- */
/* Horizontal statistics per level (ParallelMPI.HorizontalMean,
 * HorizontalMeanofSquares, HorizontalMeanofCubes, HorizontalMaximum,
 * HorizontalMinimum, HorizontalMeanConditional and
 * HorizontalMeanofSquaresConditional, ParallelMPI.pyx:196-540) in one pass.
 * For every level k in [0, nzg), over the interior columns, with
 * v = (a + base_a[k]) * (b + base_b[k]) (the second factor only when b is
 * given; a NULL base counts as 0) and m = mask (1 where mask is NULL):
 *   out[0*nzg + k] = sum v m / D,  out[1*nzg + k] = sum v^2 m / D,
 *   out[2*nzg + k] = sum v^3 m / D,
 *   out[3*nzg + k] = max(-9e12, max of v where m != 0),
 *   out[4*nzg + k] = min(9e12, min of v where m != 0),  out[5*nzg + k] = sum m,
 * with D = nx*ny (multiplied by its reciprocal, as the reference) without a
 * mask and max(sum m, 1) with one.  a, b and mask are device-layout fields of
 * the given precision; out is an FP64 device array of 6*nzg entries.  The
 * sums are FP64 in a fixed order (deterministic, not the reference's serial
 * order).  workspace: pycles_horizontal_stats_workspace_bytes(dims) bytes of
 * device memory, overwritten. */
size_t pycles_horizontal_stats_workspace_bytes(const pycles_dims *dims);
int pycles_horizontal_stats(const pycles_dims *dims, const void *a, const double *base_a, const void *b,
                            const double *base_b, const void *mask, int precision, double *out,
                            void *workspace, void *stream);
/* pycles_horizontal_stats without a workspace argument, for callers that
 * cannot allocate device memory (code under stream capture, serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  The
 * partials live in a device array inside the library (1.5 MiB per GPU, held
 * until the process ends), so nzg = nz + 2 gw must be <= 2048
 * (PYCLES_ERR_UNSUPPORTED otherwise).  The call allocates nothing.  Every
 * call on one GPU shares that array: calls that overlap on one GPU (other
 * streams, other host threads) race.  Results are bitwise those of
 * pycles_horizontal_stats. */
int pycles_horizontal_stats_noalloc(const pycles_dims *dims, const void *a, const double *base_a, const void *b,
                                    const double *base_b, const void *mask, int precision, double *out,
                                    void *stream);
/* The fields of ThermodynamicsSA.stats_io (ThermodynamicsSA.pyx:256-389) at
 * every stored point: thetas = thetas_c(s, qt), theta = T / exner(p0_half),
 * rh = pv_c(p0_half, qt, qv) / pv*(T); outputs in the field precision, any
 * of them may be NULL.  base: the FP32 storage profiles of s and qt. */
int pycles_sa_stat_fields(const pycles_dims *dims, const pycles_ref *ref, const pycles_lookup *lt,
                          const void *s, const void *qt, const void *T, const void *qv,
                          const pycles_scalar_base *base, int precision, void *thetas, void *theta,
                          void *rh, void *stream);
/* ThermodynamicsSA.liquid_stats (ThermodynamicsSA.pyx:395-492) of the
 * liquid water ql: out[k] (k < nz) = fraction of cells with ql > 0 at
 * interior level k; out[nz] = fraction of columns with any; out[nz+1],
 * out[nz+2] = cloud base and top, the lowest and highest (k + 1/2) dz with
 * a cloudy cell (99999.9 / -99999.9 if none); out[nz+3] = liquid water path,
 * the column mean of sum_k rho0_half[k] ql dz with the interior level k
 * weighted by the ghosted profile's k-th entry, as the reference indexes it.
 * out: FP64 device array of nz + 4 entries.  workspace:
 * pycles_cloud_stats_workspace_bytes(dims) bytes, overwritten.  nz <= 12287. */
size_t pycles_cloud_stats_workspace_bytes(const pycles_dims *dims);
int pycles_cloud_stats(const pycles_dims *dims, const double *rho0_half, const void *ql, int precision,
                       double *out, void *workspace, void *stream);

/* ---- launch geometry (diagnostic) -------------------------------------- */
/* z chunks of the z-chunked kernels for dims: kernel 0 = pressure
 * correction (chunks of the interior levels [gw, nz+gw), the last one taking
 * the remainder: chunk b is [gw + b*len, b+1 == n ? nz+gw : gw + (b+1)*len)),
 * 1 = advection y pass (rows [0, ny), chunk b is [b*len, min((b+1)*len, ny))),
 * 2 = advection z pass (interior levels, as 1 from gw), 3 = thermodynamics
 * (all nzg levels, as 1), 4 = the fused y-z advection pass (as 2, no chunk
 * longer than 128 levels).  No reference counterpart. */
int pycles_launch_chunks(int kernel, const pycles_dims *dims, int *nchunks, int *chunk_len);

/* ---- HBM calibration (no reference counterpart) ------------------------ */
/* Streaming kernels of known traffic, for the practical HBM ceiling and for
 * scaling the FETCH_SIZE / WRITE_SIZE counters.  mode 0 = copy src to dst
 * (bytes read and bytes written), 1 = read src (dst receives one 4-byte word
 * per block, so it holds >= 4*blocks bytes), 2 = fill dst (bytes written);
 * mode + 4 uses non-temporal loads and stores.  bytes: a positive multiple
 * of 16; src, dst: 16-byte aligned device pointers; blocks: grid of
 * 256-thread blocks, 1 .. 2^20, each thread moving four 16-byte words per
 * iteration of a grid-stride loop. */
int pycles_stream(int mode, const void *src, void *dst, long long bytes, int blocks, void *stream);

/* ---- learned first guess of the saturation adjustment -----------------
 * No counterpart in the reference: its eos_c (thermodynamics_sa.h:17-66)
 * starts the secant from the unsaturated temperature.  The network maps
 * normalised (s, qt, p0) to T with layers 3 -> 64 -> 64 -> 1 (tanh).
 * Training state lives in caller-owned device buffers: params and the two
 * Adam moments (pycles_mlp_param_count() floats each), the normalisation
 * (pycles_mlp_norm_bytes()), a workspace (pycles_mlp_workspace_bytes(batch))
 * and the packed inference block (pycles_mlp_weights_bytes(), 16-byte
 * aligned) that pycles_sa_options.mlp points to.  Training data is a device
 * array [n][4] of (s, qt, p0, T) in FP64. */

/* n training pairs from the FP64 saturation adjustment (eos_c) on
 * qt ~ U(0, 0.025), p0 ~ U(5e4, 1.02e5) Pa, T ~ U(230, 310) K; with
 * saturated_only, the same box conditioned on qt > qv*(T) (the guess is
 * only used on the saturated branch). */
int pycles_sa_training_pairs(const pycles_lookup *lt, long long n, unsigned long long seed, int max_iter,
                             int saturated_only, double *out, void *stream);
size_t pycles_mlp_param_count(void);
size_t pycles_mlp_norm_bytes(void);
size_t pycles_mlp_weights_bytes(void);
size_t pycles_mlp_workspace_bytes(int batch);
/* Glorot-uniform weights, zero biases; zeroes the Adam moments when given. */
int pycles_mlp_init(float *params, float *adam_m, float *adam_v, unsigned long long seed, void *stream);
/* Mean and standard deviation of the four columns of data. */
int pycles_mlp_normalisation(const double *data, long long n, void *norm, void *workspace, void *stream);
/* Gradient of the mean of 0.5 (y - t)^2 (normalised T) over the first n
 * samples (n a multiple of 128); *loss receives the sum of 0.5 (y - t)^2. */
int pycles_mlp_grad(const float *params, const void *norm, const double *data, long long n, float *grad,
                    double *loss, void *workspace, void *stream);
/* `epochs` passes of Adam over shuffled minibatches (n a power of two, batch a
 * multiple of 128); cosine learning rate from lr_max to lr_min over
 * total_steps (<= 0: this call's steps); step0 = Adam steps already taken.
 * epoch_loss (device, epochs doubles, may be NULL) accumulates the summed
 * 0.5 (y - t)^2 of each epoch. */
int pycles_mlp_train(float *params, float *adam_m, float *adam_v, const void *norm, const double *data,
                     long long n, int batch, int epochs, float lr_max, float lr_min, unsigned long long seed,
                     long long step0, long long total_steps, double *epoch_loss, void *workspace, void *stream);
/* FP32 forward pass with the FP32 weights: out[i] = T guess (K). */
int pycles_mlp_forward(const float *params, const void *norm, const double *data, long long n, float *out,
                       void *stream);
/* Pack params (weights rounded to bf16) and normalisation for inference. */
int pycles_mlp_pack(const float *params, const void *norm, void *weights, void *stream);
/* The thermodynamics kernel's inference path (bf16 MFMA, FP32 accumulate)
 * applied to the (s, qt, p0) columns of data. */
int pycles_mlp_infer(const void *weights, const double *data, long long n, float *out, void *stream);

#ifdef __cplusplus
}
#endif
#endif
