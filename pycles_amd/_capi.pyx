# cython: language_level=3, boundscheck=False, wraparound=False
"""Thin Cython binding of the C ABI in include/pycles_amd.h.

Device pointers cross this layer as Python ints (``tensor.data_ptr()``) and
the stream as the integer handle of a HIP stream.  Every call checks the
returned status and raises ``PyclesError`` with the library's message.  This
is the counterpart of the reference's ``cdef extern from "<header>.h"``
declarations (e.g. ScalarAdvection.pyx:23-29, MomentumAdvection.pyx:17-21,
ThermodynamicsSA.pyx:21-29).
"""
from libc.stdint cimport int64_t, uintptr_t
from libc.stdlib cimport malloc, free

cdef extern from "pycles_amd.h":
    ctypedef struct pycles_dims:
        int64_t n[3]
        int64_t gw
        double dx[3]
        double dxi[3]
    ctypedef struct pycles_ref:
        const double *rho0
        const double *rho0_half
        const double *alpha0
        const double *alpha0_half
        const double *p0
        const double *p0_half
    ctypedef struct pycles_scalar_base:
        const double *base[2]
    ctypedef struct pycles_lookup:
        const double *y
        int64_t n
        double x_min
        double dx
        double dxi
    ctypedef struct pycles_sa_options:
        int do_qt_clipping
        int max_iter
        int latent_heat
        const void *mlp
        int newton_steps
        long long *counters
        int wt_overwrite
        double secant_tol
    ctypedef struct pycles_pressure_plan:
        pass

    int PYCLES_F64
    int PYCLES_F32
    int PYCLES_LH_VARIABLE
    int PYCLES_LH_CONSTANT
    int PYCLES_ADV_TEND_ZERO
    int PYCLES_ADV_PASS_X
    int PYCLES_ADV_PASS_Y
    int PYCLES_ADV_PASS_Z
    int PYCLES_ADV_FUSE_YZ
    int PYCLES_DIFF_SCALARS
    int PYCLES_DIFF_MOMENTUM
    int PYCLES_DIFF_ENTROPY_SOURCE

    const char *pycles_last_error()
    const char *pycles_version()
    int pycles_device_count()
    int pycles_advection(const pycles_dims *dims, const pycles_ref *ref, const void *const *vel,
                         void *const *tend_vel, int nscal, const void *const *scal,
                         void *const *tend_scal, const pycles_scalar_base *base, int scalar_order,
                         int momentum_order, int precision, int flags, void *stream)
    int pycles_advection_rk3(const pycles_dims *dims, const pycles_ref *ref, const void *const *vel,
                             void *const *tend_vel, int nscal, const void *const *scal,
                             void *const *tend_scal, const pycles_scalar_base *base, int scalar_order,
                             int momentum_order, int precision, int flags, void *const *out_vel,
                             void *const *out_scal, void *const *copy_vel, void *const *copy_scal,
                             int rk_step, double dt, void *stream)
    int pycles_thermo_sa(const pycles_dims *dims, const pycles_ref *ref, const pycles_lookup *lt,
                         void *s, void *qt, void *wt, void *T, void *qv, void *ql, void *alpha,
                         void *buoyancy, void *theta_rho, void *bvf, void *thetali,
                         const pycles_scalar_base *base, const pycles_sa_options *opt, int precision,
                         void *stream)
    int pycles_thermo_dry(const pycles_dims *dims, const pycles_ref *ref, void *s, void *wt, void *T,
                          void *alpha, void *buoyancy, void *theta, void *bvf,
                          const pycles_scalar_base *base, int precision, int wt_overwrite, void *stream)
    int pycles_rk3_stage(const pycles_dims *dims, int nv, void *const *values, void *const *tend,
                         void *const *copies, const double *bc_type, int rk_step, double dt,
                         int precision, int zero_tendencies, void *stream)
    int pycles_apply_z_bcs(const pycles_dims *dims, int nv, void *const *values,
                           const double *bc_type, int precision, void *stream)
    int pycles_level_mean(const pycles_dims *dims, const void *field, double *mean, int precision,
                          void *stream)
    int pycles_cfl_max(const pycles_dims *dims, const void *u, const void *v, const void *w,
                       double dt, double *out, int precision, void *stream)
    int pycles_cfl_read(const double *dev_cfl, double *host_cfl, void *stream)
    int pycles_surface_bomex(const pycles_dims *dims, const pycles_ref *ref, const void *u, const void *v,
                             const void *qt, const void *T, const void *qv, const void *ql,
                             const pycles_scalar_base *base, void *const *tend, double u0, double v0,
                             const double *t_mean, double *flux2d, int precision, void *stream)
    int pycles_forcing_bomex(const pycles_dims *dims, const pycles_ref *ref, const double *profiles,
                             const double *uv_mean, const void *u, const void *v, const void *s, const void *qt,
                             const void *T, const void *ql, const pycles_scalar_base *base, void *const *tend,
                             double coriolis, double u0, double v0, int precision, void *stream)
    int pycles_rayleigh_damping(const pycles_dims *dims, int nv, const void *const *fields, void *const *tend,
                                const double *const *gamma, const double *means, int precision, void *stream)
    int pycles_sgs_smagorinsky(const pycles_dims *dims, const void *const *vel, const void *bvf, double cs,
                               double prt, void *strain_mag, void *visc, void *diff, int precision, void *stream)
    int pycles_diffusion(const pycles_dims *dims, const pycles_ref *ref, const void *const *vel,
                         void *const *tend_vel, int nscal, const void *const *scal, void *const *tend_scal,
                         const pycles_scalar_base *base, const void *visc, const void *diff,
                         const void *strain_mag, const void *temperature, int flags, int precision, void *stream)
    size_t pycles_horizontal_stats_workspace_bytes(const pycles_dims *dims)
    int pycles_horizontal_stats(const pycles_dims *dims, const void *a, const double *base_a, const void *b,
                                const double *base_b, const void *mask, int precision, double *out,
                                void *workspace, void *stream)
    int pycles_sa_stat_fields(const pycles_dims *dims, const pycles_ref *ref, const pycles_lookup *lt,
                              const void *s, const void *qt, const void *T, const void *qv,
                              const pycles_scalar_base *base, int precision, void *thetas, void *theta,
                              void *rh, void *stream)
    size_t pycles_cloud_stats_workspace_bytes(const pycles_dims *dims)
    int pycles_cloud_stats(const pycles_dims *dims, const double *rho0_half, const void *ql, int precision,
                           double *out, void *workspace, void *stream)
    int pycles_launch_chunks(int kernel, const pycles_dims *dims, int *nchunks, int *chunk_len)
    int pycles_stream(int mode, const void *src, void *dst, long long nbytes, int blocks, void *stream)
    int pycles_pressure_plan_create(const pycles_dims *dims, const double *rho0_host,
                                    const double *rho0_half_host, int precision,
                                    pycles_pressure_plan **plan)
    void pycles_pressure_plan_destroy(pycles_pressure_plan *plan)
    int pycles_pressure_update(pycles_pressure_plan *plan, const pycles_ref *ref, void *u, void *v,
                               void *w, void *p, void *div, double dt, double *cfl_out, void *stream)
    int pycles_divergence(const pycles_dims *dims, const pycles_ref *ref, const void *u,
                          const void *v, const void *w, void *div, int precision, void *stream)
    int pycles_layout_ingress(const pycles_dims *dims, const double *ref_layout, void *dev,
                              const double *base, int precision, void *stream)
    int pycles_layout_egress(const pycles_dims *dims, const void *dev, double *ref_layout,
                             const double *base, int precision, void *stream)
    int pycles_sa_training_pairs(const pycles_lookup *lt, long long n, unsigned long long seed,
                                 int max_iter, int saturated_only, double *out, void *stream)
    size_t pycles_mlp_param_count()
    size_t pycles_mlp_norm_bytes()
    size_t pycles_mlp_weights_bytes()
    size_t pycles_mlp_workspace_bytes(int batch)
    int pycles_mlp_init(float *params, float *adam_m, float *adam_v, unsigned long long seed, void *stream)
    int pycles_mlp_normalisation(const double *data, long long n, void *norm, void *workspace, void *stream)
    int pycles_mlp_grad(const float *params, const void *norm, const double *data, long long n, float *grad,
                        double *loss, void *workspace, void *stream)
    int pycles_mlp_train(float *params, float *adam_m, float *adam_v, const void *norm, const double *data,
                         long long n, int batch, int epochs, float lr_max, float lr_min,
                         unsigned long long seed, long long step0, long long total_steps,
                         double *epoch_loss, void *workspace, void *stream)
    int pycles_mlp_forward(const float *params, const void *norm, const double *data, long long n,
                           float *out, void *stream)
    int pycles_mlp_pack(const float *params, const void *norm, void *weights, void *stream)
    int pycles_mlp_infer(const void *weights, const double *data, long long n, float *out, void *stream)

F64 = PYCLES_F64
F32 = PYCLES_F32
LH_VARIABLE = PYCLES_LH_VARIABLE
LH_CONSTANT = PYCLES_LH_CONSTANT
ADV_TEND_ZERO = PYCLES_ADV_TEND_ZERO
ADV_PASS_X = PYCLES_ADV_PASS_X
ADV_PASS_Y = PYCLES_ADV_PASS_Y
ADV_PASS_Z = PYCLES_ADV_PASS_Z
ADV_FUSE_YZ = PYCLES_ADV_FUSE_YZ
DIFF_SCALARS = PYCLES_DIFF_SCALARS
DIFF_MOMENTUM = PYCLES_DIFF_MOMENTUM
DIFF_ENTROPY_SOURCE = PYCLES_DIFF_ENTROPY_SOURCE


class PyclesError(RuntimeError):
    pass


cdef inline void _check(int rc) except *:
    if rc != 0:
        raise PyclesError("pycles_amd error %d: %s" % (rc, pycles_last_error().decode()))


cdef inline void *_p(object x):
    if x is None:
        return NULL
    return <void *> <uintptr_t> x


def version():
    return pycles_version().decode()


def device_count():
    return pycles_device_count()


cdef class Dims:
    cdef pycles_dims d

    def __init__(self, n, gw, dx):
        for i in range(3):
            self.d.n[i] = n[i]
            self.d.dx[i] = dx[i]
            self.d.dxi[i] = 1.0 / dx[i]
        self.d.gw = gw


cdef class Ref:
    cdef pycles_ref r

    def __init__(self, rho0, rho0_half, alpha0, alpha0_half, p0, p0_half):
        self.r.rho0 = <const double *> _p(rho0)
        self.r.rho0_half = <const double *> _p(rho0_half)
        self.r.alpha0 = <const double *> _p(alpha0)
        self.r.alpha0_half = <const double *> _p(alpha0_half)
        self.r.p0 = <const double *> _p(p0)
        self.r.p0_half = <const double *> _p(p0_half)


cdef class ScalarBase:
    cdef pycles_scalar_base b

    def __init__(self, b0=None, b1=None):
        self.b.base[0] = <const double *> _p(b0)
        self.b.base[1] = <const double *> _p(b1)


cdef class Lookup:
    cdef pycles_lookup t

    def __init__(self, y_ptr, n, x_min, dx, dxi):
        self.t.y = <const double *> _p(y_ptr)
        self.t.n = n
        self.t.x_min = x_min
        self.t.dx = dx
        self.t.dxi = dxi


cdef class SAOptions:
    cdef pycles_sa_options o

    def __init__(self, do_qt_clipping=True, max_iter=100, latent_heat=0, mlp=None,
                 newton_steps=0, counters=None, wt_overwrite=False, secant_tol=0.0):
        self.o.do_qt_clipping = 1 if do_qt_clipping else 0
        self.o.max_iter = max_iter
        self.o.latent_heat = latent_heat
        self.o.mlp = _p(mlp)
        self.o.newton_steps = newton_steps
        self.o.counters = <long long *> _p(counters)
        self.o.wt_overwrite = 1 if wt_overwrite else 0
        self.o.secant_tol = secant_tol


cdef class PressurePlan:
    cdef pycles_pressure_plan *plan

    def __cinit__(self):
        self.plan = NULL

    def __init__(self, Dims dims, rho0_host_ptr, rho0_half_host_ptr, int precision):
        _check(pycles_pressure_plan_create(&dims.d, <const double *> _p(rho0_host_ptr),
                                           <const double *> _p(rho0_half_host_ptr), precision,
                                           &self.plan))

    def __dealloc__(self):
        if self.plan != NULL:
            pycles_pressure_plan_destroy(self.plan)
            self.plan = NULL

    def update(self, Ref ref, u, v, w, p, div, double dt, cfl_out, stream):
        _check(pycles_pressure_update(self.plan, &ref.r, _p(u), _p(v), _p(w), _p(p), _p(div), dt,
                                      <double *> _p(cfl_out), _p(stream)))


def advection(Dims dims, Ref ref, vel, tend_vel, scal, tend_scal, ScalarBase base,
              int scalar_order, int momentum_order, int precision, stream, int flags=0):
    cdef const void *v3[3]
    cdef void *t3[3]
    cdef const void *sc[2]
    cdef void *ts[2]
    cdef int ns = len(scal)
    for i in range(3):
        v3[i] = _p(vel[i])
        t3[i] = _p(tend_vel[i]) if tend_vel is not None else NULL
    for i in range(ns):
        sc[i] = _p(scal[i])
        ts[i] = _p(tend_scal[i])
    _check(pycles_advection(&dims.d, &ref.r, v3, t3 if tend_vel is not None else NULL, ns, sc, ts,
                            &base.b if base is not None else NULL, scalar_order, momentum_order,
                            precision, flags, _p(stream)))


def advection_rk3(Dims dims, Ref ref, vel, tend_vel, scal, tend_scal, ScalarBase base, int scalar_order,
                  int momentum_order, int precision, stream, int flags, out_vel, out_scal, copy_vel, copy_scal,
                  int rk_step, double dt):
    cdef const void *v3[3]
    cdef void *t3[3]
    cdef void *o3[3]
    cdef void *c3[3]
    cdef const void *sc[2]
    cdef void *ts[2]
    cdef void *osc[2]
    cdef void *csc[2]
    cdef int ns = len(scal)
    # copy_vel / copy_scal None: no stage copies (allowed at rk_step 0)
    cdef bint has_copy = copy_vel is not None
    for i in range(3):
        v3[i] = _p(vel[i])
        t3[i] = _p(tend_vel[i])
        o3[i] = _p(out_vel[i])
        c3[i] = _p(copy_vel[i]) if has_copy else NULL
    for i in range(ns):
        sc[i] = _p(scal[i])
        ts[i] = _p(tend_scal[i])
        osc[i] = _p(out_scal[i])
        csc[i] = _p(copy_scal[i]) if has_copy else NULL
    _check(pycles_advection_rk3(&dims.d, &ref.r, v3, t3, ns, sc, ts, &base.b if base is not None else NULL,
                                scalar_order, momentum_order, precision, flags, o3, osc,
                                c3 if has_copy else NULL, csc if has_copy else NULL, rk_step, dt, _p(stream)))


def thermo_sa(Dims dims, Ref ref, Lookup lt, s, qt, wt, T, qv, ql, alpha, buoyancy, theta_rho,
              bvf, thetali, ScalarBase base, SAOptions opt, int precision, stream):
    _check(pycles_thermo_sa(&dims.d, &ref.r, &lt.t, _p(s), _p(qt), _p(wt), _p(T), _p(qv), _p(ql),
                            _p(alpha), _p(buoyancy), _p(theta_rho), _p(bvf), _p(thetali),
                            &base.b if base is not None else NULL, &opt.o, precision, _p(stream)))


def thermo_dry(Dims dims, Ref ref, s, wt, T, alpha, buoyancy, theta, bvf, ScalarBase base,
               int precision, stream, int wt_overwrite=0):
    _check(pycles_thermo_dry(&dims.d, &ref.r, _p(s), _p(wt), _p(T), _p(alpha), _p(buoyancy),
                             _p(theta), _p(bvf), &base.b if base is not None else NULL, precision,
                             wt_overwrite, _p(stream)))


def rk3_stage(Dims dims, values, tend, copies, bc_type, int rk_step, double dt, int precision, stream,
              int zero_tendencies=1):
    cdef int nv = len(values)
    cdef void *v[8]
    cdef void *t[8]
    cdef void *c[8]
    cdef double bc[8]
    if nv > 8:
        raise ValueError("at most 8 fields")
    for i in range(nv):
        v[i] = _p(values[i])
        t[i] = _p(tend[i])
        c[i] = _p(copies[i])
        bc[i] = bc_type[i]
    _check(pycles_rk3_stage(&dims.d, nv, v, t, c, bc, rk_step, dt, precision, zero_tendencies, _p(stream)))


def apply_z_bcs(Dims dims, values, bc_type, int precision, stream):
    cdef int nv = len(values)
    cdef void *v[8]
    cdef double bc[8]
    if nv > 8:
        raise ValueError("at most 8 fields")
    for i in range(nv):
        v[i] = _p(values[i])
        bc[i] = bc_type[i]
    _check(pycles_apply_z_bcs(&dims.d, nv, v, bc, precision, _p(stream)))


def level_mean(Dims dims, field, mean, int precision, stream):
    _check(pycles_level_mean(&dims.d, _p(field), <double *> _p(mean), precision, _p(stream)))


def cfl_max(Dims dims, u, v, w, double dt, out, int precision, stream):
    _check(pycles_cfl_max(&dims.d, _p(u), _p(v), _p(w), dt, <double *> _p(out), precision, _p(stream)))


def cfl_read(dev_cfl, stream):
    """Synchronise the stream and return the device CFL value; raises
    PyclesError (status PYCLES_ERR_NONFINITE) if it is NaN or infinite."""
    cdef double h = 0.0
    _check(pycles_cfl_read(<const double *> _p(dev_cfl), &h, _p(stream)))
    return h


def horizontal_stats_workspace_bytes(Dims dims):
    return pycles_horizontal_stats_workspace_bytes(&dims.d)


def horizontal_stats(Dims dims, a, base_a, b, base_b, mask, int precision, out, workspace, stream):
    _check(pycles_horizontal_stats(&dims.d, _p(a), <const double *> _p(base_a), _p(b),
                                   <const double *> _p(base_b), _p(mask), precision, <double *> _p(out),
                                   _p(workspace), _p(stream)))


def sa_stat_fields(Dims dims, Ref ref, Lookup lt, s, qt, T, qv, ScalarBase base, int precision, thetas, theta,
                   rh, stream):
    _check(pycles_sa_stat_fields(&dims.d, &ref.r, &lt.t, _p(s), _p(qt), _p(T), _p(qv),
                                 &base.b if base is not None else NULL, precision, _p(thetas), _p(theta),
                                 _p(rh), _p(stream)))


def cloud_stats_workspace_bytes(Dims dims):
    return pycles_cloud_stats_workspace_bytes(&dims.d)


def cloud_stats(Dims dims, rho0_half, ql, int precision, out, workspace, stream):
    _check(pycles_cloud_stats(&dims.d, <const double *> _p(rho0_half), _p(ql), precision, <double *> _p(out),
                              _p(workspace), _p(stream)))


def sgs_smagorinsky(Dims dims, vel, bvf, double cs, double prt, strain_mag, visc, diff, int precision, stream):
    cdef const void *v3[3]
    for i in range(3):
        v3[i] = _p(vel[i])
    _check(pycles_sgs_smagorinsky(&dims.d, v3, _p(bvf), cs, prt, _p(strain_mag), _p(visc), _p(diff), precision,
                                  _p(stream)))


def diffusion(Dims dims, Ref ref, vel, tend_vel, scal, tend_scal, ScalarBase base, visc, diff, strain_mag,
              temperature, int flags, int precision, stream):
    cdef const void *v3[3]
    cdef void *t3[3]
    cdef const void *sc[2]
    cdef void *ts[2]
    cdef int ns = len(scal)
    for i in range(3):
        v3[i] = _p(vel[i])
        t3[i] = _p(tend_vel[i]) if tend_vel is not None else NULL
    for i in range(ns):
        sc[i] = _p(scal[i])
        ts[i] = _p(tend_scal[i])
    _check(pycles_diffusion(&dims.d, &ref.r, v3, t3 if tend_vel is not None else NULL, ns, sc, ts,
                            &base.b if base is not None else NULL, _p(visc), _p(diff), _p(strain_mag),
                            _p(temperature), flags, precision, _p(stream)))


def surface_bomex(Dims dims, Ref ref, u, v, qt, T, qv, ql, ScalarBase base, tend, double u0, double v0, t_mean,
                  flux2d, int precision, stream):
    cdef void *t4[4]
    for i in range(4):
        t4[i] = _p(tend[i])
    _check(pycles_surface_bomex(&dims.d, &ref.r, _p(u), _p(v), _p(qt), _p(T), _p(qv), _p(ql),
                                &base.b if base is not None else NULL, t4, u0, v0, <const double *> _p(t_mean),
                                <double *> _p(flux2d), precision, _p(stream)))


def forcing_bomex(Dims dims, Ref ref, profiles, uv_mean, u, v, s, qt, T, ql, ScalarBase base, tend,
                  double coriolis, double u0, double v0, int precision, stream):
    cdef void *t4[4]
    for i in range(4):
        t4[i] = _p(tend[i])
    _check(pycles_forcing_bomex(&dims.d, &ref.r, <const double *> _p(profiles), <const double *> _p(uv_mean),
                                _p(u), _p(v), _p(s), _p(qt), _p(T), _p(ql), &base.b if base is not None else NULL,
                                t4, coriolis, u0, v0, precision, _p(stream)))


def rayleigh_damping(Dims dims, fields, tend, gamma, means, int precision, stream):
    cdef int nv = len(fields)
    cdef const void *f[8]
    cdef void *t[8]
    cdef const double *gm[8]
    if nv > 8:
        raise ValueError("at most 8 fields")
    for i in range(nv):
        f[i] = _p(fields[i])
        t[i] = _p(tend[i])
        gm[i] = <const double *> _p(gamma[i])
    _check(pycles_rayleigh_damping(&dims.d, nv, f, t, gm, <const double *> _p(means), precision, _p(stream)))


def launch_chunks(int kernel, Dims dims):
    cdef int nch = 0, ln = 0
    _check(pycles_launch_chunks(kernel, &dims.d, &nch, &ln))
    return nch, ln


def stream_kernel(int mode, src, dst, long long nbytes, int blocks, stream):
    """HBM calibration kernel: 0 copy, 1 read, 2 fill (+4: non-temporal)."""
    _check(pycles_stream(mode, _p(src), _p(dst), nbytes, blocks, _p(stream)))


def divergence(Dims dims, Ref ref, u, v, w, div, int precision, stream):
    _check(pycles_divergence(&dims.d, &ref.r, _p(u), _p(v), _p(w), _p(div), precision, _p(stream)))


def layout_ingress(Dims dims, ref_layout, dev, base, int precision, stream):
    _check(pycles_layout_ingress(&dims.d, <const double *> _p(ref_layout), _p(dev),
                                 <const double *> _p(base), precision, _p(stream)))


def layout_egress(Dims dims, dev, ref_layout, base, int precision, stream):
    _check(pycles_layout_egress(&dims.d, _p(dev), <double *> _p(ref_layout),
                                <const double *> _p(base), precision, _p(stream)))


# ---- learned first guess of the saturation adjustment --------------------

def sa_training_pairs(Lookup lt, long long n, unsigned long long seed, int max_iter, out, stream,
                      int saturated_only=1):
    _check(pycles_sa_training_pairs(&lt.t, n, seed, max_iter, saturated_only, <double *> _p(out), _p(stream)))


def mlp_param_count():
    return pycles_mlp_param_count()


def mlp_norm_bytes():
    return pycles_mlp_norm_bytes()


def mlp_weights_bytes():
    return pycles_mlp_weights_bytes()


def mlp_workspace_bytes(int batch):
    return pycles_mlp_workspace_bytes(batch)


def mlp_init(params, adam_m, adam_v, unsigned long long seed, stream):
    _check(pycles_mlp_init(<float *> _p(params), <float *> _p(adam_m), <float *> _p(adam_v), seed, _p(stream)))


def mlp_normalisation(data, long long n, norm, workspace, stream):
    _check(pycles_mlp_normalisation(<const double *> _p(data), n, _p(norm), _p(workspace), _p(stream)))


def mlp_grad(params, norm, data, long long n, grad, loss, workspace, stream):
    _check(pycles_mlp_grad(<const float *> _p(params), _p(norm), <const double *> _p(data), n,
                           <float *> _p(grad), <double *> _p(loss), _p(workspace), _p(stream)))


def mlp_train(params, adam_m, adam_v, norm, data, long long n, int batch, int epochs, float lr_max,
              float lr_min, unsigned long long seed, long long step0, long long total_steps, epoch_loss,
              workspace, stream):
    _check(pycles_mlp_train(<float *> _p(params), <float *> _p(adam_m), <float *> _p(adam_v), _p(norm),
                            <const double *> _p(data), n, batch, epochs, lr_max, lr_min, seed, step0,
                            total_steps, <double *> _p(epoch_loss), _p(workspace), _p(stream)))


def mlp_forward(params, norm, data, long long n, out, stream):
    _check(pycles_mlp_forward(<const float *> _p(params), _p(norm), <const double *> _p(data), n,
                              <float *> _p(out), _p(stream)))


def mlp_pack(params, norm, weights, stream):
    _check(pycles_mlp_pack(<const float *> _p(params), _p(norm), _p(weights), _p(stream)))


def mlp_infer(weights, data, long long n, out, stream):
    _check(pycles_mlp_infer(_p(weights), <const double *> _p(data), n, <float *> _p(out), _p(stream)))
