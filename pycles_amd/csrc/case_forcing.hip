// Case physics of the BOMEX configuration: surface fluxes, large-scale
// forcing (geostrophic pressure gradient with Coriolis, radiative cooling and
// drying, subsidence) and Rayleigh damping of the top layer.
//
//   SurfaceBomex.update + SurfaceBase.update   Surface.pyx:297-373, 100-183;
//                                              surface.h:95-106, 157-185
//   ForcingBomex.update                        Forcing.pyx:147-203, 1982-2040;
//                                              entropies.h:17-21
//   Rayleigh.update                            Damping.pyx:95-133
//
// All three add to the tendencies after the advection, in the reference's
// order within each variable.  The horizontal means they need (u and v for
// the geostrophic term, every variable for the damping, T at the first level
// for the surface diagnostics) come from pycles_level_mean beforehand.
#include "common.h"
#include "thermo_funcs.h"

extern int pycles_check_dims(const pycles_dims *d);

namespace {

template <typename T>
__device__ __forceinline__ double val(const T *f, const double *base, long long o, int k) {
    return (double)f[o] + (base ? base[k] : 0.0);
}

// One thread per column: the surface fluxes at the first interior level.
template <typename TS>
__global__ __launch_bounds__(256) void surface_bomex_kernel(Geo g, const double *p0_half, const double *rho0_half,
                                                            const double *alpha0, const double *alpha0_half,
                                                            const TS *u, const TS *v, const TS *qt, const TS *T,
                                                            const TS *qv, const TS *ql, const double *qbase, TS *ut,
                                                            TS *vt, TS *st, TS *qtt, double u0, double v0,
                                                            const double *t_mean, double *flux2d) {
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.plane) return;
    const int i = (int)(c % g.nx), j = (int)(c / g.nx);
    const int gw = g.gw;
    const long long lev = (long long)gw * g.plane;
    const int im = i == 0 ? g.nx - 1 : i - 1, ip = i + 1 == g.nx ? 0 : i + 1;
    const int jm = j == 0 ? g.ny - 1 : j - 1, jp = j + 1 == g.ny ? 0 : j + 1;
    const double theta_flux = 8.0e-3, ustar = 0.28, qt_flux = 5.2e-5, gustiness = 0.0;
    const long long o = lev + c;
    const double Tb = (double)T[o], qb = val(qt, qbase, o, gw), qvb = (double)qv[o];
    // entropyflux_from_thetaflux_qtflux (surface.h:95-106)
    const double p0b = p0_half[gw];
    const double exner_b = thermo::exner_c(p0b);
    const double pd_b = thermo::pd_c(p0b, qb, qvb), pv_b = thermo::pv_c(p0b, qb, qvb);
    const double sd_b = thermo::sd_c(pd_b, Tb), sv_b = thermo::sv_c(pv_b, Tb);
    const double sflux = thermo::cpm_c(qb) * theta_flux * exner_b / Tb + qt_flux * (sv_b - sd_b);
    // compute_windspeed (surface.h:157-185) at this column and its +x, +y neighbours
    auto speed = [&](int ii, int jj) {
        const int iim = ii == 0 ? g.nx - 1 : ii - 1, jjm = jj == 0 ? g.ny - 1 : jj - 1;
        const long long cc = lev + (long long)jj * g.nx + ii;
        const double ui = 0.5 * ((double)u[lev + (long long)jj * g.nx + iim] + (double)u[cc]) + u0;
        const double vi = 0.5 * ((double)v[lev + (long long)jjm * g.nx + ii] + (double)v[cc]) + v0;
        return fmax(sqrt(ui * ui + vi * vi), gustiness);
    };
    (void)im;
    (void)jm;
    const double ws = speed(i, j);
    const double uflux = -(ustar * ustar) / (0.5 * (ws + speed(ip, j))) * ((double)u[o] + u0);
    const double vflux = -(ustar * ustar) / (0.5 * (ws + speed(i, jp))) * ((double)v[o] + v0);
    // SurfaceBase.update, moist branch (Surface.pyx:150-183)
    const double tf = alpha0_half[gw] / alpha0[gw - 1] * (1.0 / g.dx[2]);
    ut[o] = (TS)((double)ut[o] + uflux * tf);
    vt[o] = (TS)((double)vt[o] + vflux * tf);
    st[o] = (TS)((double)st[o] + sflux * tf);
    qtt[o] = (TS)((double)qtt[o] + qt_flux * tf);
    if (flux2d) {
        const double lv = thermo::latent_heat_variable(Tb);
        const double lhf = qt_flux * rho0_half[gw] * lv;
        const double qlb = (double)ql[o];
        const double pv = thermo::pv_c(p0b, qb, qb - qlb), pd = thermo::pd_c(p0b, qb, qb - qlb);
        const double sv = thermo::sv_c(pv, Tb), sd = thermo::sd_c(pd, Tb);
        const double shf = (sflux * rho0_half[gw] - lhf / lv * (sv - sd)) * Tb;
        const double cp_ = thermo::cpm_c(qb);
        const double tm = t_mean[gw];
        const double bf = PC_G * alpha0_half[gw] / cp_ / tm * (shf + (PC_EPS_VI - 1.0) * cp_ * tm * lhf / lv);
        const double q[8] = {uflux, vflux, sflux, qt_flux, shf, lhf, bf, -ustar * ustar * ustar / bf / 0.35};
        for (int r = 0; r < 8; ++r) flux2d[r * g.plane + c] = q[r];
    }
}

// One thread per interior cell: ForcingBomex.update.  prof = [5][nzg] ug, vg,
// dtdt, dqtdt, subsidence; mean = [2][nzg] horizontal means of u, v.
template <typename TS>
__global__ __launch_bounds__(256) void forcing_bomex_kernel(Geo g, const double *p0_half, const double *prof,
                                                            const double *mean, const TS *u, const TS *v,
                                                            const TS *s, const TS *qt, const TS *T, const TS *ql,
                                                            const double *sbase, const double *qbase, TS *ut, TS *vt,
                                                            TS *st, TS *qtt, double coriolis, double u0, double v0) {
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y + g.gw;
    if (c >= g.plane) return;
    const long long o = (long long)k * g.plane + c;
    const int nzg = g.nzg;
    const double *ug = prof, *vg = prof + nzg, *dtdt = prof + 2 * nzg, *dqtdt = prof + 3 * nzg,
                 *sub = prof + 4 * nzg;
    const double dxi = g.dxi[2];
    // the last interior level repeats the tendency of the level below
    // (apply_subsidence, Forcing.pyx:2035-2038)
    const int ks = k < g.nz + g.gw - 1 ? k : k - 1;
    const long long os = (long long)ks * g.plane + c;
    auto subs = [&](const TS *f, const double *base) {
        double d;
        if (base)
            d = (base[ks + 1] - base[ks]) + ((double)f[os + g.plane] - (double)f[os]);
        else
            d = (double)f[os + g.plane] - (double)f[os];
        return d * dxi * sub[ks];
    };
    // large_scale_p_gradient (Forcing.pyx:1982-2011), then subsidence
    double tu = (double)ut[o], tv = (double)vt[o];
    tu = tu - coriolis * (vg[k] - (mean[nzg + k] + v0));
    tv = tv + coriolis * (ug[k] - (mean[k] + u0));
    tu -= subs(u, nullptr);
    tv -= subs(v, nullptr);
    ut[o] = (TS)tu;
    vt[o] = (TS)tv;
    // s_tendency_c (entropies.h:17-21) and the drying, then subsidence
    const double p0 = p0_half[k], q = val(qt, qbase, o, k), qv = q - (double)ql[o], t = (double)T[o];
    const double pv = thermo::pv_c(p0, q, qv), pd = thermo::pd_c(p0, q, qv);
    double ts = (double)st[o], tq = (double)qtt[o];
    ts += thermo::cpm_c(q) * dtdt[k] / t + (thermo::sv_c(pv, t) - thermo::sd_c(pd, t)) * dqtdt[k];
    tq += dqtdt[k];
    ts -= subs(s, sbase);
    tq -= subs(qt, qbase);
    st[o] = (TS)ts;
    qtt[o] = (TS)tq;
}

// Rayleigh damping of nv fields: tend -= (phi - mean(phi)) gamma.
struct DampArgs {
    const void *f[8];
    void *t[8];
    const double *base[8];
    const double *gamma[8];
    const double *mean;  // [nv][nzg]
    int nv;
};

template <typename TS>
__global__ __launch_bounds__(256) void rayleigh_kernel(Geo g, DampArgs a) {
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y + g.gw;
    if (c >= g.plane) return;
    const long long o = (long long)k * g.plane + c;
    for (int q = 0; q < a.nv; ++q) {
        const double gm = a.gamma[q][k];
        if (gm == 0.0) continue;  // subtracting 0 * x leaves every finite tendency unchanged
        const TS *f = (const TS *)a.f[q];
        TS *t = (TS *)a.t[q];
        // in FP32 the level profile cancels: phi - mean = dev - mean(dev)
        const double d = (double)f[o] - a.mean[q * g.nzg + k];
        t[o] = (TS)((double)t[o] - d * gm);
    }
}

template <typename TS>
int launch_surface(const Geo &g, const pycles_ref *ref, const void *u, const void *v, const void *qt, const void *T,
                   const void *qv, const void *ql, const double *qbase, void *const tend[4], double u0, double v0,
                   const double *t_mean, double *flux2d, hipStream_t st) {
    hipLaunchKernelGGL(surface_bomex_kernel<TS>, dim3((unsigned)((g.plane + 255) / 256)), dim3(256), 0, st, g,
                       ref->p0_half, ref->rho0_half, ref->alpha0, ref->alpha0_half, (const TS *)u, (const TS *)v,
                       (const TS *)qt, (const TS *)T, (const TS *)qv, (const TS *)ql, qbase, (TS *)tend[0],
                       (TS *)tend[1], (TS *)tend[2], (TS *)tend[3], u0, v0, t_mean, flux2d);
    PYCLES_CHECK_LAUNCH();
    return PYCLES_OK;
}

}  // namespace

static int prec_ok(int precision, const char *who) {
    if (precision != PYCLES_F64 && precision != PYCLES_F32) {
        pycles_set_error("%s: bad precision %d", who, precision);
        return PYCLES_ERR_ARG;
    }
    return PYCLES_OK;
}

extern "C" int pycles_surface_bomex(const pycles_dims *dims, const pycles_ref *ref, const void *u, const void *v,
                                    const void *qt, const void *T, const void *qv, const void *ql,
                                    const pycles_scalar_base *base, void *const tend[4], double u0, double v0,
                                    const double *t_mean, double *flux2d, int precision, void *stream) {
    int rc = pycles_check_dims(dims);
    if (rc) return rc;
    if ((rc = prec_ok(precision, "surface_bomex"))) return rc;
    if (!ref || !ref->p0_half || !ref->rho0_half || !ref->alpha0 || !ref->alpha0_half || !u || !v || !qt || !T ||
        !qv || !tend || !tend[0] || !tend[1] || !tend[2] || !tend[3] || (flux2d && (!t_mean || !ql))) {
        pycles_set_error("surface_bomex: NULL argument");
        return PYCLES_ERR_ARG;
    }
    const Geo g = make_geo(dims);
    const double *qb = (precision == PYCLES_F32 && base) ? base->base[1] : nullptr;
    hipStream_t st = (hipStream_t)stream;
    if (precision == PYCLES_F64)
        return launch_surface<double>(g, ref, u, v, qt, T, qv, ql, qb, tend, u0, v0, t_mean, flux2d, st);
    return launch_surface<float>(g, ref, u, v, qt, T, qv, ql, qb, tend, u0, v0, t_mean, flux2d, st);
}

extern "C" int pycles_forcing_bomex(const pycles_dims *dims, const pycles_ref *ref, const double *profiles,
                                    const double *uv_mean, const void *u, const void *v, const void *s,
                                    const void *qt, const void *T, const void *ql, const pycles_scalar_base *base,
                                    void *const tend[4], double coriolis, double u0, double v0, int precision,
                                    void *stream) {
    int rc = pycles_check_dims(dims);
    if (rc) return rc;
    if ((rc = prec_ok(precision, "forcing_bomex"))) return rc;
    if (!ref || !ref->p0_half || !profiles || !uv_mean || !u || !v || !s || !qt || !T || !ql || !tend || !tend[0] ||
        !tend[1] || !tend[2] || !tend[3]) {
        pycles_set_error("forcing_bomex: NULL argument");
        return PYCLES_ERR_ARG;
    }
    const Geo g = make_geo(dims);
    const double *sb = (precision == PYCLES_F32 && base) ? base->base[0] : nullptr;
    const double *qb = (precision == PYCLES_F32 && base) ? base->base[1] : nullptr;
    hipStream_t st = (hipStream_t)stream;
    const dim3 grid((unsigned)((g.plane + 255) / 256), g.nz);
    if (precision == PYCLES_F64)
        hipLaunchKernelGGL(forcing_bomex_kernel<double>, grid, dim3(256), 0, st, g, ref->p0_half, profiles, uv_mean,
                           (const double *)u, (const double *)v, (const double *)s, (const double *)qt,
                           (const double *)T, (const double *)ql, sb, qb, (double *)tend[0], (double *)tend[1],
                           (double *)tend[2], (double *)tend[3], coriolis, u0, v0);
    else
        hipLaunchKernelGGL(forcing_bomex_kernel<float>, grid, dim3(256), 0, st, g, ref->p0_half, profiles, uv_mean,
                           (const float *)u, (const float *)v, (const float *)s, (const float *)qt, (const float *)T,
                           (const float *)ql, sb, qb, (float *)tend[0], (float *)tend[1], (float *)tend[2],
                           (float *)tend[3], coriolis, u0, v0);
    PYCLES_CHECK_LAUNCH();
    return PYCLES_OK;
}

extern "C" int pycles_rayleigh_damping(const pycles_dims *dims, int nv, const void *const *fields,
                                       void *const *tend, const double *const *gamma, const double *means,
                                       int precision, void *stream) {
    int rc = pycles_check_dims(dims);
    if (rc) return rc;
    if ((rc = prec_ok(precision, "rayleigh_damping"))) return rc;
    if (nv < 1 || nv > 8 || !fields || !tend || !gamma || !means) {
        pycles_set_error("rayleigh_damping: bad arguments (nv = %d)", nv);
        return PYCLES_ERR_ARG;
    }
    DampArgs a;
    a.nv = nv;
    a.mean = means;
    for (int q = 0; q < nv; ++q) {
        if (!fields[q] || !tend[q] || !gamma[q]) {
            pycles_set_error("rayleigh_damping: NULL field %d", q);
            return PYCLES_ERR_ARG;
        }
        a.f[q] = fields[q];
        a.t[q] = tend[q];
        a.gamma[q] = gamma[q];
    }
    const Geo g = make_geo(dims);
    hipStream_t st = (hipStream_t)stream;
    const dim3 grid((unsigned)((g.plane + 255) / 256), g.nz);
    if (precision == PYCLES_F64)
        hipLaunchKernelGGL(rayleigh_kernel<double>, grid, dim3(256), 0, st, g, a);
    else
        hipLaunchKernelGGL(rayleigh_kernel<float>, grid, dim3(256), 0, st, g, a);
    PYCLES_CHECK_LAUNCH();
    return PYCLES_OK;
}
