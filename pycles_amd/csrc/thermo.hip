// Thermodynamics kernels: saturation adjustment (SA) and dry.
//
// One thread per (i, j) column marching k over all nzg levels, so that the
// vertical neighbours needed by the buoyancy average into the w tendency and
// by the buoyancy frequency are this thread's own previous levels.
//
//   ThermodynamicsSA.update       ThermodynamicsSA.pyx:176-213
//   clip_qt                       thermodynamics_sa.h:213-220
//   eos_c / eos_update            thermodynamics_sa.h:17-96
//   buoyancy_update_sa            thermodynamics_sa.h:98-132
//   bvf_sa                        thermodynamics_sa.h:134-182
//   thetali_update                thermodynamics_sa.h:185-211
//   ThermodynamicsDry.update      ThermodynamicsDry.pyx:105-120, thermodynamics_dry.h
//   lookup                        lookup.h:48-54 (table staged in LDS)
//   latent heat                   Microphysics.pxd:65-75
#include "common.h"
#include "thermo_funcs.h"

#define LOOKUP_MAX 2048

struct SAArgs {
    const double *p0, *alpha0_half;
    const double *tab;
    int ntab;
    double x_min, dx, dxi;
    void *s, *qt, *wt, *T, *qv, *ql, *alpha, *buoy, *thr, *bvf, *thl;
    const double *sbase, *qbase;
    int clip, max_iter, lh;
    const void *mlp;
    int newton;
    unsigned long long *counters;
};

template <typename TS>
__global__ __launch_bounds__(256) void thermo_sa_kernel(Geo g, SAArgs a) {
    __shared__ double tab[LOOKUP_MAX];
    __shared__ unsigned long long cnt[4];
    const int tid = threadIdx.y * blockDim.x + threadIdx.x;
    for (int t = tid; t < a.ntab; t += blockDim.x * blockDim.y) tab[t] = a.tab[t];
    if (tid < 4) cnt[tid] = 0;
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    unsigned long long n_sat = 0, n_it = 0, n_cap = 0;
    if (i < g.nx && j < g.ny) {
        const long long col = (long long)j * g.nx + i;
        TS *s = (TS *)a.s;
        TS *qtp = (TS *)a.qt;
        TS *wt = (TS *)a.wt;
        TS *To = (TS *)a.T;
        TS *qvo = (TS *)a.qv;
        TS *qlo = (TS *)a.ql;
        TS *alo = (TS *)a.alpha;
        TS *bo = (TS *)a.buoy;
        TS *thro = (TS *)a.thr;
        TS *bvfo = (TS *)a.bvf;
        TS *thlo = (TS *)a.thl;
        const thermo::Table lt{tab, a.x_min, a.dx, a.dxi, a.ntab};
        const double dzi = 1.0 / g.dx[2];
        // windows of the two previous levels (index 0 = k-2, 1 = k-1)
        double Tw[3], qtw[3], qvw[3], thw[3], bw[2];
        for (int k = 0; k < g.nzg; ++k) {
            const long long o = (long long)k * g.plane + col;
            const double p0 = a.p0[k];
            double qt = (double)qtp[o];
            if (a.qbase) qt += a.qbase[k];
            if (a.clip) {
                const double c = fmax(qt, 1e-11);
                if (!(c == qt)) {
                    qt = c;
                    qtp[o] = (TS)(a.qbase ? c - a.qbase[k] : c);
                }
            }
            double sv = (double)s[o];
            if (a.sbase) sv += a.sbase[k];
            double T, qv, ql, qi;
            const int it = thermo::eos_sa(lt, p0, sv, qt, &T, &qv, &ql, &qi, a.max_iter);
            if (it > 0) {
                n_sat += 1;
                n_it += it;
                if (a.max_iter > 0 && it >= a.max_iter) n_cap += 1;
            }
            const double al = thermo::alpha_c(p0, T, qt, qv);
            To[o] = (TS)T;
            qvo[o] = (TS)qv;
            qlo[o] = (TS)ql;
            alo[o] = (TS)al;
            // buoyancy for k in [0, nzg-1); w tendency at k-1 for k-1 in [1, nzg-3)
            double b = 0.0;
            if (k < g.nzg - 1) {
                b = thermo::buoyancy_c(a.alpha0_half[k], al);
                if (bo) bo[o] = (TS)b;
            }
            if (k >= 2 && k - 1 < g.nzg - 3) {
                const long long om = o - g.plane;
                wt[om] = (TS)((double)wt[om] + 0.5 * (bw[1] + b));
            }
            // theta_rho for k in [1, nzg-1)
            double thr = 0.0;
            if (k >= 1 && k < g.nzg - 1) {
                thr = thermo::theta_rho_c(p0, T, qt, qv);
                if (thro) thro[o] = (TS)thr;
            }
            // shift windows: slot 2 = level k
            Tw[0] = Tw[1]; Tw[1] = Tw[2]; Tw[2] = T;
            qtw[0] = qtw[1]; qtw[1] = qtw[2]; qtw[2] = qt;
            qvw[0] = qvw[1]; qvw[1] = qvw[2]; qvw[2] = qv;
            thw[0] = thw[1]; thw[1] = thw[2]; thw[2] = thr;
            bw[0] = bw[1]; bw[1] = b;
            // bvf at m = k-1 for m in [2, nzg-2)
            const int m = k - 1;
            if (bvfo && m >= 2 && m < g.nzg - 2) {
                bvfo[o - g.plane] = (TS)thermo::bvf_sa(lt, a.p0[m], Tw, qtw, qvw, thw, dzi);
            }
            if (thlo) {
                const double L = thermo::latent_heat_variable(T);
                thlo[o] = (TS)thermo::thetali_c(p0, T, qt, ql, qi, L);
            }
        }
    }
    if (a.counters) {
        atomicAdd(&cnt[0], n_sat);
        atomicAdd(&cnt[1], n_it);
        atomicAdd(&cnt[2], n_cap);
        __syncthreads();
        if (tid < 3 && cnt[tid]) atomicAdd(&a.counters[tid], cnt[tid]);
    }
}

extern int pycles_check_dims(const pycles_dims *d);

extern "C" int pycles_thermo_sa(const pycles_dims *dims, const pycles_ref *ref, const pycles_lookup *lt, void *s,
                                void *qt, void *wt, void *T, void *qv, void *ql, void *alpha, void *buoyancy,
                                void *theta_rho, void *bvf, void *thetali, const pycles_scalar_base *base,
                                const pycles_sa_options *opt, int precision, void *stream) {
    int rc = pycles_check_dims(dims);
    if (rc) return rc;
    if (!ref || !lt || !s || !qt || !wt || !T || !qv || !ql || !alpha || !opt) {
        pycles_set_error("thermo_sa: NULL argument");
        return PYCLES_ERR_ARG;
    }
    if (lt->n < 2 || lt->n > LOOKUP_MAX || !lt->y) {
        pycles_set_error("thermo_sa: lookup table size %lld not in [2, %d]", (long long)lt->n, LOOKUP_MAX);
        return PYCLES_ERR_ARG;
    }
    if (opt->latent_heat != PYCLES_LH_VARIABLE) {
        pycles_set_error("thermo_sa: only the variable latent heat (None_SA) is supported");
        return PYCLES_ERR_UNSUPPORTED;
    }
    Geo g = make_geo(dims);
    SAArgs a;
    a.p0 = ref->p0_half;
    a.alpha0_half = ref->alpha0_half;
    a.tab = lt->y;
    a.ntab = (int)lt->n;
    a.x_min = lt->x_min;
    a.dx = lt->dx;
    a.dxi = lt->dxi;
    a.s = s; a.qt = qt; a.wt = wt; a.T = T; a.qv = qv; a.ql = ql; a.alpha = alpha;
    a.buoy = buoyancy; a.thr = theta_rho; a.bvf = bvf; a.thl = thetali;
    a.sbase = (precision == PYCLES_F32 && base) ? base->base[0] : nullptr;
    a.qbase = (precision == PYCLES_F32 && base) ? base->base[1] : nullptr;
    a.clip = opt->do_qt_clipping;
    a.max_iter = opt->max_iter;
    a.lh = opt->latent_heat;
    a.mlp = opt->mlp;
    a.newton = opt->newton_steps;
    a.counters = (unsigned long long *)opt->counters;
    dim3 blk(64, 4), grd((g.nx + 63) / 64, (g.ny + 3) / 4);
    hipStream_t st = (hipStream_t)stream;
    if (precision == PYCLES_F64)
        hipLaunchKernelGGL(thermo_sa_kernel<double>, grd, blk, 0, st, g, a);
    else if (precision == PYCLES_F32)
        hipLaunchKernelGGL(thermo_sa_kernel<float>, grd, blk, 0, st, g, a);
    else {
        pycles_set_error("bad precision %d", precision);
        return PYCLES_ERR_ARG;
    }
    PYCLES_CHECK_LAUNCH();
    return PYCLES_OK;
}

// ---------------------------------------------------------------------------
struct DryArgs {
    const double *p0, *alpha0_half;
    void *s, *wt, *T, *alpha, *buoy, *theta, *bvf;
    const double *sbase;
};

template <typename TS>
__global__ __launch_bounds__(256) void thermo_dry_kernel(Geo g, DryArgs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= g.nx || j >= g.ny) return;
    const long long col = (long long)j * g.nx + i;
    TS *s = (TS *)a.s;
    TS *wt = (TS *)a.wt;
    TS *To = (TS *)a.T;
    TS *alo = (TS *)a.alpha;
    TS *bo = (TS *)a.buoy;
    TS *tho = (TS *)a.theta;
    TS *bvfo = (TS *)a.bvf;
    const double dzi = 1.0 / g.dx[2];
    double thw[3] = {0.0, 0.0, 0.0}, bprev = 0.0;
    for (int k = 0; k < g.nzg; ++k) {
        const long long o = (long long)k * g.plane + col;
        const double p0 = a.p0[k];
        double sv = (double)s[o];
        if (a.sbase) sv += a.sbase[k];
        const double T = thermo::eos_dry(p0, sv);
        const double al = thermo::alpha_c(p0, T, 0.0, 0.0);
        To[o] = (TS)T;
        alo[o] = (TS)al;
        double b = 0.0;
        if (k < g.nzg - 1) {
            b = thermo::buoyancy_c(a.alpha0_half[k], al);
            if (bo) bo[o] = (TS)b;
        }
        // wt[k-1] += 0.5*(b[k-1] + b[k]) for k-1 in [0, nzg-2)
        if (k >= 1 && k - 1 < g.nzg - 2) {
            const long long om = o - g.plane;
            wt[om] = (TS)((double)wt[om] + 0.5 * (bprev + b));
        }
        bprev = b;
        double th = 0.0;
        if (k >= 1) {
            th = T / thermo::exner_c(p0);
            if (tho) tho[o] = (TS)th;
        }
        thw[0] = thw[1]; thw[1] = thw[2]; thw[2] = th;
        const int m = k - 1;
        if (bvfo && m >= 2 && m < g.nzg - 1)
            bvfo[o - g.plane] = (TS)(PC_G / thw[1] * (0.5 * (thw[1] + thw[2]) - 0.5 * (thw[0] + thw[1])) * dzi);
    }
}

extern "C" int pycles_thermo_dry(const pycles_dims *dims, const pycles_ref *ref, void *s, void *wt, void *T,
                                 void *alpha, void *buoyancy, void *theta, void *bvf,
                                 const pycles_scalar_base *base, int precision, void *stream) {
    int rc = pycles_check_dims(dims);
    if (rc) return rc;
    if (!ref || !s || !wt || !T || !alpha) {
        pycles_set_error("thermo_dry: NULL argument");
        return PYCLES_ERR_ARG;
    }
    Geo g = make_geo(dims);
    DryArgs a;
    a.p0 = ref->p0_half;
    a.alpha0_half = ref->alpha0_half;
    a.s = s; a.wt = wt; a.T = T; a.alpha = alpha; a.buoy = buoyancy; a.theta = theta; a.bvf = bvf;
    a.sbase = (precision == PYCLES_F32 && base) ? base->base[0] : nullptr;
    dim3 blk(64, 4), grd((g.nx + 63) / 64, (g.ny + 3) / 4);
    hipStream_t st = (hipStream_t)stream;
    if (precision == PYCLES_F64)
        hipLaunchKernelGGL(thermo_dry_kernel<double>, grd, blk, 0, st, g, a);
    else if (precision == PYCLES_F32)
        hipLaunchKernelGGL(thermo_dry_kernel<float>, grd, blk, 0, st, g, a);
    else {
        pycles_set_error("bad precision %d", precision);
        return PYCLES_ERR_ARG;
    }
    PYCLES_CHECK_LAUNCH();
    return PYCLES_OK;
}
