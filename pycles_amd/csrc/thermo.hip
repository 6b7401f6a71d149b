// Thermodynamics kernels: saturation adjustment (SA) and dry.
//
// Each thread marches one (i, j) column in k, so the vertical neighbours
// needed by the buoyancy average into the w tendency and by the buoyancy
// frequency are its own previous levels.  Columns are split into z chunks
// (grid z) for parallelism: a thread owns levels [k0, k1), starts one level
// early to fill its windows and reads one level past k1; it writes only the
// outputs of the levels it owns.  The exner function of every level is
// computed once per block into LDS, with the other per-level values (p0,
// alpha0): a vector load of one inside the level loop would make each level
// wait for all the loads and stores before it (the memory counter completes
// in order).  s and qt of the next level are loaded one level ahead.
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
#include "thermo_common.h"

// ---------------------------------------------------------------------------
// FP64: the reference's arithmetic.  Dynamic LDS: the pv* table, then the
// chunk's exner, p0 and alpha0 values.  WTO: the w tendency is overwritten,
// not added to (no read of it).
template <bool WTO>
__global__ __launch_bounds__(256) void thermo_sa_f64_kernel(Geo g, SAArgs a) {
    extern __shared__ double dyn64[];
    const int nl = a.kchunk + 2;  // levels of a chunk with its two neighbours
    double *tab = dyn64;
    double *exner = tab + a.ntab;
    double *p0l = exner + nl;
    double *a0l = p0l + nl;
    __shared__ unsigned long long cnt[4];
    const int tid = threadIdx.y * blockDim.x + threadIdx.x;
    const int nthr = blockDim.x * blockDim.y;
    const Chunk ch = chunk_of(g, a.kchunk);
    for (int t = tid; t < a.ntab; t += nthr) tab[t] = a.tab[t];
    for (int k = ch.kw + tid; k < ch.ke; k += nthr) {
        const int l = k - ch.kw;
        exner[l] = thermo::exner_c(a.p0[k]);
        p0l[l] = a.p0[k];
        a0l[l] = a.alpha0_half[k];
    }
    if (tid < 4) cnt[tid] = 0;
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    unsigned long long ctr[4] = {0, 0, 0, 0};
    if (i < g.nx && j < g.ny) {
        const long long col = (long long)j * g.nx + i;
        double *s = (double *)a.s;
        double *qtp = (double *)a.qt;
        double *wt = (double *)a.wt;
        double *To = (double *)a.T, *qvo = (double *)a.qv, *qlo = (double *)a.ql, *alo = (double *)a.alpha;
        double *bo = (double *)a.buoy, *thro = (double *)a.thr, *bvfo = (double *)a.bvf, *thlo = (double *)a.thl;
        const thermo::Table lt{tab, a.x_min, a.dx, a.dxi, a.ntab};
        const double dzi = 1.0 / g.dx[2];
        double Tw[3] = {0, 0, 0}, qtw[3] = {0, 0, 0}, qvw[3] = {0, 0, 0}, thw[3] = {0, 0, 0}, bprev = 0.0;
        // the next level's s and qt are loaded at every level (the level
        // clamped to the chunk), so the wait for them counts only these loads
        double s_nx = s[(long long)ch.kw * g.plane + col];
        double q_nx = qtp[(long long)ch.kw * g.plane + col];
        for (int k = ch.kw; k < ch.ke; ++k) {
            const int l = k - ch.kw;
            const bool own = k >= ch.k0 && k < ch.k1;
            const bool own_m = k - 1 >= ch.k0 && k - 1 < ch.k1;
            const long long o = (long long)k * g.plane + col;
            const double p0 = p0l[l];
            const double sk = s_nx;
            double qt = q_nx;
            {
                const long long on = (long long)min(k + 1, ch.ke - 1) * g.plane + col;
                s_nx = s[on];
                q_nx = qtp[on];
            }
            if (a.clip) {
                const double c = fmax(qt, 1e-11);
                if (!(c == qt)) {
                    qt = c;
                    if (own) qtp[o] = c;
                }
            }
            double T, qv, ql, qi;
            const int it = thermo::eos_sa<thermo::DivFast>(lt, p0, sk, qt, &T, &qv, &ql, &qi, a.max_iter, a.sa_tol);
            if (it > 0 && own) {
                ctr[0] += 1;
                ctr[1] += it;
                if (a.max_iter > 0 && it >= a.max_iter) ctr[2] += 1;
            }
            const double al = thermo::alpha_c<thermo::DivFast>(p0, T, qt, qv);
            double b = 0.0;
            if (k < g.nzg - 1) b = thermo::buoyancy_c<thermo::DivFast>(a0l[l], al);
            const double ex = exner[l];
            double thr = 0.0;
            if (k >= 1 && k < g.nzg - 1) thr = T * (1.0 - qt + PC_EPS_VI * qv) / ex;
            if (own) {
                To[o] = T;
                qvo[o] = qv;
                qlo[o] = ql;
                alo[o] = al;
                if (bo && k < g.nzg - 1) bo[o] = b;
                if (thro && k >= 1 && k < g.nzg - 1) thro[o] = thr;
                if (thlo) {
                    // exp(-L*0/(T cpd)) is exactly 1 for unsaturated cells
                    if (ql == 0.0 && qi == 0.0) thlo[o] = T / ex * 1.0;
                    else {
                        const double L = thermo::latent_heat_variable(T);
                        thlo[o] = T / ex * exp(-L * (ql / (1.0 - qt) + qi / (1.0 - qt)) / (T * PC_CPD));
                    }
                }
            }
            // w tendency at k-1 for k-1 in [1, nzg-3): wt + interp_2(b[k-1], b[k])
            if (own_m && k >= 2 && k - 1 < g.nzg - 3) {
                const double w0 = WTO ? 0.0 : wt[o - g.plane];
                wt[o - g.plane] = w0 + 0.5 * (bprev + b);
            }
            bprev = b;
            Tw[0] = Tw[1]; Tw[1] = Tw[2]; Tw[2] = T;
            qtw[0] = qtw[1]; qtw[1] = qtw[2]; qtw[2] = qt;
            qvw[0] = qvw[1]; qvw[1] = qvw[2]; qvw[2] = qv;
            thw[0] = thw[1]; thw[1] = thw[2]; thw[2] = thr;
            const int m = k - 1;
            if (bvfo && own_m && m >= 2 && m < g.nzg - 2)
                bvfo[o - g.plane] = thermo::bvf_sa<thermo::DivFast>(lt, p0l[l - 1], Tw, qtw, qvw, thw, dzi);
        }
    }
    flush_counters(cnt, ctr, a.counters, tid);
}

extern int pycles_check_dims(const pycles_dims *d);

int thermo_sa_f32_launch(const Geo &g, const SAArgs &a, dim3 grd, dim3 blk, hipStream_t st);


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
    if (g.nzg > LEVELS_MAX) {
        pycles_set_error("thermo_sa: %d levels exceed %d", g.nzg, LEVELS_MAX);
        return PYCLES_ERR_UNSUPPORTED;
    }
    if (opt->mlp && precision != PYCLES_F32) {
        pycles_set_error("thermo_sa: the learned first guess is part of the FP32 form only");
        return PYCLES_ERR_ARG;
    }
    if (opt->mlp && (opt->newton_steps < 1 || opt->newton_steps > 4)) {
        pycles_set_error("thermo_sa: newton_steps = %d not in [1, 4]", opt->newton_steps);
        return PYCLES_ERR_ARG;
    }
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
    a.wt_over = opt->wt_overwrite;
    a.mlp = (const mlp::Weights *)opt->mlp;
    a.newton = opt->newton_steps;
    a.sa_tol = opt->secant_tol > 0.0 ? opt->secant_tol : 1.0e-3;
    a.counters = (unsigned long long *)opt->counters;
    const ZSplit zs = thermo_zsplit(g.nx, g.ny, g.nzg);
    a.kchunk = zs.len;
    const int nch = zs.nch;
    dim3 blk(64, 4), grd((g.nx + 63) / 64, (g.ny + 3) / 4, nch);
    hipStream_t st = (hipStream_t)stream;
    if (precision == PYCLES_F64) {
        const size_t smem = (size_t)(a.ntab + 3 * (a.kchunk + 2)) * sizeof(double);
        if (a.wt_over)
            hipLaunchKernelGGL(thermo_sa_f64_kernel<true>, grd, blk, smem, st, g, a);
        else
            hipLaunchKernelGGL(thermo_sa_f64_kernel<false>, grd, blk, smem, st, g, a);
    }
    else if (precision == PYCLES_F32)
        return thermo_sa_f32_launch(g, a, grd, blk, st);
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
    int wt_over, kchunk;
};

template <typename TS>
__global__ __launch_bounds__(256) void thermo_dry_kernel(Geo g, DryArgs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= g.nx || j >= g.ny) return;
    const Chunk ch = chunk_of(g, a.kchunk);
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
    for (int k = ch.kw; k < ch.ke; ++k) {
        const bool own = k >= ch.k0 && k < ch.k1;
        const bool own_m = k - 1 >= ch.k0 && k - 1 < ch.k1;
        const long long o = (long long)k * g.plane + col;
        const double p0 = a.p0[k];
        double sv = (double)s[o];
        if (a.sbase) sv += a.sbase[k];
        const double T = thermo::eos_dry<thermo::DivFast>(p0, sv);
        const double al = thermo::alpha_c<thermo::DivFast>(p0, T, 0.0, 0.0);
        double b = 0.0;
        if (k < g.nzg - 1) b = thermo::buoyancy_c<thermo::DivFast>(a.alpha0_half[k], al);
        double th = 0.0;
        if (k >= 1) th = T / thermo::exner_c(p0);
        if (own) {
            To[o] = (TS)T;
            alo[o] = (TS)al;
            if (bo && k < g.nzg - 1) bo[o] = (TS)b;
            if (tho && k >= 1) tho[o] = (TS)th;
        }
        // wt[k-1] += 0.5*(b[k-1] + b[k]) for k-1 in [0, nzg-2)
        if (own_m && k >= 1 && k - 1 < g.nzg - 2) {
            const long long om = o - g.plane;
            const double w0 = a.wt_over ? 0.0 : (double)wt[om];
            wt[om] = (TS)(w0 + 0.5 * (bprev + b));
        }
        bprev = b;
        thw[0] = thw[1]; thw[1] = thw[2]; thw[2] = th;
        const int m = k - 1;
        if (bvfo && own_m && m >= 2 && m < g.nzg - 1)
            bvfo[o - g.plane] = (TS)(PC_G / thw[1] * (0.5 * (thw[1] + thw[2]) - 0.5 * (thw[0] + thw[1])) * dzi);
    }
}

extern "C" int pycles_thermo_dry(const pycles_dims *dims, const pycles_ref *ref, void *s, void *wt, void *T,
                                 void *alpha, void *buoyancy, void *theta, void *bvf,
                                 const pycles_scalar_base *base, int precision, int wt_overwrite, void *stream) {
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
    a.wt_over = wt_overwrite;
    const ZSplit zs = thermo_zsplit(g.nx, g.ny, g.nzg);
    a.kchunk = zs.len;
    const int nch = zs.nch;
    dim3 blk(64, 4), grd((g.nx + 63) / 64, (g.ny + 3) / 4, nch);
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
