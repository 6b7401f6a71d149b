// Entry point of the fused advection passes (see advection_impl.h).
#include "advection_impl.h"

extern int pycles_check_dims(const pycles_dims *d);
int adv_dispatch_f64(const Geo &g, const adv::Prof &pr, const adv::Fields &fl, int nsc, bool mom, int so, int mo,
                     hipStream_t st);
int adv_dispatch_f32(const Geo &g, const adv::Prof &pr, const adv::Fields &fl, int nsc, bool mom, int so, int mo,
                     hipStream_t st);

// Shared by both entry points; rk: optional fused RK3 stage (rk->rk >= 0).
static int advection_run(const pycles_dims *dims, const pycles_ref *ref, const void *const vel[3],
                         void *const tend_vel[3], int nscal, const void *const *scal, void *const *tend_scal,
                         const pycles_scalar_base *base, int scalar_order, int momentum_order, int precision,
                         int flags, const adv::Fields *rk, void *stream) {
    int rc = pycles_check_dims(dims);
    if (rc) return rc;
    if (!ref || !vel || !vel[0] || !vel[1] || !vel[2]) {
        pycles_set_error("advection: NULL velocity or profile argument");
        return PYCLES_ERR_ARG;
    }
    if (nscal < 0 || nscal > 2) {
        pycles_set_error("advection: nscal = %d not in [0, 2]", nscal);
        return PYCLES_ERR_ARG;
    }
    const bool mom = tend_vel != nullptr;
    if (mom && (!tend_vel[0] || !tend_vel[1] || !tend_vel[2])) {
        pycles_set_error("advection: NULL velocity tendency");
        return PYCLES_ERR_ARG;
    }
    if (!mom && nscal == 0) return PYCLES_OK;
    for (int s = 0; s < nscal; ++s)
        if (!scal || !scal[s] || !tend_scal || !tend_scal[s]) {
            pycles_set_error("advection: NULL scalar %d", s);
            return PYCLES_ERR_ARG;
        }
    const int so = nscal ? scalar_order : momentum_order;
    const int mo = mom ? momentum_order : scalar_order;
    for (int o : {so, mo})
        if (o != 5 && o != 7 && o != 9) {
            pycles_set_error("advection: order %d not supported (WENO 5, 7, 9)", o);
            return PYCLES_ERR_UNSUPPORTED;
        }
    const int r = ((so > mo ? so : mo) + 1) / 2;
    if (dims->gw < r) {
        pycles_set_error("advection: order %d needs gw >= %d (gw = %lld)", so > mo ? so : mo, r, (long long)dims->gw);
        return PYCLES_ERR_ARG;
    }
    const Geo g = make_geo(dims);
    if (g.nx < 2 * r || g.ny < 2 * r) {
        pycles_set_error("advection: nx, ny must be at least %d", 2 * r);
        return PYCLES_ERR_ARG;
    }
    adv::Prof pr{ref->rho0, ref->rho0_half, ref->alpha0, ref->alpha0_half};
    adv::Fields fl;
    memset(&fl, 0, sizeof(fl));
    fl.tzero = (flags & PYCLES_ADV_TEND_ZERO) ? 1 : 0;
    fl.passes = (flags >> 1) & 7;
    if (!fl.passes) fl.passes = 7;
    fl.fuse_yz = (flags & PYCLES_ADV_FUSE_YZ) ? 1 : 0;
    for (int s = 0; s < nscal; ++s) {
        fl.q[s] = scal[s];
        fl.t[s] = tend_scal[s];
        fl.base[s] = (precision == PYCLES_F32 && base) ? base->base[s] : nullptr;
    }
    for (int d = 0; d < 3; ++d) {
        fl.q[adv::VU + d] = vel[d];
        fl.t[adv::VU + d] = mom ? tend_vel[d] : nullptr;
    }
    fl.rk = -1;
    if (rk) {
        fl.rk = rk->rk;
        fl.dt = rk->dt;
        for (int v = 0; v < 5; ++v) {
            fl.out[v] = rk->out[v];
            fl.cp[v] = rk->cp[v];
        }
    }
    hipStream_t st = (hipStream_t)stream;
    const size_t esz = precision == PYCLES_F64 ? 8 : 4;
    if ((size_t)g.plane * g.nzg * esz >= ((size_t)1 << 32)) {
        pycles_set_error("advection: one field exceeds 4 GiB (32-bit offsets)");
        return PYCLES_ERR_UNSUPPORTED;
    }
    if (precision == PYCLES_F64) return adv_dispatch_f64(g, pr, fl, nscal, mom, so, mo, st);
    if (precision == PYCLES_F32) return adv_dispatch_f32(g, pr, fl, nscal, mom, so, mo, st);
    pycles_set_error("bad precision %d", precision);
    return PYCLES_ERR_ARG;
}

extern "C" int pycles_advection(const pycles_dims *dims, const pycles_ref *ref, const void *const vel[3],
                                void *const tend_vel[3], int nscal, const void *const *scal, void *const *tend_scal,
                                const pycles_scalar_base *base, int scalar_order, int momentum_order, int precision,
                                int flags, void *stream) {
    return advection_run(dims, ref, vel, tend_vel, nscal, scal, tend_scal, base, scalar_order, momentum_order,
                         precision, flags, nullptr, stream);
}

extern "C" int pycles_advection_rk3(const pycles_dims *dims, const pycles_ref *ref, const void *const vel[3],
                                    void *const tend_vel[3], int nscal, const void *const *scal,
                                    void *const *tend_scal, const pycles_scalar_base *base, int scalar_order,
                                    int momentum_order, int precision, int flags, void *const out_vel[3],
                                    void *const *out_scal, void *const copy_vel[3], void *const *copy_scal,
                                    int rk_step, double dt, void *stream) {
    if (rk_step < 0 || rk_step > 2) {
        pycles_set_error("advection_rk3: rk_step %d not in [0, 2]", rk_step);
        return PYCLES_ERR_ARG;
    }
    const bool need_copy = rk_step > 0;  // at stage 0 the copies may be NULL (none written)
    if (!tend_vel || !out_vel || (need_copy && !copy_vel) || nscal < 0 || nscal > 2 ||
        (nscal > 0 && (!out_scal || (need_copy && !copy_scal)))) {
        pycles_set_error("advection_rk3: needs the momentum and an output (and, at rk_step 1 and 2, a copy) "
                         "array for every field");
        return PYCLES_ERR_ARG;
    }
    if (flags & (PYCLES_ADV_PASS_X | PYCLES_ADV_PASS_Y | PYCLES_ADV_PASS_Z)) {
        if (!(flags & PYCLES_ADV_PASS_Z)) {
            // passes without the z pass do no stage update
            return advection_run(dims, ref, vel, tend_vel, nscal, scal, tend_scal, base, scalar_order,
                                 momentum_order, precision, flags, nullptr, stream);
        }
    }
    adv::Fields rk;
    memset(&rk, 0, sizeof(rk));
    rk.rk = rk_step;
    rk.dt = dt;
    for (int s = 0; s < nscal; ++s) {
        rk.out[s] = out_scal[s];
        rk.cp[s] = copy_scal ? copy_scal[s] : nullptr;
    }
    for (int d = 0; d < 3; ++d) {
        rk.out[adv::VU + d] = out_vel[d];
        rk.cp[adv::VU + d] = copy_vel ? copy_vel[d] : nullptr;
    }
    for (int v = 0; v < 5; ++v) {
        if (v < nscal || v >= adv::VU) {
            const void *in = v < nscal ? scal[v] : vel[v - adv::VU];
            if (!rk.out[v] || (need_copy && !rk.cp[v]) || rk.out[v] == in || rk.cp[v] == in) {
                pycles_set_error("advection_rk3: output or copy %d is NULL or the input array itself", v);
                return PYCLES_ERR_ARG;
            }
        }
    }
    return advection_run(dims, ref, vel, tend_vel, nscal, scal, tend_scal, base, scalar_order, momentum_order,
                         precision, flags, &rk, stream);
}
