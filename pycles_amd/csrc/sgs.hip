// Kinematics, Smagorinsky SGS closure and second-order diffusion.
//
//   Kinematics.update                Kinematics.pyx:41-58, kinematics.h:6-138
//   Smagorinsky.update               SGS.pyx:146-152, sgs.h:4-19
//   ScalarDiffusion.update           ScalarDiffusion.pyx:76-119,
//                                    scalar_diffusion.h:6-58, flux_divergence.h:4-32
//   MomentumDiffusion.update         MomentumDiffusion.pyx:44-66,
//                                    momentum_diffusion.h:5-118, flux_divergence.h:34-80
//
// Two kernels per stage, both one thread per cell with x-fastest rows:
//  * sgs_kernel: the strain-rate magnitude from the velocity gradients at the
//    staggered points around the cell, then the eddy viscosity and
//    diffusivity.  The reference stores the nine gradients and the nine
//    strain-rate components (18 field passes written and read again); here
//    they are formed from the velocities in registers and only |S|, nu and D
//    are written.
//  * diffusion_kernel: for each cell, the diffusive fluxes of s and qt on
//    its six faces and the nine stress fluxes on the faces the momentum
//    divergence needs (each flux formed twice, once for each cell sharing
//    the face, with no flux array), the divergences accumulated into the
//    tendencies in the reference's per-variable order, and the entropy
//    source of the dissipation.
// x and y are periodic by index wrap; z reads the stored ghost levels.  In
// FP64 every expression follows the reference's operation order.
#include "common.h"

extern int pycles_check_dims(const pycles_dims *d);

namespace {

// Offset of the point (di, dj, dk) from cell (i, j, k), each in {-2..1}.
__device__ __forceinline__ long long off(const Geo &g, int i, int j, int k, int di, int dj, int dk) {
    int ii = i + di, jj = j + dj;
    ii = ii < 0 ? ii + g.nx : (ii >= g.nx ? ii - g.nx : ii);
    jj = jj < 0 ? jj + g.ny : (jj >= g.ny ? jj - g.ny : jj);
    return ((long long)(k + dk) * g.ny + jj) * g.nx + ii;
}

// Velocity gradient dv_a/dx_b at the point p = (i, j, k) + d: the
// reference's vgrad[a][b][p] = (v_a[p + s_b] - v_a[p]) * dxi_b.
template <typename T>
struct Vel {
    const T *v[3];
    __device__ __forceinline__ T grad(const Geo &g, int i, int j, int k, int a, int b, int di, int dj, int dk,
                                      T dxi) const {
        const long long p = off(g, i, j, k, di, dj, dk);
        const long long q = off(g, i, j, k, di + (b == 0), dj + (b == 1), dk + (b == 2));
        return (v[a][q] - v[a][p]) * dxi;
    }
    // strain rate S_ab at p (kinematics.h:54-62): 0.5 (grad_ab + grad_ba)
    __device__ __forceinline__ T strain(const Geo &g, int i, int j, int k, int a, int b, int di, int dj, int dk,
                                        const T *dxi) const {
        return T(0.5) * (grad(g, i, j, k, a, b, di, dj, dk, dxi[b]) + grad(g, i, j, k, b, a, di, dj, dk, dxi[a]));
    }
};

template <typename T>
__global__ __launch_bounds__(256) void sgs_kernel(Geo g, Vel<T> vel, const T *bvf, T *smag_out, T *visc, T *diff,
                                                  double csd2, double prt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    const int k = blockIdx.z;
    if (i >= g.nx) return;
    const long long o = ((long long)k * g.ny + j) * g.nx + i;
    const T dxi[3] = {(T)g.dxi[0], (T)g.dxi[1], (T)g.dxi[2]};
    T m = T(0);
    // the reference computes |S| on [1, nlg-1) in z (0 elsewhere); x and y
    // are periodic, so every column has its neighbours
    if (k >= 1 && k < g.nzg - 1) {
        // on-diagonal components at ijk - s_d (kinematics.h:96-108)
        for (int d = 0; d < 3; ++d) {
            const T s = vel.strain(g, i, j, k, d, d, -(d == 0), -(d == 1), -(d == 2), dxi);
            m += s * s;
        }
        // the (v1, d) pairs with d from v1 to 2 (kinematics.h:111-128): four
        // points ijk, ijk - s_v1, ijk - s_d, ijk - s_d - s_v1
        for (int v1 = 0; v1 < 2; ++v1)
            for (int d = v1; d < 3; ++d) {
                const int a0 = -(v1 == 0), a1 = -(v1 == 1), a2 = -(v1 == 2);
                const int b0 = -(d == 0), b1 = -(d == 1), b2 = -(d == 2);
                const T s1 = vel.strain(g, i, j, k, v1, d, 0, 0, 0, dxi);
                const T s2 = vel.strain(g, i, j, k, v1, d, a0, a1, a2, dxi);
                const T s3 = vel.strain(g, i, j, k, v1, d, b0, b1, b2, dxi);
                const T s4 = vel.strain(g, i, j, k, v1, d, a0 + b0, a1 + b1, a2 + b2, dxi);
                const T si = T(0.25) * (s1 + s2 + s3 + s4);
                m += T(2.0) * si * si;
            }
    }
    m = sqrt(T(2.0) * m);
    // smagorinsky_update (sgs.h:4-19); csd2 = cs*cs*delta*delta as the
    // reference evaluates it left to right
    const T S = m;
    T nu = (T)csd2 * S;
    const T n2 = bvf[o];
    if (n2 > T(0)) {
        const T fb = sqrt(fmax(T(1.0) - n2 / ((T)prt * S * S), T(0.0)));
        nu = nu * fb;
    }
    smag_out[o] = S;
    visc[o] = nu;
    diff[o] = nu / (T)prt;
}

template <typename T>
struct DiffArgs {
    Vel<T> vel;
    T *tv[3];
    const T *phi[2];
    T *tphi[2];
    const double *base[2];
    int nscal, flags;
    const T *visc, *diff, *smag, *temp;
    const double *rho0, *rho0_half, *alpha0, *alpha0_half;
};

// Scalar flux -interp_2(D[p], D[p+s]) (phi[p+s] - phi[p]) rho dxi on the face
// above p along d (scalar_diffusion.h:21-43), zero on the surface face.
template <typename T>
__device__ __forceinline__ T scalar_flux(const Geo &g, const DiffArgs<T> &a, int sc, int i, int j, int k, int d,
                                         int di, int dj, int dk) {
    const int kp = k + dk;
    if (d == 2 && kp == g.gw - 1) return T(0);
    const long long p = off(g, i, j, k, di, dj, dk);
    const long long q = off(g, i, j, k, di + (d == 0), dj + (d == 1), dk + (d == 2));
    const T dxi = (T)(1.0 / g.dx[d]);
    const T K = T(0.5) * (a.diff[p] + a.diff[q]);
    T dphi;
    if (sizeof(T) == 4 && d == 2 && a.base[sc]) {
        // FP32 deviations from the level profile: the profile's difference in FP64
        dphi = (T)((a.base[sc][kp + 1] - a.base[sc][kp]) + (double)(a.phi[sc][q] - a.phi[sc][p]));
    } else {
        dphi = a.phi[sc][q] - a.phi[sc][p];
    }
    const T rho = d == 2 ? (T)a.rho0[kp] : (T)a.rho0_half[kp];
    return T(-1.0) * K * dphi * rho * dxi;
}

// Stress flux of velocity i1 along i2 at p (compute_diffusive_flux_m,
// momentum_diffusion.h:5-88), zero on the surface face.
template <typename T>
__device__ __forceinline__ T stress_flux(const Geo &g, const DiffArgs<T> &a, int i1, int i2, int i, int j, int k,
                                         int di, int dj, int dk, const T *dxi) {
    const int kp = k + dk;
    if (kp == g.gw - 1) return T(0);
    const T S = a.vel.strain(g, i, j, k, i1, i2, di, dj, dk, dxi);
    const int e0 = (i1 == 0), e1 = (i1 == 1), e2 = (i1 == 2);
    const int f0 = (i2 == 0), f1 = (i2 == 1), f2 = (i2 == 2);
    if (i1 == i2) {
        const T nu = a.visc[off(g, i, j, k, di + e0, dj + e1, dk + e2)];
        const double rho = i1 == 2 ? a.rho0_half[kp + 1] : a.rho0_half[kp];
        return T(-2.0) * S * nu * (T)rho;
    }
    const T nu = T(0.25) * (a.visc[off(g, i, j, k, di, dj, dk)] + a.visc[off(g, i, j, k, di + e0, dj + e1, dk + e2)] +
                            a.visc[off(g, i, j, k, di + f0, dj + f1, dk + f2)] +
                            a.visc[off(g, i, j, k, di + e0 + f0, dj + e1 + f1, dk + e2 + f2)]);
    const double rho = (i1 != 2 && i2 != 2) ? a.rho0_half[kp] : a.rho0[kp];
    return T(-2.0) * S * nu * (T)rho;
}

template <typename T>
__global__ __launch_bounds__(256) void diffusion_kernel(Geo g, DiffArgs<T> a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    const int k = blockIdx.z + g.gw;
    if (i >= g.nx) return;
    const long long o = ((long long)k * g.ny + j) * g.nx + i;
    const T dxi[3] = {(T)g.dxi[0], (T)g.dxi[1], (T)g.dxi[2]};
    const T ah = (T)a.alpha0_half[k], af = (T)a.alpha0[k];
    // ScalarDiffusion: each scalar, directions x, y, z (flux_divergence.h:4-32)
    for (int sc = 0; (a.flags & PYCLES_DIFF_SCALARS) && sc < a.nscal; ++sc) {
        T t = a.tphi[sc][o];
        for (int d = 0; d < 3; ++d) {
            const T fu = scalar_flux(g, a, sc, i, j, k, d, 0, 0, 0);
            const T fd = scalar_flux(g, a, sc, i, j, k, d, -(d == 0), -(d == 1), -(d == 2));
            t -= ah * (fu - fd) * dxi[d];
        }
        a.tphi[sc][o] = t;
    }
    // MomentumDiffusion: component i1, directions i2 (flux_divergence.h:34-80)
    for (int i1 = 0; (a.flags & PYCLES_DIFF_MOMENTUM) && i1 < 3; ++i1) {
        const T al = i1 == 2 ? af : ah;
        T t = a.tv[i1][o];
        for (int i2 = 0; i2 < 3; ++i2) {
            const T fu = stress_flux(g, a, i1, i2, i, j, k, 0, 0, 0, dxi);
            const T fd = stress_flux(g, a, i1, i2, i, j, k, -(i2 == 0), -(i2 == 1), -(i2 == 2), dxi);
            t -= al * (fu - fd) * dxi[i2];
        }
        a.tv[i1][o] = t;
    }
    // compute_entropy_source (momentum_diffusion.h:90-118) into the s tendency
    if (a.flags & PYCLES_DIFF_ENTROPY_SOURCE) {
        const T S = a.smag[o];
        a.tphi[0][o] += S * S * a.visc[o] / a.temp[o];
    }
}

}  // namespace

extern "C" int pycles_sgs_smagorinsky(const pycles_dims *dims, const void *const vel[3], const void *bvf, double cs,
                                      double prt, void *strain_mag, void *visc, void *diff, int precision,
                                      void *stream) {
    int rc = pycles_check_dims(dims);
    if (rc) return rc;
    if (!vel || !vel[0] || !vel[1] || !vel[2] || !bvf || !strain_mag || !visc || !diff) {
        pycles_set_error("sgs_smagorinsky: NULL argument");
        return PYCLES_ERR_ARG;
    }
    if (precision != PYCLES_F64 && precision != PYCLES_F32) {
        pycles_set_error("sgs_smagorinsky: bad precision %d", precision);
        return PYCLES_ERR_ARG;
    }
    if (!(prt > 0.0)) {
        pycles_set_error("sgs_smagorinsky: the turbulent Prandtl number must be positive");
        return PYCLES_ERR_ARG;
    }
    const Geo g = make_geo(dims);
    // sgs.h:7 on the host: cbrt and the product in the reference's order
    const double delta = cbrt(g.dx[0] * g.dx[1] * g.dx[2]);
    const double csd2 = cs * cs * delta * delta;
    const dim3 grid((g.nx + 255) / 256, g.ny, g.nzg);
    hipStream_t st = (hipStream_t)stream;
    if (precision == PYCLES_F64) {
        Vel<double> v{{(const double *)vel[0], (const double *)vel[1], (const double *)vel[2]}};
        hipLaunchKernelGGL(sgs_kernel<double>, grid, dim3(256), 0, st, g, v, (const double *)bvf,
                           (double *)strain_mag, (double *)visc, (double *)diff, csd2, prt);
    } else {
        Vel<float> v{{(const float *)vel[0], (const float *)vel[1], (const float *)vel[2]}};
        hipLaunchKernelGGL(sgs_kernel<float>, grid, dim3(256), 0, st, g, v, (const float *)bvf, (float *)strain_mag,
                           (float *)visc, (float *)diff, csd2, prt);
    }
    PYCLES_CHECK_LAUNCH();
    return PYCLES_OK;
}

template <typename T>
static int launch_diffusion(const Geo &g, const pycles_ref *ref, const void *const vel[3], void *const tend_vel[3],
                            int nscal, const void *const *scal, void *const *tend_scal,
                            const pycles_scalar_base *base, const void *visc, const void *diff,
                            const void *strain_mag, const void *T_, int flags, hipStream_t st) {
    DiffArgs<T> a;
    for (int q = 0; q < 3; ++q) {
        a.vel.v[q] = (const T *)vel[q];
        a.tv[q] = tend_vel ? (T *)tend_vel[q] : nullptr;
    }
    for (int q = 0; q < 2; ++q) {
        a.phi[q] = q < nscal ? (const T *)scal[q] : nullptr;
        a.tphi[q] = q < nscal ? (T *)tend_scal[q] : nullptr;
        a.base[q] = (sizeof(T) == 4 && base && q < nscal) ? base->base[q] : nullptr;
    }
    a.nscal = nscal;
    a.flags = flags;
    a.visc = (const T *)visc;
    a.diff = (const T *)diff;
    a.smag = (const T *)strain_mag;
    a.temp = (const T *)T_;
    a.rho0 = ref->rho0;
    a.rho0_half = ref->rho0_half;
    a.alpha0 = ref->alpha0;
    a.alpha0_half = ref->alpha0_half;
    const dim3 grid((g.nx + 255) / 256, g.ny, g.nz);
    hipLaunchKernelGGL(diffusion_kernel<T>, grid, dim3(256), 0, st, g, a);
    PYCLES_CHECK_LAUNCH();
    return PYCLES_OK;
}

extern "C" int pycles_diffusion(const pycles_dims *dims, const pycles_ref *ref, const void *const vel[3],
                                void *const tend_vel[3], int nscal, const void *const *scal, void *const *tend_scal,
                                const pycles_scalar_base *base, const void *visc, const void *diff,
                                const void *strain_mag, const void *temperature, int flags, int precision,
                                void *stream) {
    int rc = pycles_check_dims(dims);
    if (rc) return rc;
    const int known = PYCLES_DIFF_SCALARS | PYCLES_DIFF_MOMENTUM | PYCLES_DIFF_ENTROPY_SOURCE;
    if ((flags & ~known) || !flags) {
        pycles_set_error("diffusion: flags %d", flags);
        return PYCLES_ERR_ARG;
    }
    if (!ref || !ref->rho0 || !ref->rho0_half || !ref->alpha0 || !ref->alpha0_half || !vel || !vel[0] || !vel[1] ||
        !vel[2]) {
        pycles_set_error("diffusion: NULL reference profile or velocity");
        return PYCLES_ERR_ARG;
    }
    if (nscal < 0 || nscal > 2 || (nscal > 0 && (!scal || !tend_scal))) {
        pycles_set_error("diffusion: nscal = %d", nscal);
        return PYCLES_ERR_ARG;
    }
    for (int q = 0; q < nscal; ++q)
        if (!scal[q] || !tend_scal[q]) {
            pycles_set_error("diffusion: NULL scalar %d", q);
            return PYCLES_ERR_ARG;
        }
    if ((flags & PYCLES_DIFF_SCALARS) && (!diff || nscal == 0)) {
        pycles_set_error("diffusion: scalar diffusion needs the diffusivity and a scalar");
        return PYCLES_ERR_ARG;
    }
    if ((flags & PYCLES_DIFF_MOMENTUM) && (!tend_vel || !tend_vel[0] || !tend_vel[1] || !tend_vel[2] || !visc)) {
        pycles_set_error("diffusion: momentum diffusion needs the three tendencies and the viscosity");
        return PYCLES_ERR_ARG;
    }
    if ((flags & PYCLES_DIFF_ENTROPY_SOURCE) && (nscal == 0 || !strain_mag || !temperature || !visc)) {
        pycles_set_error("diffusion: the entropy source needs s, |S|, the viscosity and T");
        return PYCLES_ERR_ARG;
    }
    if (precision != PYCLES_F64 && precision != PYCLES_F32) {
        pycles_set_error("diffusion: bad precision %d", precision);
        return PYCLES_ERR_ARG;
    }
    if (dims->gw < 2) {
        pycles_set_error("diffusion: gw must be at least 2");
        return PYCLES_ERR_ARG;
    }
    const Geo g = make_geo(dims);
    hipStream_t st = (hipStream_t)stream;
    if (precision == PYCLES_F64)
        return launch_diffusion<double>(g, ref, vel, tend_vel, nscal, scal, tend_scal, base, visc, diff, strain_mag,
                                        temperature, flags, st);
    return launch_diffusion<float>(g, ref, vel, tend_vel, nscal, scal, tend_scal, base, visc, diff, strain_mag,
                                   temperature, flags, st);
}
