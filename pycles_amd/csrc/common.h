// Shared device definitions for the PyCLES dynamical-core kernels (gfx950).
//
// Device layout of one field: [k][j][i], x fastest, over nzg = nz + 2*gw
// levels (z ghost levels kept, as the reference keeps them) and the nx*ny
// interior columns.  x and y are periodic, so their neighbours come from
// index wrap instead of ghost copies.
#pragma once
#include <hip/hip_runtime.h>
#include <string.h>
#include <stdint.h>

#include <type_traits>

#include "../../include/pycles_amd.h"

// Physical constants, restated from reference generate_parameters.py:20-39.
// Derived values use the same double operations as the reference generator.
#define PC_PI 3.141592653589793
#define PC_G 9.80665
#define PC_RD 287.1
#define PC_RV 461.5
#define PC_EPS_V (PC_RD / PC_RV)
#define PC_EPS_VI (1.0 / PC_EPS_V)
#define PC_CPD 1004.0
#define PC_CPV 1859.0
#define PC_KAPPA (PC_RD / PC_CPD)
#define PC_T_TILDE 298.15
#define PC_P_TILDE 100000.0
#define PC_SD_TILDE 6864.8
#define PC_SV_TILDE 10513.6

struct Geo {
    int nx, ny, nz, gw, nzg;
    long long plane;  // nx*ny
    double dx[3], dxi[3];
};

static inline Geo make_geo(const pycles_dims *d) {
    Geo g;
    g.nx = (int)d->n[0];
    g.ny = (int)d->n[1];
    g.nz = (int)d->n[2];
    g.gw = (int)d->gw;
    g.nzg = g.nz + 2 * g.gw;
    g.plane = (long long)g.nx * g.ny;
    for (int i = 0; i < 3; ++i) {
        g.dx[i] = d->dx[i];
        g.dxi[i] = d->dxi[i];
    }
    return g;
}

__device__ __forceinline__ long long gidx(const Geo &g, int i, int j, int k) {
    return ((long long)k * g.ny + j) * g.nx + i;
}

// Split of n levels (or rows) into the chunks of a kernel grid's z dimension:
// nch chunks of len levels each, the last one holding what is left.
struct ZSplit {
    int nch, len;
};
// Chunks of ceil(n / want) levels, recounted so that none is empty; the last
// chunk holds n - (nch - 1) * len levels, between 1 and len.
static inline ZSplit zsplit_ceil(int n, int want) {
    want = want < 1 ? 1 : (want > n ? n : want);
    const int len = (n + want - 1) / want;
    return ZSplit{(n + len - 1) / len, len};
}
// Chunks of floor(n / nch) levels with the last one also taking the
// remainder (n - (nch - 1) * len levels), so that every chunk holds at least
// minlen levels; a single chunk when n < 2 * minlen.
static inline ZSplit zsplit_floor(int n, int want, int minlen) {
    const int cap = minlen > 0 ? n / minlen : n;
    int nch = want < cap ? want : cap;
    if (nch < 1) nch = 1;
    return ZSplit{nch, n / nch};
}
// The pressure correction's chunks of the interior levels (~16k waves; the
// last chunk holds the top gw levels, the first the bottom gw, since the
// boundary fill of each end reads them).
static inline ZSplit correct_zsplit(long long plane, int nz, int gw) {
    int want = 1;
    while (plane / 64 * want < 16384 && nz / (2 * want) >= 2 * gw && want < 16) want *= 2;
    return zsplit_floor(nz, want, gw);
}
// The advection y pass's chunks of the ny rows and the z pass's chunks of the
// interior levels (at least 2048 blocks, chunks of at least 32).
static inline ZSplit advy_zsplit(int nx, int ny, int nz) {
    const long long cols = (long long)((nx + 63) / 64) * ((nz + 3) / 4);
    int want = 1;
    while (cols * want < 2048 && ny / (want * 2) >= 32) want *= 2;
    return zsplit_ceil(ny, want);
}
static inline ZSplit advz_zsplit(int nx, int ny, int nz) {
    const long long cols = (long long)((nx + 63) / 64) * ((ny + 3) / 4);
    int want = 1;
    while (cols * want < 2048 && nz / (want * 2) >= 32) want *= 2;
    return zsplit_ceil(nz, want);
}
// The fused y-z advection pass: the z pass's chunks, none longer than
// ADV_YZ_KMAX levels (the LDS copy of a chunk's level profiles holds that many).
constexpr int ADV_YZ_KMAX = 128;
static inline ZSplit advyz_zsplit(int nx, int ny, int nz) {
    const ZSplit z = advz_zsplit(nx, ny, nz);
    return z.len > ADV_YZ_KMAX ? zsplit_ceil(nz, (nz + ADV_YZ_KMAX - 1) / ADV_YZ_KMAX) : z;
}
// The thermodynamics kernels' chunks of the nzg ghosted levels (~4096 blocks;
// their LDS holds len + 2 levels, so no chunk is longer than len).
static inline ZSplit thermo_zsplit(int nx, int ny, int nzg) {
    const long long cols = (long long)((nx + 63) / 64) * ((ny + 3) / 4);
    int want = 1;
    while (cols * want < 4096 && nzg / (want * 2) >= 16) want *= 2;
    return zsplit_ceil(nzg, want);
}

// CFL maximum of non-negative values.  A NaN or infinite value makes the
// result NaN: fmax would drop a NaN, and the host reads the result as the
// step's check that the velocities are finite.
__device__ __forceinline__ double cfl_fold(double m, double c) {
    return (c <= 1.7976931348623157e308 && m == m) ? fmax(m, c) : __longlong_as_double(0x7ff8000000000000LL);
}
// Block maximum into *out by an integer atomic max: non-negative doubles
// order like their bit patterns, and the positive quiet NaN above them all.
__device__ __forceinline__ void cfl_block_max(double m, double *out) {
    __shared__ double red[256];
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = cfl_fold(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0)
        atomicMax((unsigned long long *)out, (unsigned long long)__double_as_longlong(red[0]));
}

// Periodic wrap for offsets smaller than the extent.
__device__ __forceinline__ int wrapi(int i, int n) {
    return i < 0 ? i + n : (i >= n ? i - n : i);
}

// XCD-aware tile order.  The dispatcher deals workgroups round-robin over
// the 8 XCDs by linear workgroup id, and each XCD has its own L2.  This maps
// the linear id to a tile so that XCD x owns a contiguous band of grid
// dimension BAND (0: x, 1: y) over the whole of the other dimensions; tiles
// that read each other's cells along that dimension then share one L2.
// A bijection of the grid for any dispatch order; the plain order when the
// band dimension is not a multiple of 8.
template <int BAND>
__device__ __forceinline__ uint3 xcd_tile() {
    const unsigned gx = gridDim.x, gy = gridDim.y;
    const unsigned gb = BAND == 0 ? gx : gy;
    if (gb % 8u != 0u) return make_uint3(blockIdx.x, blockIdx.y, blockIdx.z);
    const unsigned L = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    const unsigned xcd = L % 8u, idx = L / 8u, band = gb / 8u;
    if (BAND == 0) {
        const unsigned t = idx / band;
        return make_uint3(xcd * band + idx % band, t % gy, t / gy);
    }
    const unsigned t = idx / gx;
    return make_uint3(idx % gx, xcd * band + t % band, t / band);
}

// ---------------------------------------------------------------------------
// Interpolants.  The expression trees follow reference
// Csrc/advection_interpolation.h (interp_4 :12-14, interp_6 :20-24,
// interp_8 :32-36, interp_weno5 :79-104, interp_weno7 :106-145,
// interp_weno9 :148-199) so that, compiled without contraction, the FP64
// instantiation reproduces the reference's rounding exactly.  q[] holds the
// stencil in the same argument order as the reference functions.

template <typename T> struct Num;
// a / b rounded to nearest, as the IEEE division gives it, for a finite
// (zero or of either sign) and b normal, with b's reciprocal and the quotient
// well inside the normal range.  Markstein's method: the hardware reciprocal estimate refined by two
// Newton steps (y += y (1 - b y)), the quotient q = a y, its exact remainder
// r = a - b q (one FMA) and q + r y, the last rounding giving the correctly
// rounded quotient.  The hardware division sequence is these FMAs between a
// range scaling of a and b (a no-op unless exponents lie within 2^53 of the
// range limits or 768 apart) and a fix-up of zero, infinite, NaN and denormal
// cases, which such operands never need: 8 instructions against 11, with
// the same result.
__device__ __forceinline__ double div64(double a, double b) {
    double y = __builtin_amdgcn_rcp(b);
    y = __builtin_fma(__builtin_fma(-b, y, 1.0), y, y);
    y = __builtin_fma(__builtin_fma(-b, y, 1.0), y, y);
    const double q = a * y;
    return __builtin_fma(__builtin_fma(-b, q, a), y, q);
}

template <> struct Num<double> {
    // The interpolants' operands: a in [1/126, 1], b >= 1e-20 and finite
    // (squared smoothness indicators plus epsilon, and sums of weights); the
    // advection tests hold the FP64 passes bitwise to the oracle's divisions.
    static __device__ __forceinline__ double div(double a, double b) { return div64(a, b); }
    static __device__ __forceinline__ double rcp(double b) { return div64(1.0, b); }
};
template <> struct Num<float> {
    // v_rcp_f32 (1 ulp): the FP32 path does not need a correctly rounded divide.
    static __device__ __forceinline__ float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
    static __device__ __forceinline__ float rcp(float b) { return __builtin_amdgcn_rcpf(b); }
};

template <typename T> __device__ __forceinline__ T sq_eps(T b) {
    return (b + T(1e-10)) * (b + T(1e-10));
}

template <int ORDER, typename T> struct Interp;

template <typename T> struct Interp<5, T> {
    static constexpr int R = 3;
    static __device__ __forceinline__ T weno(const T *q) {
        const T m2 = q[0], m1 = q[1], c = q[2], p1 = q[3], p2 = q[4];
        const T s0 = T(1.0 / 3.0) * m2 - T(7.0 / 6.0) * m1 + T(11.0 / 6.0) * c;
        const T s1 = T(-1.0 / 6.0) * m1 + T(5.0 / 6.0) * c + T(1.0 / 3.0) * p1;
        const T s2 = T(1.0 / 3.0) * c + T(5.0 / 6.0) * p1 - T(1.0 / 6.0) * p2;
        const T b2 = (T(13.0 / 12.0) * (c - T(2.0) * p1 + p2) * (c - T(2.0) * p1 + p2) +
                      T(0.25) * (T(3.0) * c - T(4.0) * p1 + p2) * (T(3.0) * c - T(4.0) * p1 + p2));
        const T b1 = (T(13.0 / 12.0) * (m1 - T(2.0) * c + p1) * (m1 - T(2.0) * c + p1) +
                      T(0.25) * (m1 - p1) * (m1 - p1));
        const T b0 = (T(13.0 / 12.0) * (m2 - T(2.0) * m1 + c) * (m2 - T(2.0) * m1 + c) +
                      T(0.25) * (m2 - T(4.0) * m1 + T(3.0) * c) * (m2 - T(4.0) * m1 + T(3.0) * c));
        const T a0 = Num<T>::div(T(0.1), sq_eps(b0));
        const T a1 = Num<T>::div(T(0.6), sq_eps(b1));
        const T a2 = Num<T>::div(T(0.3), sq_eps(b2));
        const T inv = Num<T>::rcp(a0 + a1 + a2);
        const T w0 = a0 * inv, w1 = a1 * inv, w2 = a2 * inv;
        return w0 * s0 + w1 * s1 + w2 * s2;
    }
    // weno() of a stencil whose centre value q[2] is zero (offsets from the
    // centre), for the FP32 form, rearranged for fewer operations (the same
    // function in exact arithmetic): the linear weights d = (0.1, 0.6, 0.3)
    // are folded into the candidate polynomials, the smoothness indicators
    // are taken times 4 with epsilon times 4 (the weights are unchanged by a
    // common factor), and the weighted sum and the sum of the weights are two
    // multiply-add chains over r_k = 1 / (4 b_k + 4 eps)^2 with one final
    // reciprocal: 34 operations and 4 reciprocals against 43 and 4.
    static __device__ __forceinline__ T weno_c0(const T *q) {
        const T m2 = q[0], m1 = q[1], p1 = q[3], p2 = q[4];
        const T s0 = T(0.1 / 3.0) * m2 - T(0.7 / 6.0) * m1;  // 0.1 ((1/3) m2 - (7/6) m1)
        const T s1 = T(-0.1) * m1 + T(0.2) * p1;             // 0.6 (-(1/6) m1 + (1/3) p1)
        const T s2 = T(0.25) * p1 - T(0.05) * p2;            // 0.3 ((5/6) p1 - (1/6) p2)
        const T e2 = p2 - T(2.0) * p1, f2 = p2 - T(4.0) * p1;
        const T e1 = m1 + p1, f1 = m1 - p1;
        const T e0 = m2 - T(2.0) * m1, f0 = m2 - T(4.0) * m1;
        const T B2 = (T(13.0 / 3.0) * e2) * e2 + f2 * f2;  // 4 b2
        const T B1 = (T(13.0 / 3.0) * e1) * e1 + f1 * f1;
        const T B0 = (T(13.0 / 3.0) * e0) * e0 + f0 * f0;
        const T t0 = B0 + T(4e-10), t1 = B1 + T(4e-10), t2 = B2 + T(4e-10);
        const T r0 = Num<T>::rcp(t0 * t0), r1 = Num<T>::rcp(t1 * t1), r2 = Num<T>::rcp(t2 * t2);
        const T num = s0 * r0 + s1 * r1 + s2 * r2;
        const T den = T(0.1) * r0 + T(0.6) * r1 + T(0.3) * r2;
        return num * Num<T>::rcp(den);
    }
    // interp_4: q = phim1, phi, phip1, phip2
    static __device__ __forceinline__ T centred(const T *q) {
        return T(7.0 / 12.0) * (q[1] + q[2]) - T(1.0 / 12.0) * (q[0] + q[3]);
    }
};

template <typename T> struct Interp<7, T> {
    static constexpr int R = 4;
    static __device__ __forceinline__ T weno(const T *q) {
        const T m3 = q[0], m2 = q[1], m1 = q[2], c = q[3], p1 = q[4], p2 = q[5], p3 = q[6];
        const T s0 = T(-1.0 / 4.0) * m3 + T(13.0 / 12.0) * m2 + T(-23.0 / 12.0) * m1 + T(25.0 / 12.0) * c;
        const T s1 = T(1.0 / 12.0) * m2 + T(-5.0 / 12.0) * m1 + T(13.0 / 12.0) * c + T(1.0 / 4.0) * p1;
        const T s2 = T(-1.0 / 12.0) * m1 + T(7.0 / 12.0) * c + T(7.0 / 12.0) * p1 + T(-1.0 / 12.0) * p2;
        const T s3 = T(1.0 / 4.0) * c + T(13.0 / 12.0) * p1 + T(-5.0 / 12.0) * p2 + T(1.0 / 12.0) * p3;
        const T b0 = (m3 * (T(547.0) * m3 - T(3882.0) * m2 + T(4642.0) * m1 - T(1854.0) * c) +
                      m2 * (T(7043.0) * m2 - T(17246.0) * m1 + T(7042.0) * c) +
                      m1 * (T(11003.0) * m1 - T(9402.0) * c) + T(2107.0) * c * c);
        const T b1 = (m2 * (T(267.0) * m2 - T(1642.0) * m1 + T(1602.0) * c - T(494.0) * p1) +
                      m1 * (T(2843.0) * m1 - T(5966.0) * c + T(1922.0) * p1) +
                      c * (T(3443.0) * c - T(2522.0) * p1) + T(547.0) * p1 * p1);
        const T b2 = (m1 * (T(547.0) * m1 - T(2522.0) * c + T(1922.0) * p1 - T(494.0) * p2) +
                      c * (T(3443.0) * c - T(5966.0) * p1 + T(1602.0) * p2) +
                      p1 * (T(2843.0) * p1 - T(1642.0) * p2) + T(267.0) * p2 * p2);
        const T b3 = (c * (T(2107.0) * c - T(9402.0) * p1 + T(7042.0) * p2 - T(1854.0) * p3) +
                      p1 * (T(11003.0) * p1 - T(17246.0) * p2 + T(4642.0) * p3) +
                      p2 * (T(7043.0) * p2 - T(3882.0) * p3) + T(547.0) * p3 * p3);
        const T a0 = Num<T>::div(T(1.0 / 35.0), sq_eps(b0));
        const T a1 = Num<T>::div(T(12.0 / 35.0), sq_eps(b1));
        const T a2 = Num<T>::div(T(18.0 / 35.0), sq_eps(b2));
        const T a3 = Num<T>::div(T(4.0 / 35.0), sq_eps(b3));
        const T inv = Num<T>::rcp(a0 + a1 + a2 + a3);
        const T w0 = a0 * inv, w1 = a1 * inv, w2 = a2 * inv, w3 = a3 * inv;
        return w0 * s0 + w1 * s1 + w2 * s2 + w3 * s3;
    }
    // interp_6: q = phim2 .. phip3
    // stencil with a zero centre value (no specialised form for this order)
    static __device__ __forceinline__ T weno_c0(const T *q) { return weno(q); }
    static __device__ __forceinline__ T centred(const T *q) {
        return (T(37.0 / 60.0) * (q[2] + q[3]) - T(2.0 / 15.0) * (q[1] + q[4]) +
                T(1.0 / 60.0) * (q[0] + q[5]));
    }
};

template <typename T> struct Interp<9, T> {
    static constexpr int R = 5;
    static __device__ __forceinline__ T weno(const T *q) {
        const T m4 = q[0], m3 = q[1], m2 = q[2], m1 = q[3], c = q[4];
        const T p1 = q[5], p2 = q[6], p3 = q[7], p4 = q[8];
        const T s0 = T(1.0 / 5.0) * m4 + T(-21.0 / 20.0) * m3 + T(137.0 / 60.0) * m2 + T(-163.0 / 60.0) * m1 + T(137.0 / 60.0) * c;
        const T s1 = T(-1.0 / 20.0) * m3 + T(17.0 / 60.0) * m2 + T(-43.0 / 60.0) * m1 + T(77.0 / 60.0) * c + T(1.0 / 5.0) * p1;
        const T s2 = T(1.0 / 30.0) * m2 + T(-13.0 / 60.0) * m1 + T(47.0 / 60.0) * c + T(9.0 / 20.0) * p1 + T(-1.0 / 20.0) * p2;
        const T s3 = T(-1.0 / 20.0) * m1 + T(9.0 / 20.0) * c + T(47.0 / 60.0) * p1 + T(-13.0 / 60.0) * p2 + T(1.0 / 30.0) * p3;
        const T s4 = T(1.0 / 5.0) * c + T(77.0 / 60.0) * p1 + T(-43.0 / 60.0) * p2 + T(17.0 / 60.0) * p3 + T(-1.0 / 20.0) * p4;
        const T b0 = (m4 * (T(22658.0) * m4 - T(208501.0) * m3 + T(364863.0) * m2 - T(288007.0) * m1 + T(86329.0) * c) +
                      m3 * (T(482963.0) * m3 - T(1704396.0) * m2 + T(1358458.0) * m1 - T(411487.0) * c) +
                      m2 * (T(1521393.0) * m2 - T(2462076.0) * m1 + T(758823.0) * c) +
                      m1 * (T(1020563.0) * m1 - T(649501.0) * c) + T(107918.0) * c * c);
        const T b1 = (m3 * (T(6908.0) * m3 - T(60871.0) * m2 + T(99213.0) * m1 - T(70237.0) * c + T(18079.0) * p1) +
                      m2 * (T(138563.0) * m2 - T(464976.0) * m1 + T(337018.0) * c - T(88297.0) * p1) +
                      m1 * (T(406293.0) * m1 - T(611976.0) * c + T(165153.0) * p1) +
                      c * (T(242723.0) * c - T(140251.0) * p1) + T(22658.0) * p1 * p1);
        const T b2 = (m2 * (T(6908.0) * m2 - T(51001.0) * m1 + T(67923.0) * c - T(38947.0) * p1 + T(8209.0) * p2) +
                      m1 * (T(104963.0) * m1 - T(299076.0) * c + T(179098.0) * p1 - T(38947.0) * p2) +
                      c * (T(231153.0) * c - T(299076.0) * p1 + T(67923.0) * p2) +
                      p1 * (T(104963.0) * p1 - T(51001.0) * p2) + T(6908.0) * p2 * p2);
        const T b3 = (m1 * (T(22658.0) * m1 - T(140251.0) * c + T(165153.0) * p1 - T(88297.0) * p2 + T(18079.0) * p3) +
                      c * (T(242723.0) * c - T(611976.0) * p1 + T(337018.0) * p2 - T(70237.0) * p3) +
                      p1 * (T(406293.0) * p1 - T(464976.0) * p2 + T(99213.0) * p3) +
                      p2 * (T(138563.0) * p2 - T(60871.0) * p3) + T(6908.0) * p3 * p3);
        const T b4 = (c * (T(107918.0) * c - T(649501.0) * p1 + T(758823.0) * p2 - T(411487.0) * p3 + T(86329.0) * p4) +
                      p1 * (T(1020563.0) * p1 - T(2462076.0) * p2 + T(1358458.0) * p3 - T(288007.0) * p4) +
                      p2 * (T(1521393.0) * p2 - T(1704396.0) * p3 + T(364863.0) * p4) +
                      p3 * (T(482963.0) * p3 - T(208501.0) * p4) + T(22658.0) * p4 * p4);
        const T a0 = Num<T>::div(T(1.0 / 126.0), sq_eps(b0));
        const T a1 = Num<T>::div(T(10.0 / 63.0), sq_eps(b1));
        const T a2 = Num<T>::div(T(10.0 / 21.0), sq_eps(b2));
        const T a3 = Num<T>::div(T(20.0 / 63.0), sq_eps(b3));
        const T a4 = Num<T>::div(T(5.0 / 126.0), sq_eps(b4));
        const T inv = Num<T>::rcp(a0 + a1 + a2 + a3 + a4);
        const T w0 = a0 * inv, w1 = a1 * inv, w2 = a2 * inv, w3 = a3 * inv, w4 = a4 * inv;
        return w0 * s0 + w1 * s1 + w2 * s2 + w3 * s3 + w4 * s4;
    }
    // interp_8: q = phim3 .. phip4
    // stencil with a zero centre value (no specialised form for this order)
    static __device__ __forceinline__ T weno_c0(const T *q) { return weno(q); }
    static __device__ __forceinline__ T centred(const T *q) {
        return (T(533. / 840.) * (q[3] + q[4]) - T(139.0 / 840.0) * (q[2] + q[5]) +
                T(29.0 / 840.0) * (q[1] + q[6]) - T(1.0 / 280.0) * (q[0] + q[7]));
    }
};

// Upwinded face flux through the face between stencil cells R-1 and R of
// win[0 .. 2R-1] (win[m] = phi at face-cell offset m-(R-1)).
//
// The reference evaluates both one-sided reconstructions and blends them as
// 0.5*((U+|U|)*phi_p + (U-|U|)*phi_m)*rho (scalar_advection.h:361).  For
// finite U and phi that equals U*phi_p*rho when U >= 0 and U*phi_m*rho when
// U < 0 to the last bit: (U+|U|) = 2U and (U-|U|) = 0 exactly, 2U*phi rounds
// to twice fl(U*phi), and the 0.5 scaling is exact.  So the kernels select the
// upwind stencil first and evaluate one reconstruction per face.
// Correctly rounded multiply and add that are never contracted into an FMA
// (the pragma drops the contraction permission from these operations), so an
// expression gives the same bits in every translation unit whatever its
// contraction flag.
__device__ __forceinline__ double mul_rn(double a, double b) {
#pragma clang fp contract(off)
    return a * b;
}
__device__ __forceinline__ double add_rn(double a, double b) {
#pragma clang fp contract(off)
    return a + b;
}
__device__ __forceinline__ float mul_rn(float a, float b) {
#pragma clang fp contract(off)
    return a * b;
}
__device__ __forceinline__ float add_rn(float a, float b) {
#pragma clang fp contract(off)
    return a + b;
}

// TimeStepping.update_third (TimeStepping.pyx:131-151), one point of stage RK:
// phi is the value, tn its tendency, cp the stage-0 copy.  Uncontracted, in
// the reference's operation order.
template <typename T, int RK>
__device__ __forceinline__ T rk_update(T phi, T tn, T cp, T dt) {
    const T x = add_rn(phi, mul_rn(tn, dt));
    if (RK == 0) return x;
    if (RK == 1) return add_rn(mul_rn(T(0.75), cp), mul_rn(T(0.25), x));
    return add_rn(mul_rn(T(1.0 / 3.0), cp), mul_rn(T(2.0 / 3.0), x));
}

template <int ORDER, typename T>
__device__ __forceinline__ T upwind_flux(const T *win, T U, T rho) {
    constexpr int R = Interp<ORDER, T>::R;
    T q[2 * R - 1];
    const bool pos = U >= T(0);
#pragma unroll
    for (int m = 0; m < 2 * R - 1; ++m) q[m] = pos ? win[m] : win[2 * R - 1 - m];
    const T ph = Interp<ORDER, T>::weno(q);
    return (U * ph) * rho;
}

// ---------------------------------------------------------------------------
// Error reporting shared by the entry points.
void pycles_set_error(const char *fmt, ...);

#define PYCLES_CHECK_HIP(expr)                                                          \
    do {                                                                                \
        hipError_t e__ = (expr);                                                        \
        if (e__ != hipSuccess) {                                                        \
            pycles_set_error("%s:%d %s: %s", __FILE__, __LINE__, #expr, hipGetErrorString(e__)); \
            return PYCLES_ERR_HIP;                                                      \
        }                                                                               \
    } while (0)

#define PYCLES_CHECK_LAUNCH()                                                           \
    do {                                                                                \
        hipError_t e__ = hipGetLastError();                                             \
        if (e__ != hipSuccess) {                                                        \
            pycles_set_error("%s:%d launch: %s", __FILE__, __LINE__, hipGetErrorString(e__)); \
            return PYCLES_ERR_HIP;                                                      \
        }                                                                               \
    } while (0)
