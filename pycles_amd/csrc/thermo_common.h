// Shared by the saturation-adjustment kernels (thermo.hip, thermo_f32.hip).
#pragma once
#include "common.h"
#include "mlp_infer.h"
#include "thermo_f32.h"
#include "thermo_funcs.h"

#define LOOKUP_MAX 2048
#define LEVELS_MAX 2048

struct SAArgs {
    const double *p0, *alpha0_half;
    const double *tab;
    int ntab;
    double x_min, dx, dxi;
    void *s, *qt, *wt, *T, *qv, *ql, *alpha, *buoy, *thr, *bvf, *thl;
    const double *sbase, *qbase;
    int clip, max_iter, lh, wt_over;
    const mlp::Weights *mlp;
    int newton;
    unsigned long long *counters;
    int kchunk;
    double sa_tol;  // the FP64 secant's stopping test (the reference: 1e-3 K)
};

// Block-level reduction of the four counters into out.
__device__ __forceinline__ void flush_counters(unsigned long long *cnt, const unsigned long long v[4],
                                               unsigned long long *out, int tid) {
    if (!out) return;
    for (int c = 0; c < 4; ++c)
        if (v[c]) atomicAdd(&cnt[c], v[c]);
    __syncthreads();
    if (tid < 4 && cnt[tid]) atomicAdd(&out[tid], cnt[tid]);
}

// A copy of a uniform pointer in an SGPR pair of its own, opaque to the
// compiler (a register move).  The kernel arguments arrive in wide scalar
// loads; the thermodynamics kernels need more SGPRs than they have, and once
// such a tuple is spilled (to VGPR lanes) every use of one of its pointers
// restores the whole tuple, one v_readlane per register, in every store
// block of the level loop.
template <typename T> __device__ __forceinline__ T *sreg(T *p) {
    T *q;
    asm volatile("s_mov_b64 %0, %1" : "=s"(q) : "s"(p));
    return q;
}

// z chunk of this block: owned [k0, k1), marched [kw, ke)
struct Chunk {
    int k0, k1, kw, ke;
};
__device__ __forceinline__ Chunk chunk_of(const Geo &g, int kchunk) {
    Chunk c;
    c.k0 = blockIdx.z * kchunk;
    c.k1 = min(c.k0 + kchunk, g.nzg);
    c.kw = max(0, c.k0 - 1);
    c.ke = min(c.k1 + 1, g.nzg);
    return c;
}

