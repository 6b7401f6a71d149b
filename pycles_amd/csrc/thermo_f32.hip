// FP32 saturation-adjustment kernel (with and without the learned first
// guess).  Built with multiply-add contraction: the FP32 form is checked
// against the FP64 form by tolerance, not bit for bit.  See thermo.hip for the
// reference citations.
#include "thermo_common.h"

// ---------------------------------------------------------------------------
// FP32 storage and arithmetic; s and qt are deviations from FP64 level
// profiles; FP64 only in the per-level constants and the residual check.
// MLP selects the learned first guess (a separate instantiation, so the plain
// kernel carries none of the network's registers or LDS).  The guess is
// evaluated by whole wavefronts (one row of 64 columns at one level) when any
// of their cells is saturated; every lane stays in the loop for that reason.
// Dynamic LDS: the pv* table in FP64, the chunk's p0 and base profiles (FP64),
// the table in FP32, then the chunk's 1/exner, alpha0, 1/alpha0 and 1/p0
// (divisions by them are multiplications).  Every per-level value comes from
// LDS: a vector load of one would make each level wait for all the loads and
// stores before it (the memory counter completes in order).
// WTO: the w tendency is overwritten, not added to (no read of it).
// The field pointers are copied into SGPR pairs of their own (sreg).
struct NoSmem {};
struct MlpSmem {
    mlp::Weights w;
};

// Waves per SIMD: at 5 the plain kernel keeps its registers (at 6 its spill
// reloads, which wait for every outstanding load and store, cost more than
// the extra wave gains); the network's registers keep the learned kernel at 4
// (at 5 it spills and runs slower).
template <bool MLP, bool WTO>
__global__ __launch_bounds__(256, MLP ? 4 : 5) void thermo_sa_f32_kernel(Geo g, SAArgs a) {
    extern __shared__ double dyn[];
    const int nl = a.kchunk + 2;  // levels of a chunk with its two neighbours
    double *tab = dyn;
    double *p0l = tab + a.ntab;
    double *sbl = p0l + nl;
    double *qbl = sbl + nl;
    float *tabf = (float *)(qbl + nl);
    float *iexl = tabf + a.ntab;
    float *a0l = iexl + nl;
    float *ia0l = a0l + nl;
    float *ip0l = ia0l + nl;
    __shared__ unsigned long long cnt[4];
    __shared__ typename std::conditional<MLP, MlpSmem, NoSmem>::type msm_store;
    MlpSmem *msm = reinterpret_cast<MlpSmem *>(&msm_store);
    const int tid = threadIdx.y * blockDim.x + threadIdx.x;
    const int nthr = blockDim.x * blockDim.y;
    const Chunk ch = chunk_of(g, a.kchunk);
    for (int t = tid; t < a.ntab; t += nthr) {
        tab[t] = a.tab[t];
        tabf[t] = (float)a.tab[t];
    }
    for (int k = ch.kw + tid; k < ch.ke; k += nthr) {
        const int l = k - ch.kw;
        iexl[l] = (float)(1.0 / thermo::exner_c(a.p0[k]));
        p0l[l] = a.p0[k];
        sbl[l] = a.sbase ? a.sbase[k] : 0.0;
        qbl[l] = a.qbase ? a.qbase[k] : 0.0;
        a0l[l] = (float)a.alpha0_half[k];
        ia0l[l] = (float)(1.0 / a.alpha0_half[k]);
        ip0l[l] = (float)(1.0 / a.p0[k]);
    }
    if (tid < 4) cnt[tid] = 0;
    if (MLP) mlp::stage(a.mlp, &msm->w, tid, nthr);
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    unsigned ctr[4] = {0, 0, 0, 0};  // one thread's counts fit in 32 bits
    const bool active = i < g.nx && j < g.ny;
    const long long col = active ? (long long)j * g.nx + i : 0;
    float *s = sreg((float *)a.s);
    float *qtp = sreg((float *)a.qt);
    float *wt = sreg((float *)a.wt);
    float *To = sreg((float *)a.T), *qvo = sreg((float *)a.qv), *qlo = sreg((float *)a.ql);
    float *alo = sreg((float *)a.alpha), *bo = sreg((float *)a.buoy), *thro = sreg((float *)a.thr);
    float *bvfo = sreg((float *)a.bvf), *thlo = sreg((float *)a.thl);
    const thermo32::TableF ltf{tabf, (float)a.x_min, (float)a.dx, (float)a.dxi, a.ntab};
    const thermo::Table lt{tab, a.x_min, a.dx, a.dxi, a.ntab};
    const float dzi = (float)(1.0 / g.dx[2]);
    float Tw[3] = {0, 0, 0}, qtw[3] = {0, 0, 0}, qvw[3] = {0, 0, 0}, thw[3] = {0, 0, 0}, bprev = 0.0f;
    // s and qt of the next level are loaded one iteration ahead, by every
    // lane and at every level (the level clamped to the chunk; lanes outside
    // the domain read column 0): a load under a branch would make the wait
    // for it count from the branch's other path
    float s_nx = s[(long long)ch.kw * g.plane + col];
    float q_nx = qtp[(long long)ch.kw * g.plane + col];
    for (int k = ch.kw; k < ch.ke; ++k) {
        const int l = k - ch.kw;
        const bool own = active && k >= ch.k0 && k < ch.k1;
        const bool own_m = active && k - 1 >= ch.k0 && k - 1 < ch.k1;
        const long long o = (long long)k * g.plane + col;
        const double p0 = p0l[l];
        const float p0f = (float)p0;
        const double sb = sbl[l];
        const double qb = qbl[l];
        float sd = s_nx;
        float qd = q_nx;
        {
            const long long on = (long long)min(k + 1, ch.ke - 1) * g.plane + col;
            s_nx = s[on];
            q_nx = qtp[on];
        }
        double qt64 = qb + (double)qd;
        if (a.clip && !(qt64 >= 1e-11)) {
            qt64 = 1e-11;
            if (own) {
                // the stored deviation is stepped down until qb + dev, formed
                // as every reader forms it, is below 1e-11: a neighbouring
                // chunk that reads this level after the store clips it again
                // to exactly 1e-11, as one that reads it before does
                float dev = (float)(1e-11 - qb);
                for (int t = 0; t < 4 && !(qb + (double)dev < 1e-11); ++t) dev = nextafterf(dev, -INFINITY);
                qtp[o] = dev;
            }
        }
        const float qt = (float)qt64;
        const double s64 = sb + (double)sd;
        // s - (1-qt) sd~ - qt sv~ formed from the level profile and the deviations
        const double s_red = (sb - PC_SD_TILDE - qb * (PC_SV_TILDE - PC_SD_TILDE)) + (double)sd -
                             (qt64 - qb) * (PC_SV_TILDE - PC_SD_TILDE);
        const thermo32::SAStart st = thermo32::eos_sa_f32_start(ltf, p0f, s_red, qt);
        const bool sat = active && st.saturated;
        float T_guess = -1.0f;
        if (MLP) {
            if (__ballot(sat)) {
                float x0, x1, x2;
                mlp::normalise(&msm->w, s64, qt, p0f, &x0, &x1, &x2);
                T_guess = mlp::guess_wave(&msm->w, x0, x1, x2);
            }
        }
        float T = st.T_1, qv = qt, ql = 0.0f;
        int it = 0, fb = 0;
        if (sat)
            it = thermo32::eos_sa_f32_finish(ltf, lt, st, p0f, p0, s64, s_red, qt, qt64, T_guess, a.newton, &T,
                                             &qv, &ql, a.max_iter, &fb);
        if (sat && own) {
            ctr[0] += 1;
            ctr[1] += it;
            if (T_guess <= 0.0f && a.max_iter > 0 && it >= a.max_iter) ctr[2] += 1;
            ctr[3] += fb;
        }
        const float al = (float)PC_RD * T * ip0l[l] * (1.0f - qt + (float)PC_EPS_VI * qv);
        const float iex = iexl[l];
        float b = 0.0f;
        if (k < g.nzg - 1) b = (float)PC_G * (al - a0l[l]) * ia0l[l];
        float thr = 0.0f;
        if (k >= 1 && k < g.nzg - 1) thr = T * (1.0f - qt + (float)PC_EPS_VI * qv) * iex;
        Tw[0] = Tw[1]; Tw[1] = Tw[2]; Tw[2] = T;
        qtw[0] = qtw[1]; qtw[1] = qtw[2]; qtw[2] = qt;
        qvw[0] = qvw[1]; qvw[1] = qvw[2]; qvw[2] = qv;
        thw[0] = thw[1]; thw[1] = thw[2]; thw[2] = thr;
        if (own) {
            To[o] = T;
            qvo[o] = qv;
            qlo[o] = ql;
            alo[o] = al;
            if (bo && k < g.nzg - 1) bo[o] = b;
            if (thro && k >= 1 && k < g.nzg - 1) thro[o] = thr;
            if (thlo) {
                if (ql == 0.0f) thlo[o] = T * iex;
                else {
                    const float L = thermo32::latent_heat(T);
                    thlo[o] = T * iex *
                              thermo32::fexp(-L * thermo32::fdiv(thermo32::fdiv(ql, 1.0f - qt), T * (float)PC_CPD));
                }
            }
        }
        if (own_m && k >= 2 && k - 1 < g.nzg - 3) {
            const float w0 = WTO ? 0.0f : wt[o - g.plane];
            wt[o - g.plane] = w0 + 0.5f * (bprev + b);
        }
        bprev = b;
        const int m = k - 1;
        if (bvfo && own_m && m >= 2 && m < g.nzg - 2) {
            float bv;
            if (qvw[1] < qtw[1]) {
                const float Tk = Tw[1];
                const float Lv = thermo32::latent_heat(Tk);
                const float pvs = thermo32::lookup(ltf, Tk);
                const float p0m = (float)p0l[l - 1];
                using thermo32::fdiv;
                using thermo32::frcp;
                const float rsl = fdiv((float)PC_EPS_V * pvs, p0m - pvs);
                const float cpm = (1.0f - qtw[1]) * (float)PC_CPD + qtw[1] * (float)PC_CPV;
                const float iqd = frcp(1.0f - qtw[1]);
                const float iT = frcp(Tk);
                const float lr = 1.0f + Lv * rsl * iT * (float)(1.0 / PC_RD);
                const float gamma_w =
                    fdiv((float)(PC_G / PC_CPD) * iqd * lr,
                         cpm * (float)(1.0 / PC_CPD) +
                             Lv * Lv * ((float)PC_EPS_V + rsl) * rsl * iT * iT * (float)(1.0 / (PC_CPD * PC_RD)));
                const float dTdz = 0.5f * (Tw[2] - Tw[0]) * dzi;
                const float dqtdz = 0.5f * (qtw[2] - qtw[0]) * dzi;
                bv = (float)PC_G * iT * (dTdz + gamma_w) * lr - dqtdz * iqd;
            } else {
                bv = (float)PC_G * thermo32::frcp(thw[1]) * 0.5f * (thw[2] - thw[0]) * dzi;
            }
            bvfo[o - g.plane] = bv;
        }
    }
    const unsigned long long ctr64[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    flush_counters(cnt, ctr64, a.counters, tid);
}

// Dynamic LDS of the FP32 kernel: table (FP64 + FP32), three FP64 and four
// FP32 values per level of the chunk.
static size_t f32_smem_bytes(const SAArgs &a) {
    return (size_t)a.ntab * (sizeof(double) + sizeof(float)) +
           (size_t)(a.kchunk + 2) * (3 * sizeof(double) + 4 * sizeof(float));
}

template <bool MLP>
static void launch_f32(const Geo &g, const SAArgs &a, dim3 grd, dim3 blk, hipStream_t st) {
    if (a.wt_over)
        hipLaunchKernelGGL((thermo_sa_f32_kernel<MLP, true>), grd, blk, f32_smem_bytes(a), st, g, a);
    else
        hipLaunchKernelGGL((thermo_sa_f32_kernel<MLP, false>), grd, blk, f32_smem_bytes(a), st, g, a);
}

int thermo_sa_f32_launch(const Geo &g, const SAArgs &a, dim3 grd, dim3 blk, hipStream_t st) {
    if (a.mlp)
        launch_f32<true>(g, a, grd, blk, st);
    else
        launch_f32<false>(g, a, grd, blk, st);
    PYCLES_CHECK_LAUNCH();
    return PYCLES_OK;
}
