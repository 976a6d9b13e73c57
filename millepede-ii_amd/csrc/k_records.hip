// Record-loop kernel: one 64-lane wavefront per Mille record.
//
// Restates loopbf (src/pede.f90:3387-4442) for records whose local fit takes the full
// (non-banded) path; records with bordered-band structure go to k_records_band.  Residuals
// after the current global parameters, local normal equations,
// symmetric Gauss-Jordan inversion with diagonal pivoting (sqminv, src/mpnum.f90:97-213),
// chi^2 / down-weighting / reject logic, then the "Millepede" update of the global system:
//   globalVector += dw * w * r * g                       (src/pede.f90:4184-4186)
//   globalMatrix += dw * w * g g^T  (per equation)       (src/pede.f90:4188-4201, 3336-3356)
//   globalMatrix -= G Gamma^-1 G^T  (Schur complement)   (dbavat, src/mpnum.f90:1389-1447)
// The per-element arithmetic of the local fit follows the reference expression by expression
// (same operand order), so local-fit decisions (rank, chi^2 cuts) match pede; sums across
// equations that pede forms sequentially are formed sequentially here as well, except the
// chi^2 and weight sums, which are wavefront reductions.  Matrix contributions are scattered
// with FP64 atomics, either into pede's packed full matrix or into the group-block layout.
#include "mpgpu_device.h"
#include <float.h>
#include <algorithm>

namespace mpgpu {

#define WAVE 64

__device__ __forceinline__ double wave_sum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
    return v;
}
__device__ __forceinline__ int wave_max(int v)
{
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, WAVE));
    return v;
}

template <typename T> __device__ __forceinline__ double wval(const DevRecords &R, int64_t w);
template <> __device__ __forceinline__ double wval<float>(const DevRecords &R, int64_t w) { return (double)R.vf[w]; }
template <> __device__ __forceinline__ double wval<double>(const DevRecords &R, int64_t w) { return R.vd[w]; }

// packed index helpers (0-based storage of 1-based (i,j), i>=j)
__device__ __forceinline__ int pk(int i, int j) { return (i * i - i) / 2 + j - 1; }
__device__ __forceinline__ int64_t pk64(int64_t i, int64_t j) { return (i * i - i) / 2 + j - 1; }

// g g^T part of element (a, b) of a record's update: dw * w_i * g_ia * g_ib summed over the
// equations i holding both variables (from the per-variable equation lists, or, for records
// that overflow them, by scanning the equations as pede's per-equation update does)
template <typename T>
__device__ __forceinline__ double dc1(int a, int b, bool perq, int QE, const int32_t *qn, const int32_t *qe,
                                      const double *qg, const double *lw, double dw1, const int32_t *les,
                                      const uint32_t *len, const int32_t *li, const T *lv, int neq, const int32_t *gpos)
{
    double c = 0.0;
    if (!perq) {
        const int na = qn[a], nb = qn[b];
        for (int x = 0; x < na; x++) {
            const int i = qe[a * QE + x];
            for (int y = 0; y < nb; y++)
                if (qe[b * QE + y] == i) c = c + dw1 * lw[i] * qg[a * QE + x] * qg[b * QE + y];
        }
        return c;
    }
    for (int i = 0; i < neq; i++) {
        const int sl = les[i];
        const uint32_t en = len[i];
        const int sg = sl + (int)(en & 0xffff) + 1;
        const int ng = en >> 16;
        for (int j = 1; j <= ng; j++) {
            if (gpos[li[sg + j]] != a) continue;
            const double gj = (double)lv[sg + j];
            if (gj == 0.0) continue;
            for (int k = 1; k <= ng; k++) {
                if (gpos[li[sg + k]] != b) continue;
                if (a == b && k > j) continue;           // each (j, k <= j) word pair once
                c = c + dw1 * lw[i] * gj * (double)lv[sg + k];
            }
        }
    }
    return c;
}

// a rejected record contributes nothing: its slot of the update buffer is cleared
__device__ __forceinline__ void zero_update(const DevRecords &R, const DevModel &M, const PassOpts &o, int64_t r, int lane)
{
    if (o.icalcm != 1 || !M.sparse || R.cbuf == nullptr) return;
    const int64_t off = R.rec_cb[r];
    if (off < 0) return;
    const int n = R.rec_cn[r];
    for (int t = lane; t < n; t += WAVE) R.cbuf[off + t] = 0.0;
}

template <typename T>
__global__ __launch_bounds__(WAVE) void k_records(DevRecords R, DevModel M, PassOpts o, RecOut out, LdsPlan P,
                                                  int64_t rec_begin, int64_t rec_end)
{
    extern __shared__ double lds[];
    const int lane = threadIdx.x;
    const double chuber = 1.345, cauchy = 2.3849;
    const int QE = P.qe;                     // equations per record variable on the merged update path

    // ---- carve LDS ----
    double *gp = lds;                        // [max_gl] parameter value per distinct label
    double *rm = gp + P.max_gl;              // [max_eq] corrected measurement / residual
    double *re = rm + P.max_eq;              // [max_eq] sigma
    double *lc = re + P.max_eq;              // [max_eq] local-fit prediction
    double *lw = lc + P.max_eq;              // [max_eq] weight of the last local-fit pass
    const int npk_max = (P.nalc_max * P.nalc_max + P.nalc_max) / 2;
    double *cl = lw + P.max_eq;              // [npk_max] local matrix -> inverse
    double *bl = cl + npk_max;               // [nalc_max] local vector -> solution
    double *dg = bl + P.nalc_max;            // [nalc_max] |diag| for pivot test
    double *cs = dg + P.nalc_max;            // [nalc_max] scaled pivot column
    double *co = cs + P.nalc_max;            // [nalc_max] original pivot column
    double *G = co + P.nalc_max;             // [max_var * nalc_max]
    double *Gh = G + (size_t)P.max_var * P.nalc_max; // [max_var * nalc_max] (full storage only)
    double *qg = Gh + (size_t)P.gh * P.max_var * P.nalc_max; // [max_var*QE] derivative in those equations
    double *stage = qg + (size_t)P.max_var * QE;   // [stage] values of one step (coalesced stores)
    T *lv = (T *)(stage + P.stage);                // [max_w] the record's word values
    double *ghs = stage + P.stage + ((size_t)P.max_w * sizeof(T) + 7) / 8; // [max_grp * nalc_max] rows of G Gamma^-1
    const int d1n = min(P.max_grp, 16);
    double *d1 = ghs + (size_t)P.max_grp * P.nalc_max;  // [min(max_grp,16)^2] g g^T of a diagonal block
    unsigned long long *segm = (unsigned long long *)(d1 + (size_t)d1n * d1n); // [max_var]
    int32_t *gv = (int32_t *)(segm + P.max_var); // [max_gl] var index
    int32_t *gpos = gv + P.max_gl;           // [max_gl] position among the record's variables
    int32_t *gvl = gpos + P.max_gl;          // [max_var] sorted variable list
    int32_t *segof = gvl + P.max_var;        // [max_var] segment of each record variable
    int32_t *segst = segof + P.max_var;      // [max_var+1] first record variable of each segment
    int32_t *qn = segst + P.max_var + 1;     // [max_var] equations of each record variable
    int32_t *qe = qn + P.max_var;            // [max_var*QE] their indices (ascending)
    int32_t *nxt = qe + P.max_var * QE;      // [nalc_max+1] pivot list
    int32_t *sh = nxt + P.nalc_max + 1;      // [8] scalars
    int32_t *li = sh + 8;                    // [max_w] the record's word indices (local / global position)
    int32_t *les = li + P.max_w;             // [max_eq] measurement word of each equation (record-local)
    uint32_t *len = (uint32_t *)(les + P.max_eq); // [max_eq] local / global word counts
    auto lwv = [&](int x) { return (double)lv[x]; };

    for (int64_t ri = rec_begin + blockIdx.x; ri < rec_end; ri += gridDim.x) {
        const int64_t r = R.order ? R.order[ri] : ri;
        if (R.band && R.band[r] != 0) continue;   // bordered-band local fit: k_records_band
        const int64_t e0 = R.rec_eq[r];
        const int neq = (int)(R.rec_eq[r + 1] - e0);
        const int64_t g0 = R.rec_gl[r];
        const int ngl = (int)(R.rec_gl[r + 1] - g0);
        const double dw1 = R.rec_w ? (double)R.rec_w[r] : 1.0;
        const double dw2 = sqrt(dw1);
        const int64_t w0 = R.rec_word[r];
        const int nw = (int)(R.rec_word[r + 1] - w0);
        if (neq > P.max_eq || ngl > P.max_gl || nw > P.max_w) {
            if (lane == 0) { out.status[r] = 9; out.chi2[r] = 0.0; out.ndf[r] = 0; }
            zero_update(R, M, o, r, lane);
            continue;
        }
        // ---- the record's words and equation table into LDS (coalesced, once) ----
        for (int t = lane; t < nw; t += WAVE) {
            lv[t] = R.vf ? (T)R.vf[w0 + t] : (T)R.vd[w0 + t];
            li[t] = R.widx[w0 + t];
        }
        for (int i = lane; i < neq; i += WAVE) {
            les[i] = (int)(R.eq_sl[e0 + i] - w0);
            len[i] = R.eq_n[e0 + i];
        }
        // ---- distinct labels: parameter value, variable index, position among variables ----
        int nalg = 0;
        for (int base = 0; base < ngl; base += WAVE) {
            int k = base + lane;
            int v = 0;
            if (k < ngl) {
                int t = R.gl_tix[g0 + k];
                gp[k] = M.par[t];
                v = M.var[t];
                gv[k] = v;
            }
            bool isv = (k < ngl) && v > 0;
            unsigned long long m = __ballot(isv);
            int pos = nalg + __popcll(m & ((1ull << lane) - 1ull));
            if (k < ngl) gpos[k] = isv ? pos : -1;
            if (isv && pos < P.max_var) gvl[pos] = v;
            nalg += __popcll(m);
        }
        if (nalg > P.max_var) {
            if (lane == 0) { out.status[r] = 9; out.chi2[r] = 0.0; out.ndf[r] = 0; }
            zero_update(R, M, o, r, lane);
            continue;
        }
        __syncthreads();
        // ---- first loop: subtract global contribution, count local parameters (3655-3693) ----
        int nalc = 0;
        for (int i = lane; i < neq; i += WAVE) {
            const int sl = les[i];
            const uint32_t en = len[i];
            int nl = en & 0xffff, ng = en >> 16;
            const int sg = sl + nl + 1;
            double rmeas = lwv(sl);
            for (int j = 1; j <= ng; j++) rmeas = rmeas - lwv(sg + j) * gp[li[sg + j]];
            rm[i] = rmeas;
            re[i] = lwv(sg);
            lc[i] = 0.0;
            for (int j = 1; j <= nl; j++) nalc = max(nalc, li[sl + j]);
        }
        nalc = wave_max(nalc);
        if (nalc <= 0) {
            if (lane == 0) { out.status[r] = 7; out.chi2[r] = 0.0; out.ndf[r] = 0; }
            zero_update(R, M, o, r, lane);
            __syncthreads();
            continue;
        }
        if (nalc > P.nalc_max) {
            if (lane == 0) { out.status[r] = 9; out.chi2[r] = 0.0; out.ndf[r] = 0; }
            zero_update(R, M, o, r, lane);
            __syncthreads();
            continue;
        }
        const int npk = (nalc * nalc + nalc) / 2;
        // ---- outlier-suppression iterations (3759-4094) ----
        const int nter = (o.nloopn != 1 && o.lhuber != 0) ? o.lhuber : 1;
        int status = 0, nrank = 0, ndf = 0;
        double summ = 0.0;
        float resing = 0.0f;
        for (int iter = 1; iter <= nter; iter++) {
            // weights
            for (int i = lane; i < neq; i += WAVE) {
                double rerr = re[i];
                double wght = 1.0 / (rerr * rerr);
                double resid = rm[i] - lc[i];
                if (o.nloopn != 1 && iter != 1 && o.lhuber != 0) {
                    if (iter <= 3) {
                        if (fabs(resid) > chuber * rerr) wght = wght * chuber * rerr / fabs(resid);
                    } else {
                        double t = resid / rerr / cauchy;
                        wght = wght / (1.0 + t * t);
                    }
                }
                lw[i] = wght;
            }
            __syncthreads();
            // local normal equations, element-parallel, equations in order (3781-3836)
            for (int t = lane; t < npk + nalc; t += WAVE) {
                int J, K;
                if (t < npk) {
                    J = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5) + 1;
                    while ((J * J - J) / 2 > t) J--;
                    while ((J * J + J) / 2 <= t) J++;
                    K = t - (J * J - J) / 2 + 1;
                } else {
                    J = t - npk + 1;
                    K = 0;
                }
                double acc = 0.0;
                for (int i = 0; i < neq; i++) {
                    const int sl = les[i];
                    int nl = len[i] & 0xffff;
                    double dj = 0.0, dk = 0.0;
                    bool hj = false, hk = false;
                    for (int j = 1; j <= nl; j++) {
                        int ij = li[sl + j];
                        if (ij == J) { dj = lwv(sl + j); hj = true; }
                        if (K > 0 && ij == K) { dk = lwv(sl + j); hk = true; }
                    }
                    if (K == 0) {
                        if (hj) acc = acc + lw[i] * rm[i] * dj;
                    } else if (hj && hk) {
                        acc = acc + lw[i] * dj * dk;
                    }
                }
                if (K == 0) bl[J - 1] = acc; else cl[t] = acc;
            }
            __syncthreads();
            // ---- sqminv (src/mpnum.f90:97-213), element-parallel Gauss-Jordan ----
            const double eps = 16.0 * DBL_EPSILON;
            for (int i = lane; i < nalc; i += WAVE) {
                dg[i] = fabs(cl[pk(i + 1, i + 1)]);
                nxt[i + 1] = i + 2;
            }
            __syncthreads();
            if (lane == 0) { nxt[nalc] = -1; sh[1] = 1; /* next0 */ sh[2] = 0; /* nrank */ }
            __syncthreads();
            for (int step = 1; step <= nalc; step++) {
                if (lane == 0) {
                    int k = 0, l = 0, last = 0;
                    double vkk = 0.0;
                    int j = sh[1];
                    while (j > 0) {
                        double vjj = cl[pk(j, j)];
                        double m = fmax(fabs(vkk), eps * dg[j - 1]);
                        if (fabs(vjj) > m) { vkk = vjj; k = j; l = last; }
                        last = j;
                        j = nxt[last];
                    }
                    if (k != 0) {
                        if (l == 0) sh[1] = nxt[k]; else nxt[l] = nxt[k];
                        nxt[k] = 0;
                        sh[2] = sh[2] + 1;
                    }
                    sh[3] = k;
                }
                __syncthreads();
                const int k = sh[3];
                if (k == 0) {
                    for (int t = lane; t < npk; t += WAVE) {
                        int J = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5) + 1;
                        while ((J * J - J) / 2 > t) J--;
                        while ((J * J + J) / 2 <= t) J++;
                        int K = t - (J * J - J) / 2 + 1;
                        if (nxt[J] != 0 && nxt[K] != 0) cl[t] = 0.0;
                    }
                    for (int i = lane; i < nalc; i += WAVE) if (nxt[i + 1] != 0) bl[i] = 0.0;
                    __syncthreads();
                    break;
                }
                const int kk = pk(k, k);
                const double vkk = 1.0 / cl[kk];
                for (int j = lane + 1; j <= nalc; j += WAVE) {
                    if (j != k) {
                        double v = cl[j < k ? pk(k, j) : pk(j, k)];
                        co[j - 1] = v;
                        cs[j - 1] = vkk * v;
                    }
                }
                __syncthreads();
                const double bk = bl[k - 1] * vkk;
                for (int t = lane; t < npk; t += WAVE) {
                    int J = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5) + 1;
                    while ((J * J - J) / 2 > t) J--;
                    while ((J * J + J) / 2 <= t) J++;
                    int K = t - (J * J - J) / 2 + 1;
                    if (J == k && K == k) cl[t] = -vkk;
                    else if (J == k) cl[t] = cs[K - 1];
                    else if (K == k) cl[t] = cs[J - 1];
                    else cl[t] = cl[t] - cs[K - 1] * co[J - 1];
                }
                for (int j = lane + 1; j <= nalc; j += WAVE) {
                    if (j == k) bl[j - 1] = bk;
                    else bl[j - 1] = bl[j - 1] - bk * co[j - 1];
                }
                __syncthreads();
            }
            for (int t = lane; t < npk; t += WAVE) cl[t] = -cl[t];
            nrank = sh[2];
            __syncthreads();
            // NaN check (3903-3906)
            int nan = 0;
            for (int i = lane; i < nalc; i += WAVE) {
                double b = bl[i];
                if (!(b <= 0.0) && !(b > 0.0)) nan++;
            }
            nan = wave_max(nan);
            // ---- third loop: residuals, chi^2 (3919-4035) ----
            double s_chi = 0.0, s_wt = 0.0;
            for (int i = lane; i < neq; i += WAVE) {
                const int sl = les[i];
                int nl = len[i] & 0xffff;
                double rerr = re[i];
                double wght = 1.0 / (rerr * rerr);
                double rmloc = 0.0;
                for (int j = 1; j <= nl; j++) rmloc = rmloc + lwv(sl + j) * bl[li[sl + j] - 1];
                lc[i] = rmloc;
                double rmeas = rm[i] - rmloc;
                double dchi2 = wght * rmeas * rmeas;
                double resid = rmeas;
                if (o.nloopn != 1 && iter != 1 && o.lhuber != 0) {
                    if (iter <= 3) {
                        if (fabs(resid) > chuber * rerr) {
                            wght = wght * chuber * rerr / fabs(resid);
                            dchi2 = 2.0 * chuber * (fabs(resid) / rerr - 0.5 * chuber);
                        }
                    } else {
                        double t = resid / rerr / cauchy;
                        wght = wght / (1.0 + t * t);
                        dchi2 = log(1.0 + t * t) * cauchy * cauchy;
                    }
                }
                double down = 1.0 / sqrt(wght);
                s_wt += rerr / down;
                s_chi += dchi2;
                if (iter == nter) rm[i] = rmeas; // remaining residual
            }
            summ = wave_sum(s_chi);
            double suwt = wave_sum(s_wt);
            __syncthreads();
            ndf = neq - nrank;
            resing = ((float)neq - (float)suwt) / (float)neq;
            if (nan > 0) { status = 1; break; }
            if (nrank != nalc) { status = 2; break; }
            if (ndf <= 0) { status = 4; break; }
        }
        double chi_add = 0.0;
        int ndf_add = 0;
        if (status == 0) {
            // chi^2 cuts with REAL(4) thresholds (4106-4131)
            float chichi;
            {
                // chindl(3, ndf) (src/mpnum.f90:2078-2115)
                const float tab3[30] = {9.0000f, 5.9146f, 4.7184f, 4.0628f, 3.6410f, 3.3436f, 3.1209f, 2.9468f,
                                        2.8063f, 2.6902f, 2.5922f, 2.5082f, 2.4352f, 2.3711f, 2.3143f, 2.2635f,
                                        2.2178f, 2.1764f, 2.1386f, 2.1040f, 2.0722f, 2.0428f, 2.0155f, 1.9901f,
                                        1.9665f, 1.9443f, 1.9235f, 1.9040f, 1.8855f, 1.8681f};
                float c;
                if (ndf < 1) c = 0.0f;
                else if (ndf <= 30) c = tab3[ndf - 1];
                else {
                    float t = 2.782170f + sqrtf((float)(ndf + ndf - 1));
                    c = t * t / (float)(ndf + ndf);
                }
                chichi = c * (float)ndf;
            }
            if (o.chicut >= 0.0f) {
                if (summ > (double)(o.chhuge * chichi)) status = 5;
                else if (o.chicut > 0.0f) {
                    float chlimt = o.chicut * chichi;
                    if (summ > (double)chlimt) { status = 6; chi_add = (double)chlimt; ndf_add = ndf; }
                }
            }
            if (status == 0 && o.lhuber > 1 && o.dwcut != 0.0f && resing > o.dwcut) {
                status = 6;
                chi_add = summ;
                ndf_add = ndf;
            }
        }
        if (status != 0) {
            if (lane == 0) { out.status[r] = (int8_t)status; out.chi2[r] = chi_add * dw1; out.ndf[r] = ndf_add; }
            zero_update(R, M, o, r, lane);
            __syncthreads();
            continue;
        }
        // ---- fourth loop: vector, counters, g g^T, G (4162-4229) ----
        // Every record variable collects the equations it occurs in (up to QE, ascending): the
        // vector and the counters get one update per variable.  Full storage: g g^T per
        // equation and the Schur complement per element with FP64 atomics.  Group-block
        // storage: the record's complete update (g g^T of the common equations minus the Schur
        // complement) is written block by block into the update buffer, to be summed per matrix
        // block by k_block_gather.  A variable in more than QE equations, or twice in one
        // equation, makes the record take the per-equation path for the vector (and, with
        // group-block storage, an equation scan for g g^T).
        const bool twophase = M.sparse != 0;   // the host always provides the update buffer then
        if (o.icalcm == 1)
            for (int t = lane; t < nalg * nalc; t += WAVE) G[t] = 0.0;
        for (int t = lane; t < nalg; t += WAVE) qn[t] = 0;
        if (lane == 0) sh[5] = 0;
        if (o.icalcm == 1 && twophase) {
            // segments: runs of record variables in the same parameter group
            int nseg = 0;
            for (int base = 0; base < nalg; base += WAVE) {
                int a = base + lane;
                int g = a < nalg ? M.group_of[gvl[a]] : -1;
                int gprev = (a > 0 && a < nalg) ? M.group_of[gvl[a - 1]] : -2;
                bool st = (a < nalg) && (a == 0 || g != gprev);
                unsigned long long m = __ballot(st);
                int s = nseg + __popcll(m & ((2ull << lane) - 1ull)) - 1;
                if (a < nalg && st) segst[s] = a;
                nseg += __popcll(m);
            }
            if (lane == 0) { segst[nseg] = nalg; sh[6] = nseg; }
        }
        __syncthreads();
        double s_chi = 0.0;
        const bool down = (o.nloopn != 1 && o.lhuber != 0);
        for (int i = lane; i < neq; i += WAVE) {
            const int sl = les[i];
            const uint32_t en = len[i];
            int nl = en & 0xffff, ng = en >> 16;
            const int sg = sl + nl + 1;
            double rmeas = rm[i], rerr = re[i];
            double wght = 1.0 / (rerr * rerr);
            double dchi2 = wght * rmeas * rmeas;
            if (down) {
                double resid = fabs(rmeas);
                if (nter <= 3) {
                    if (resid > chuber * rerr) dchi2 = 2.0 * chuber * (resid / rerr - 0.5 * chuber);
                } else {
                    double t = resid / rerr / cauchy;
                    dchi2 = log(1.0 + t * t) * cauchy * cauchy;
                }
                wght = lw[i];
            }
            s_chi += dchi2;
            lw[i] = wght; // weight used by the global update
            for (int j = 1; j <= ng; j++) {
                double gj = lwv(sg + j);
                int pj = li[sg + j];
                int vj = gv[pj];
                if (gj == 0.0 || vj <= 0) continue;
                int aj = gpos[pj];
                int slot = atomicAdd(&qn[aj], 1);
                if (slot < QE) { qe[aj * QE + slot] = i; qg[aj * QE + slot] = gj; }
                if (o.icalcm == 1 && !M.sparse) {
                    // full storage: g g^T of this equation
                    for (int k = 1; k <= j; k++) {
                        int pk_ = li[sg + k];
                        int vk = gv[pk_];
                        if (vk > 0) {
                            double val = dw1 * wght * gj * lwv(sg + k);
                            if (vj >= vk) atomicAdd(&M.mat[pk64(vj, vk)], val);
                            else atomicAdd(&M.mat[pk64(vk, vj)], val);
                        }
                    }
                }
            }
        }
        summ = wave_sum(s_chi);
        __syncthreads();
        for (int a = lane; a < nalg; a += WAVE) {
            const int n = qn[a];
            if (n > QE) { sh[5] = 1; continue; }
            for (int x = 1; x < n; x++)
                for (int y = x; y > 0 && qe[a * QE + y - 1] >= qe[a * QE + y]; y--) {
                    if (qe[a * QE + y - 1] == qe[a * QE + y]) sh[5] = 1;
                    int te = qe[a * QE + y]; qe[a * QE + y] = qe[a * QE + y - 1]; qe[a * QE + y - 1] = te;
                    double tg = qg[a * QE + y]; qg[a * QE + y] = qg[a * QE + y - 1]; qg[a * QE + y - 1] = tg;
                }
        }
        __syncthreads();
        const bool perq = sh[5] != 0;
        if (!perq) {
            for (int a = lane; a < nalg; a += WAVE) {
                const int n = qn[a];
                double v = 0.0;
                for (int x = 0; x < n; x++) {
                    const int i = qe[a * QE + x];
                    v = v + dw1 * lw[i] * rm[i] * qg[a * QE + x];
                }
                atomicAdd(&M.vec[gvl[a] - 1], v);
                atomicAdd(&M.cnt[gvl[a] - 1], (unsigned long long)n);
            }
        } else {
            for (int i = lane; i < neq; i += WAVE) {
                const int sl = les[i];
                const uint32_t en = len[i];
                int nl = en & 0xffff, ng = en >> 16;
                const int sg = sl + nl + 1;
                for (int j = 1; j <= ng; j++) {
                    double gj = lwv(sg + j);
                    int vj = gv[li[sg + j]];
                    if (gj == 0.0 || vj <= 0) continue;
                    atomicAdd(&M.vec[vj - 1], dw1 * lw[i] * rm[i] * gj);
                    atomicAdd(&M.cnt[vj - 1], 1ull);
                }
            }
        }
        if (o.icalcm == 1) {
            // G = sum over equations of sqrt(dw) w g l^T, equations in order (4206-4219)
            for (int i = 0; i < neq; i++) {
                const int sl = les[i];
                const uint32_t en = len[i];
                int nl = en & 0xffff, ng = en >> 16;
                const int sg = sl + nl + 1;
                double wght = lw[i];
                for (int t = lane; t < ng * nl; t += WAVE) {
                    int j = t / nl + 1, k = t % nl + 1;
                    int pj = li[sg + j];
                    int a = gpos[pj];
                    if (a >= 0) {
                        int ik = li[sl + k];
                        G[a * nalc + ik - 1] = G[a * nalc + ik - 1] + dw2 * wght * lwv(sg + j) * lwv(sl + k);
                    }
                }
                __syncthreads();
            }
            // Gh = G Gamma^-1 (row-wise, l ascending as dbavat); group-block storage forms the
            // rows of Gh in registers where they are used
            if (!twophase) {
                for (int a = lane; a < nalg; a += WAVE) {
                    for (int k = 1; k <= nalc; k++) {
                        double c = 0.0;
                        for (int l = 1; l <= nalc; l++) c = c + G[a * nalc + l - 1] * cl[l >= k ? pk(l, k) : pk(k, l)];
                        Gh[a * nalc + k - 1] = c;
                    }
                }
            }
            __syncthreads();
            if (!twophase) {
                // full storage: element (a>=b) -= sum_k Gh[a,k] G[b,k]; packed (a, b) enumeration,
                // lane starts at element `lane`, then steps by 64
                const int ngg = (nalg * nalg + nalg) / 2;
                int a = (int)((sqrt(8.0 * lane + 1.0) - 1.0) * 0.5);
                while ((a * a + a) / 2 > lane) a--;
                while (((a + 1) * (a + 1) + (a + 1)) / 2 <= lane) a++;
                int b = lane - (a * a + a) / 2;
                for (int t = lane; t < ngg; t += WAVE) {
                    double c = 0.0;
                    for (int k = 0; k < nalc; k++) c = c + Gh[a * nalc + k] * G[b * nalc + k];
                    atomicAdd(&M.mat[pk64(gvl[a], gvl[b])], -c);
                    b += WAVE;
                    while (b > a) { b -= a + 1; a++; }
                }
            } else {
                // Group-block storage.  Layout of a record's update: for every segment sa (in
                // order) the columns b = 0 .. end of sa (all record variables up to and including
                // segment sa), each column holding the na values of the segment's rows
                // (value (ra, b) at b * na + ra); the diagonal block is kept square, its upper
                // part is zero.  The block of segment pair (sa, sb) is thus contiguous
                // (columns of sb), which the block gather reads.  Lanes take columns, keep the
                // segment's rows of G Gamma^-1 in registers, and stage 64 columns in LDS so that
                // the buffer is written in contiguous runs.
                const int nseg = sh[6];
                double *dst = R.cbuf + R.rec_cb[r];
                // equations touching each segment (bit i for equation i < 64; bit 63 also stands
                // for all equations >= 63): pairs without a common equation skip g g^T
                for (int q = lane; q < nseg; q += WAVE) {
                    unsigned long long m = 0ull;
                    for (int a = segst[q]; a < segst[q + 1]; a++) {
                        if (perq) { m = ~0ull; break; }
                        for (int x = 0; x < qn[a]; x++) m |= 1ull << min(qe[a * QE + x], 63);
                    }
                    segm[q] = m;
                }
                // segment of every record variable
                for (int q = 0; q < nseg; q++)
                    for (int a = segst[q] + lane; a < segst[q + 1]; a += WAVE) segof[a] = q;
                __syncthreads();
                int64_t roff = 0;                              // start of segment sa's columns
                for (int sa = 0; sa < nseg; sa++) {
                    const int a0 = segst[sa], na = segst[sa + 1] - a0;
                    const int ncolumns = a0 + na;
                    // rows of G Gamma^-1 of the segment (l ascending, as dbavat)
                    for (int t = lane; t < na * nalc; t += WAVE) {
                        const int ra = t / nalc, k = t - ra * nalc + 1;
                        double c = 0.0;
                        for (int l = 1; l <= nalc; l++) c = c + G[(a0 + ra) * nalc + l - 1] * cl[l >= k ? pk(l, k) : pk(k, l)];
                        ghs[t] = c;
                    }
                    // g g^T of the diagonal block, one element per lane
                    const bool dcommon = segm[sa] != 0ull;
                    const bool use_d1 = na <= 16;           // d1 holds blocks of up to 16 x 16
                    for (int t = lane; use_d1 && t < na * na; t += WAVE) {
                        const int ra = t / na, rb = t - ra * na;
                        d1[t] = (dcommon && rb <= ra)
                                    ? dc1<T>(a0 + ra, a0 + rb, perq, QE, qn, qe, qg, lw, dw1, les, len, li, lv, neq, gpos)
                                    : 0.0;
                    }
                    __syncthreads();
                    const bool fast = na <= 8 && nalc <= 8 && 64 * na <= P.stage;
                    for (int B0 = 0; B0 < ncolumns; B0 += WAVE) {
                        const int bcol = B0 + lane;
                        const int nb_here = min(WAVE, ncolumns - B0);
                        if (fast) {
                            if (bcol < ncolumns) {
                                double g[8];
#pragma unroll
                                for (int k = 0; k < 8; k++) g[k] = k < nalc ? G[bcol * nalc + k] : 0.0;
                                const int sb = segof[bcol];
                                const bool diagblk = sb == sa;
                                const bool common = !diagblk && (segm[sa] & segm[sb]) != 0ull;
#pragma unroll
                                for (int ra = 0; ra < 8; ra++) {
                                    if (ra >= na) break;
                                    const int aa = a0 + ra;
                                    double v = 0.0;
                                    if (bcol <= aa) {
                                        double d = 0.0;
#pragma unroll
                                        for (int k = 0; k < 8; k++)
                                            if (k < nalc) d = d + ghs[ra * nalc + k] * g[k];
                                        double c = 0.0;
                                        if (diagblk) c = d1[ra * na + bcol - a0];     // na <= 8 here
                                        else if (common)
                                            c = dc1<T>(aa, bcol, perq, QE, qn, qe, qg, lw, dw1, les, len, li, lv, neq, gpos);
                                        v = c - d;
                                    }
                                    stage[lane * na + ra] = v;
                                }
                            }
                            __syncthreads();
                            for (int t = lane; t < nb_here * na; t += WAVE) dst[roff + (int64_t)B0 * na + t] = stage[t];
                            __syncthreads();
                        } else if (bcol < ncolumns) {
                            const int sb = segof[bcol];
                            const bool diagblk = sb == sa;
                            const bool common = !diagblk && (segm[sa] & segm[sb]) != 0ull;
                            for (int ra = 0; ra < na; ra++) {
                                const int aa = a0 + ra;
                                double v = 0.0;
                                if (bcol <= aa) {
                                    double d = 0.0;
                                    for (int k = 0; k < nalc; k++) d = d + ghs[ra * nalc + k] * G[bcol * nalc + k];
                                    double c = 0.0;
                                    if (diagblk && use_d1) c = d1[ra * na + bcol - a0];
                                    else if (diagblk || common)
                                        c = dc1<T>(aa, bcol, perq, QE, qn, qe, qg, lw, dw1, les, len, li, lv, neq, gpos);
                                    v = c - d;
                                }
                                dst[roff + (int64_t)bcol * na + ra] = v;
                            }
                        }
                    }
                    roff += (int64_t)ncolumns * na;
                    __syncthreads();
                }
            }
        }
        if (lane == 0) { out.status[r] = 0; out.chi2[r] = summ * dw1; out.ndf[r] = ndf; }
        __syncthreads();
    }
}

// Block gather of the two-phase matrix build: every stored block adds the values that the
// records of the batch wrote for it (in processing order).  One wavefront per block; lane e
// handles element e of the r x c block (row-major in the matrix, column-major in the records'
// updates); diagonal blocks use their lower triangle.
__global__ __launch_bounds__(256) void k_block_gather(int64_t nblock, const int64_t *boff, const int32_t *brg,
                                                      const int32_t *bcol, const int32_t *gfirst, const int64_t *ptr,
                                                      const int64_t *src, const double *cbuf, double *mat)
{
    const int64_t k = blockIdx.x * (int64_t)(blockDim.x >> 6) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (k >= nblock) return;
    const int64_t i0 = ptr[k], i1 = ptr[k + 1];
    if (i1 <= i0) return;
    const int ga = brg[k], gb = bcol[k];
    const int nr = gfirst[ga + 1] - gfirst[ga], nc = gfirst[gb + 1] - gfirst[gb];
    const bool diag = ga == gb;
    double *B = mat + boff[k];
    for (int e = lane; e < nr * nc; e += 64) {
        const int rr = e / nc, cc = e - rr * nc;
        if (diag && cc > rr) continue;
        const int se = cc * nr + rr;        // records store their blocks column by column
        // four independent partial sums keep four loads in flight
        double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0, acc3 = 0.0;
        int64_t q = i0;
        for (; q + 4 <= i1; q += 4) {
            acc0 += cbuf[src[q] + se];
            acc1 += cbuf[src[q + 1] + se];
            acc2 += cbuf[src[q + 2] + se];
            acc3 += cbuf[src[q + 3] + se];
        }
        for (; q < i1; q++) acc0 += cbuf[src[q] + se];
        B[e] += (acc0 + acc1) + (acc2 + acc3);
    }
}

// host launcher
hipError_t launch_records(const DevRecords &R, const DevModel &M, const PassOpts &o, const RecOut &out,
                          const LdsPlan &P, int64_t rec_begin, int64_t rec_end, int grid, hipStream_t s)
{
    if (rec_end <= rec_begin) return hipSuccess;
    if (P.bytes > 65536) {
        hipError_t e = R.vf ? hipFuncSetAttribute((const void *)k_records<float>, hipFuncAttributeMaxDynamicSharedMemorySize, P.bytes)
                            : hipFuncSetAttribute((const void *)k_records<double>, hipFuncAttributeMaxDynamicSharedMemorySize, P.bytes);
        if (e != hipSuccess) return e;
    }
    if (R.vf)
        hipLaunchKernelGGL(k_records<float>, dim3(grid), dim3(WAVE), P.bytes, s, R, M, o, out, P, rec_begin, rec_end);
    else
        hipLaunchKernelGGL(k_records<double>, dim3(grid), dim3(WAVE), P.bytes, s, R, M, o, out, P, rec_begin, rec_end);
    return hipGetLastError();
}

hipError_t launch_block_gather(int64_t nblock, const int64_t *boff, const int32_t *brg, const int32_t *bcol,
                               const int32_t *gfirst, const int64_t *ptr, const int64_t *src, const double *cbuf,
                               double *mat, hipStream_t s)
{
    if (nblock <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_block_gather, dim3((unsigned)((nblock + 3) / 4)), dim3(256), 0, s, nblock, boff, brg, bcol, gfirst,
                       ptr, src, cbuf, mat);
    return hipGetLastError();
}

int lds_bytes(LdsPlan &P)
{
    int npk = (P.nalc_max * P.nalc_max + P.nalc_max) / 2;
    size_t d = (size_t)P.max_gl + 4 * (size_t)P.max_eq + npk + 4 * (size_t)P.nalc_max +
               (size_t)(1 + P.gh) * P.max_var * P.nalc_max + (size_t)P.max_var * P.qe;
    size_t i = 2 * (size_t)P.max_gl + 3 * (size_t)P.max_var + 1 + (size_t)P.max_var * P.qe + (P.nalc_max + 1) + 8;
    i += (size_t)P.max_w + 2 * (size_t)P.max_eq + P.max_var;
    size_t b = d * 8 + (size_t)P.stage * 8 + (((size_t)P.max_w * 8 + 7) / 8) * 8 + (size_t)P.max_grp * P.nalc_max * 8 +
               (size_t)std::min(P.max_grp, 16) * std::min(P.max_grp, 16) * 8 + (size_t)P.max_var * 8 + i * 4;
    P.bytes = (int)b;
    return P.bytes;
}

} // namespace mpgpu
