// Record-loop kernel for local fits with bordered-band structure: one workgroup of 256
// threads per record.
//
// Restates loopbf (src/pede.f90:3387-4442) for the records whose local normal matrix pede
// solves with sqmibb / sqmibb2 (src/mpnum.f90:3118-3628): a band of half-width mbnd plus mbdr
// border rows at the start (side 1) or at the end (side 2), as chosen per record by the cost
// model of src/pede.f90:3838-3874 (evaluated on the host when the structure is built).
//
//  * the local matrix is accumulated directly into band and border storage (equations in
//    order, the same per-element arithmetic as the full path);
//  * root-free band Cholesky (dbcdec), band solves for the border columns and the right-hand
//    side (dbcslv), border Schur complement inverted with sqminv, back-substitution - the
//    operation order of sqmibb, so rank and condition decisions are pede's;
//  * the complete inverse (inv = 2) is formed column by column from the right with the dbcinv
//    recursion (the entries of a column need only the mbnd columns to its right, held in an
//    LDS ring), plus the border corrections, into a per-workgroup global scratch area;
//  * G Gamma^-1 G^T (dbavats order, src/mpnum.f90:1470-1528) uses support classes: runs of
//    record globals whose rows of G have the same local-parameter support share the same
//    block of Gamma^-1, so each matrix element costs |S| FMAs and one FP64 atomic.
#include "mpgpu_device.h"
#include <float.h>

namespace mpgpu {

#define BNT 256
#define BWAVES (BNT / 64)
#define NWM 4          // 64-bit words of a local-parameter support mask (nalc <= 256)
#define QCH 16         // support entries of one class handled per register chunk

__device__ __forceinline__ int64_t bpk64(int64_t i, int64_t j) { return (i * i - i) / 2 + j - 1; }
// packed index of (i, j) in either order, 1-based
__device__ __forceinline__ int64_t bsym(int i, int j) { return i >= j ? bpk64(i, j) : bpk64(j, i); }

template <typename T> __device__ __forceinline__ double bval(const DevRecords &R, int64_t w);
template <> __device__ __forceinline__ double bval<float>(const DevRecords &R, int64_t w) { return (double)R.vf[w]; }
template <> __device__ __forceinline__ double bval<double>(const DevRecords &R, int64_t w) { return R.vd[w]; }

__device__ __forceinline__ int64_t band_block_addr(const DevModel &M, int32_t blk, int va, int vb, int ga, int gb)
{
    int fa = M.gfirst[ga], fb = M.gfirst[gb];
    int nc = M.gfirst[gb + 1] - fb;
    return M.boff[blk] + (int64_t)(va - fa) * nc + (vb - fb);
}

__device__ __forceinline__ void band_mat_add(const DevModel &M, int va, int vb, double v, const int32_t *seg_of,
                                             const int32_t *seg_grp, const int32_t *pblk, int pos_a, int pos_b)
{
    if (!M.sparse) {
        atomicAdd(&M.mat[bpk64(va, vb)], v);
    } else {
        int sa = seg_of[pos_a], sb = seg_of[pos_b];
        int blk = pblk[(sa * sa + sa) / 2 + sb];
        if (blk < 0) return;
        atomicAdd(&M.mat[band_block_addr(M, blk, va, vb, seg_grp[sa], seg_grp[sb])], v);
    }
}

// block-wide sum (all threads return the total)
__device__ double block_sum(double v, double *red)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
    for (int k = 0; k < BWAVES; k++) s += red[k];
    __syncthreads();
    return s;
}

// symmetric Gauss-Jordan with diagonal pivoting (sqminv, src/mpnum.f90:97-213), one thread,
// on a small packed matrix (the border Schur complement); 1-based v, b
__device__ void sqminv_small(double *v, double *b, int n, int *nrank, double *diag, int *next)
{
    const double eps = 16.0 * DBL_EPSILON;
    int next0 = 1, l = 1;
    for (int i = 1; i <= n; i++) {
        next[i] = i + 1;
        diag[i] = fabs(v[(i * i + i) / 2]);
    }
    next[n] = -1;
    *nrank = 0;
    for (int i = 1; i <= n; i++) {
        int k = 0;
        double vkk = 0.0;
        int j = next0, last = 0;
        while (j > 0) {
            int jj = (j * j + j) / 2;
            double m = fabs(vkk) > eps * diag[j] ? fabs(vkk) : eps * diag[j];
            if (fabs(v[jj]) > m) { vkk = v[jj]; k = j; l = last; }
            last = j;
            j = next[last];
        }
        if (k != 0) {
            int kk = (k * k + k) / 2;
            if (l == 0) next0 = next[k]; else next[l] = next[k];
            next[k] = 0;
            (*nrank)++;
            vkk = 1.0 / vkk;
            v[kk] = -vkk;
            b[k] = b[k] * vkk;
            int jk = kk - k, jl = 0;
            for (int jx = 1; jx <= n; jx++) {
                if (jx == k) { jk = kk; jl = jl + jx; }
                else {
                    if (jx < k) jk = jk + 1; else jk = jk + jx - 1;
                    double vjk = v[jk];
                    v[jk] = vkk * vjk;
                    b[jx] = b[jx] - b[k] * vjk;
                    int lk = kk - k;
                    for (int lx = 1; lx <= jx; lx++) {
                        jl = jl + 1;
                        if (lx == k) lk = kk;
                        else {
                            if (lx < k) lk = lk + 1; else lk = lk + lx - 1;
                            v[jl] = v[jl] - v[lk] * vjk;
                        }
                    }
                }
            }
        } else {
            for (int kx = 1; kx <= n; kx++) {
                if (next[kx] != 0) {
                    b[kx] = 0.0;
                    for (int jx = 1; jx <= kx; jx++)
                        if (next[jx] != 0) v[(kx * kx - kx) / 2 + jx] = 0.0;
                }
            }
            break;
        }
    }
    for (int i = 1; i <= (n * n + n) / 2; i++) v[i] = -v[i];
}

// band solve with the root-free factor (dbcslv, src/Dbandmatrix.f90:521-560), one thread;
// w column-major w(mp1, n) 1-based through W(r, c) = w[(c - 1) * mp1 + r]; x 1-based, in place
__device__ void dbcslv_thread(const double *w, int mp1, int n, double *x)
{
    for (int i = 1; i <= n; i++) {
        const double xi = x[i];
        const int jm = min(mp1 - 1, n - i);
        for (int j = 1; j <= jm; j++) x[j + i] = x[j + i] - w[(i - 1) * mp1 + j + 1] * xi;
    }
    for (int i = n; i >= 1; i--) {
        double rxw = x[i] * w[(i - 1) * mp1 + 1];
        const int jm = min(mp1 - 1, n - i);
        for (int j = 1; j <= jm; j++) rxw = rxw - w[(i - 1) * mp1 + j + 1] * x[j + i];
        x[i] = rxw;
    }
}

template <typename T>
__global__ __launch_bounds__(BNT) void k_records_band(DevRecords R, DevModel M, PassOpts o, RecOut out, BandPlan P,
                                                      const int64_t *ids, int64_t nb, const int32_t *binfo,
                                                      double *scratch)
{
    extern __shared__ double lds[];
    const int tid = threadIdx.x;
    const double chuber = 1.345, cauchy = 2.3849;
    const int NA = P.nalc_max;

    // ---- carve LDS (1-based arrays get one spare element) ----
    double *gp = lds;                              // [max_gl]
    double *rm = gp + P.max_gl;                    // [max_eq]
    double *re = rm + P.max_eq;
    double *lc = re + P.max_eq;
    double *lw = lc + P.max_eq;
    double *vbnd = lw + P.max_eq;                  // [NA*mp1_max + 1] band, W(r,c) = vbnd[(c-1)*mp1 + r]
    double *vbdr = vbnd + NA * P.mp1_max + 1;      // [nbdr_max*NA + 1]
    double *aux = vbdr + P.nbdr_max * NA + 1;      // [nbdr_max*NA + NA + 1]
    double *bl = aux + (P.nbdr_max + 1) * NA + 1;  // [NA + 1]
    double *bs = bl + NA + 1;                      // [NA + 1]
    double *vbk = bs + NA + 1;                     // [nbdr_max(nbdr_max+1)/2 + 1]
    double *vzru = vbk + (P.nbdr_max * P.nbdr_max + P.nbdr_max) / 2 + 1; // [nbdr_max + 1]
    double *sdiag = vzru + P.nbdr_max + 1;         // [nbdr_max + 1]
    double *ring = sdiag + P.nbdr_max + 1;         // [mp1_max * (NA + 1)]
    double *gval = ring + P.mp1_max * (NA + 1);    // [max_sup]
    double *red = gval + P.max_sup;                // [8]
    unsigned long long *mask = (unsigned long long *)(red + 8); // [max_var * NWM]
    int32_t *gv = (int32_t *)(mask + (size_t)P.max_var * NWM);  // [max_gl]
    int32_t *gpos = gv + P.max_gl;                 // [max_gl]
    int32_t *gvl = gpos + P.max_gl;                // [max_var]
    int32_t *segof = gvl + P.max_var;              // [max_var]
    int32_t *seggr = segof + P.max_var;            // [max_var]
    int32_t *clsof = seggr + P.max_var;            // [max_var]
    int32_t *gptr = clsof + P.max_var;             // [max_var + 1]
    int32_t *cls0 = gptr + P.max_var + 1;          // [max_var + 1] first global of each class
    int32_t *sptr = cls0 + P.max_var + 1;          // [max_var + 1] support offsets of each class
    int32_t *upre = sptr + P.max_var + 1;          // [max_var + 1] work-unit prefix per class
    int32_t *sidx = upre + P.max_var + 1;          // [max_sup] support (local indices) of each class
    int32_t *pblk = sidx + P.max_sup;              // [max_pair]
    int32_t *snext = pblk + P.max_pair;            // [nbdr_max + 1]
    int32_t *sh = snext + P.nbdr_max + 1;          // [32] scalars

    double *vinv = scratch + (size_t)blockIdx.x * P.vinv_stride;   // packed inverse, 0-based bpk64

    for (int64_t ri = blockIdx.x; ri < nb; ri += gridDim.x) {
        const int64_t r = ids[ri];
        const int32_t info = binfo[r];
        const int mbnd = (info & 0x3ff) - 1, mbdr = (info >> 10) & 0x3ff, mside = (info >> 20) & 3;
        const int64_t e0 = R.rec_eq[r];
        const int neq = (int)(R.rec_eq[r + 1] - e0);
        const int64_t g0 = R.rec_gl[r];
        const int ngl = (int)(R.rec_gl[r + 1] - g0);
        const double dw1 = R.rec_w ? (double)R.rec_w[r] : 1.0;
        const double dw2 = sqrt(dw1);
        if (neq > P.max_eq || ngl > P.max_gl || mbnd + 1 > P.mp1_max || mbdr > P.nbdr_max) {
            if (tid == 0) { out.status[r] = 9; out.chi2[r] = 0.0; out.ndf[r] = 0; }
            continue;
        }
        // ---- distinct labels (wave 0) ----
        if (tid < 64) {
            int nalg = 0;
            for (int base = 0; base < ngl; base += 64) {
                int k = base + tid;
                int v = 0;
                if (k < ngl) {
                    int t = R.gl_tix[g0 + k];
                    gp[k] = M.par[t];
                    v = M.var[t];
                    gv[k] = v;
                }
                bool isv = (k < ngl) && v > 0;
                unsigned long long m = __ballot(isv);
                int pos = nalg + __popcll(m & ((1ull << tid) - 1ull));
                if (k < ngl) gpos[k] = isv ? pos : -1;
                if (isv && pos < P.max_var) gvl[pos] = v;
                nalg += __popcll(m);
            }
            if (tid == 0) { sh[0] = nalg; sh[1] = 0; }
        }
        __syncthreads();
        const int nalg = sh[0];
        // ---- first loop: residuals after the global parameters, number of locals ----
        for (int i = tid; i < neq; i += BNT) {
            int64_t sl = R.eq_sl[e0 + i];
            uint32_t en = R.eq_n[e0 + i];
            int nl = en & 0xffff, ng = en >> 16;
            int64_t sg = sl + nl + 1;
            double rmeas = bval<T>(R, sl);
            for (int j = 1; j <= ng; j++) rmeas = rmeas - bval<T>(R, sg + j) * gp[R.widx[sg + j]];
            rm[i] = rmeas;
            // 'scaleerrors' (prepareRecords, src/mppar.f90:312-316)
            re[i] = bval<T>(R, sg) * (ng > 0 ? o.esc[0] : o.esc[1]);
            lc[i] = 0.0;
            int mx = 0;
            for (int j = 1; j <= nl; j++) mx = max(mx, R.widx[sl + j]);
            if (mx > 0) atomicMax(&sh[1], mx);
        }
        __syncthreads();
        const int nalc = sh[1];
        const int nmb = nalc - mbdr, mp1 = mbnd + 1;
        if (nalg > P.max_var || nalc > NA || nalc <= 0 || nmb <= 0) {
            if (tid == 0) { out.status[r] = 9; out.chi2[r] = 0.0; out.ndf[r] = 0; }
            __syncthreads();
            continue;
        }
        const int nb1 = mside == 1 ? mbdr + 1 : nmb + 1;     // first band (side 1) / border (side 2) index
        const int bo = mside == 1 ? mbdr : 0;                  // band offset in local numbering
        // ---- local fit iterations (3759-4094) ----
        const int nter = (o.nloopn != 1 && o.lhuber != 0) ? o.lhuber : 1;
        int status = 0, nrank = 0, ndf = 0;
        double summ = 0.0;
        float resing = 0.0f;
        for (int iter = 1; iter <= nter; iter++) {
            for (int i = tid; i < neq; i += BNT) {
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
            for (int t = tid; t < nmb * mp1 + 1; t += BNT) vbnd[t] = 0.0;
            for (int t = tid; t < mbdr * nalc + 1; t += BNT) vbdr[t] = 0.0;
            for (int t = tid; t <= nalc; t += BNT) bl[t] = 0.0;
            if (tid == 0) sh[2] = 0;
            __syncthreads();
            // local normal equations into band / border storage, equations in order
            for (int i = 0; i < neq; i++) {
                const int64_t sl = R.eq_sl[e0 + i];
                const int nl = R.eq_n[e0 + i] & 0xffff;
                const int npair = (nl * nl + nl) / 2;
                const double wght = lw[i];
                for (int t = tid; t < npair + nl; t += BNT) {
                    if (t >= npair) {
                        int j = t - npair + 1;
                        int ij = R.widx[sl + j];
                        bl[ij] = bl[ij] + wght * rm[i] * bval<T>(R, sl + j);
                        continue;
                    }
                    int j = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5) + 1;
                    while ((j * j - j) / 2 > t) j--;
                    while ((j * j + j) / 2 <= t) j++;
                    int k = t - (j * j - j) / 2 + 1;
                    int ij = R.widx[sl + j], ik = R.widx[sl + k];
                    double val = wght * bval<T>(R, sl + j) * bval<T>(R, sl + k);
                    int J = max(ij, ik), K = min(ij, ik);
                    if (mside == 1) {
                        if (K <= mbdr) vbdr[(K - 1) * nalc + J] += val;          // column K of the border
                        else if (J - K <= mbnd) vbnd[(K - nb1) * mp1 + J - K + 1] += val;
                        else sh[2] = 1;
                    } else {
                        if (J > nmb) vbdr[(J - nb1) * nalc + K] += val;           // row J of the border
                        else if (J - K <= mbnd) vbnd[(K - 1) * mp1 + J - K + 1] += val;
                        else sh[2] = 1;
                    }
                }
                __syncthreads();
            }
            if (sh[2]) { status = 9; break; }   // element outside the detected structure
            // ---- sqmibb / sqmibb2: root-free band Cholesky (dbcdec) ----
            for (int i = tid; i < nmb; i += BNT) aux[i + 1] = 16.0 * vbnd[i * mp1 + 1];
            __syncthreads();
            for (int i = 1; i <= nmb; i++) {
                const int jm = min(mp1 - 1, nmb - i);
                double piv = vbnd[(i - 1) * mp1 + 1];
                piv = (piv + aux[i] != aux[i]) ? 1.0 / piv : 0.0;
                // W(k, i+j) -= W(k+j, i) * rxw, rxw = W(j+1, i) * W(1, i); pairs (j, k) independent
                const int nup = (jm * (jm + 1)) / 2;
                double upd = 0.0;
                int tgt = -1;
                if (tid < nup) {
                    int j = 1, rem = tid;
                    while (rem >= jm + 1 - j) { rem -= jm + 1 - j; j++; }
                    int k = rem + 1;
                    double rxw = vbnd[(i - 1) * mp1 + j + 1] * piv;
                    tgt = (i + j - 1) * mp1 + k;
                    upd = vbnd[tgt] - vbnd[(i - 1) * mp1 + k + j] * rxw;
                }
                __syncthreads();
                if (tgt >= 0) vbnd[tgt] = upd;
                if (tid < jm) vbnd[(i - 1) * mp1 + tid + 2] = vbnd[(i - 1) * mp1 + tid + 2] * piv;
                if (tid == 0) vbnd[(i - 1) * mp1 + 1] = piv;
                __syncthreads();
            }
            // rank of the band part and its condition (inverted pivots, as sqmibb)
            if (tid == 0) {
                int ok = 1;
                double emin = 0.0, emax = 0.0;
                for (int i = 1; i <= nmb; i++) {
                    double d = vbnd[(i - 1) * mp1 + 1];
                    if (d <= 0.0) { ok = 0; break; }
                    if (i == 1) { emin = d; emax = d; }
                    else { emin = fmin(emin, d); emax = fmax(emax, d); }
                }
                sh[3] = ok;
                red[4] = emin;
                red[5] = emax;
            }
            __syncthreads();
            const bool bandok = sh[3] != 0;
            const double evdmin = red[4], evdmax = red[5];
            __syncthreads();
            int nrankb = 0;
            if (!bandok) {
                for (int t = tid; t <= nalc; t += BNT) bl[t] = 0.0;
                nrank = 0;
                __syncthreads();
            } else if (mbdr == 0) {
                if (tid == 0) dbcslv_thread(vbnd, mp1, nmb, bl);
                nrank = nmb;
                __syncthreads();
            } else {
                // border columns and the band part of the right-hand side, one thread each
                const int n = nalc;
                if (tid < mbdr) {
                    const int ib = tid + 1;
                    double *ax = (mside == 1) ? &aux[nb1 + (ib - 1) * n - 1] : &aux[(ib - 1) * n];
                    const double *src = (mside == 1) ? &vbdr[nb1 + (ib - 1) * n - 1] : &vbdr[(ib - 1) * n];
                    for (int i = 1; i <= nmb; i++) ax[i] = src[i];
                    dbcslv_thread(vbnd, mp1, nmb, ax);
                } else if (tid == mbdr) {
                    const double *src = (mside == 1) ? &bl[nb1 - 1] : &bl[0];
                    for (int i = 1; i <= nmb; i++) bs[i] = src[i];
                    dbcslv_thread(vbnd, mp1, nmb, bs);
                }
                __syncthreads();
                // VZRU = b_border - AUX^T b_band (original b), VBK = C_bb - C_bn AUX
                const int nq = (mbdr * mbdr + mbdr) / 2;
                if (tid < mbdr) {
                    const int ib = tid + 1;
                    double z;
                    if (mside == 1) {
                        const int ioff = nb1 + (ib - 1) * n;
                        z = bl[ib];
                        for (int i = 0; i <= nmb - 1; i++) z = z - bl[nb1 + i] * aux[ioff + i];
                    } else {
                        const int ioff = (ib - 1) * n;
                        z = bl[nmb + ib];
                        for (int i = 1; i <= nmb; i++) z = z - bl[i] * aux[ioff + i];
                    }
                    vzru[ib] = z;
                } else if (tid >= 64 && tid < 64 + nq) {
                    int q = tid - 64 + 1, ib = 1;
                    while ((ib * ib + ib) / 2 < q) ib++;
                    const int jb = q - (ib * ib - ib) / 2;
                    double s;
                    if (mside == 1) {
                        const int ioff = nb1 + (ib - 1) * n, joff = nb1 + (jb - 1) * n;
                        s = vbdr[(jb - 1) * n + ib];
                        for (int i = 0; i <= nmb - 1; i++) s = s - vbdr[ioff + i] * aux[joff + i];
                    } else {
                        const int ioff = (ib - 1) * n, joff = (jb - 1) * n;
                        s = vbdr[(ib - 1) * n + nmb + jb];
                        for (int i = 1; i <= nmb; i++) s = s - vbdr[ioff + i] * aux[joff + i];
                    }
                    vbk[q] = s;
                }
                __syncthreads();
                if (tid == 0) {
                    int nr = 0;
                    sqminv_small(vbk, vzru, mbdr, &nr, sdiag, snext);
                    if (nr != mbdr) {
                        for (int ib = 1; ib <= mbdr; ib++) vzru[ib] = 0.0;
                        for (int q = 1; q <= nq; q++) vbk[q] = 0.0;
                    }
                    sh[4] = nr;
                }
                __syncthreads();
                nrankb = sh[4];
                nrank = nmb + (nrankb == mbdr ? mbdr : 0);
                // back-substitution of the band part with the border solution
                for (int i = tid + 1; i <= nmb; i += BNT) {
                    double x = bs[i];
                    if (mside == 1) {
                        for (int ib = 1; ib <= mbdr; ib++) x = x - vzru[ib] * aux[nb1 + (ib - 1) * n + i - 1];
                        bl[nb1 + i - 1] = x;
                    } else {
                        for (int ib = 1; ib <= mbdr; ib++) x = x - vzru[ib] * aux[(ib - 1) * n + i];
                        bl[i] = x;
                    }
                }
                if (tid < mbdr) bl[(mside == 1 ? 0 : nmb) + tid + 1] = vzru[tid + 1];
                __syncthreads();
            }
            float cndl10 = 0.0f;
            if (bandok && evdmin > 0.0) cndl10 = log10f((float)(evdmax / evdmin));
            // NaN check and third loop: residuals, chi^2 (3903-4035)
            int nan = 0;
            for (int i = tid + 1; i <= nalc; i += BNT) {
                double b = bl[i];
                if (!(b <= 0.0) && !(b > 0.0)) nan++;
            }
            double s_chi = 0.0, s_wt = 0.0;
            for (int i = tid; i < neq; i += BNT) {
                int64_t sl = R.eq_sl[e0 + i];
                int nl = R.eq_n[e0 + i] & 0xffff;
                double rerr = re[i];
                double wght = 1.0 / (rerr * rerr);
                double rmloc = 0.0;
                for (int j = 1; j <= nl; j++) rmloc = rmloc + bval<T>(R, sl + j) * bl[R.widx[sl + j]];
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
                if (iter == nter) rm[i] = rmeas;
            }
            nan = (int)block_sum((double)nan, red);
            summ = block_sum(s_chi, red);
            double suwt = block_sum(s_wt, red);
            ndf = neq - nrank;
            resing = ((float)neq - (float)suwt) / (float)neq;
            if (nan > 0) { status = 1; break; }
            if (nrank != nalc) { status = 2; break; }
            if (cndl10 > o.cndlmx) { status = 3; break; }
            if (ndf <= 0) { status = 4; break; }
        }
        double chi_add = 0.0;
        int ndf_add = 0;
        if (status == 0) {
            float chichi;
            {
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
            if (tid == 0) { out.status[r] = (int8_t)status; out.chi2[r] = chi_add * dw1; out.ndf[r] = ndf_add; }
            __syncthreads();
            continue;
        }
        // ---- segments of the record's variables (group-block storage) ----
        if (o.icalcm == 1 && M.sparse && tid < 64) {
            int nseg = 0;
            for (int base = 0; base < nalg; base += 64) {
                int a = base + tid;
                int g = a < nalg ? M.group_of[gvl[a]] : -1;
                int gprev = (a > 0 && a < nalg) ? M.group_of[gvl[a - 1]] : -2;
                bool st = (a < nalg) && (a == 0 || g != gprev);
                unsigned long long m = __ballot(st);
                int s = nseg + __popcll(m & ((2ull << tid) - 1ull)) - 1;
                if (a < nalg) { segof[a] = s; if (st) seggr[s] = g; }
                nseg += __popcll(m);
            }
            int npair = (nseg * nseg + nseg) / 2;
            const int64_t p0 = R.rec_pb[r];
            for (int t = tid; t < npair; t += 64) pblk[t] = t < P.max_pair ? R.pair_blk[p0 + t] : -1;
        }
        if (o.icalcm == 1)
            for (int t = tid; t < nalg * NWM; t += BNT) mask[t] = 0ull;
        __syncthreads();
        // ---- fourth loop: vector, counters, g g^T (4162-4201), support masks of G ----
        double s_chi = 0.0;
        const bool down = (o.nloopn != 1 && o.lhuber != 0);
        for (int i = tid; i < neq; i += BNT) {
            int64_t sl = R.eq_sl[e0 + i];
            uint32_t en = R.eq_n[e0 + i];
            int nl = en & 0xffff, ng = en >> 16;
            int64_t sg = sl + nl + 1;
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
            lw[i] = wght;
            for (int j = 1; j <= ng; j++) {
                double gj = bval<T>(R, sg + j);
                int pj = R.widx[sg + j];
                int vj = gv[pj];
                if (gj == 0.0) continue;
                if (vj > 0) {
                    atomicAdd(&M.vec[vj - 1], dw1 * wght * rmeas * gj);
                    atomicAdd(&M.cnt[vj - 1], 1ull);
                    if (o.icalcm == 1) {
                        int aj = gpos[pj];
                        for (int l = 1; l <= nl; l++) {
                            int kk = R.widx[sl + l] - 1;
                            atomicOr(&mask[aj * NWM + (kk >> 6)], 1ull << (kk & 63));
                        }
                        for (int k = 1; k <= j; k++) {
                            int pk_ = R.widx[sg + k];
                            int vk = gv[pk_];
                            if (vk > 0) {
                                double val = dw1 * wght * gj * bval<T>(R, sg + k);
                                int ak = gpos[pk_];
                                if (vj >= vk) band_mat_add(M, vj, vk, val, segof, seggr, pblk, aj, ak);
                                else band_mat_add(M, vk, vj, val, segof, seggr, pblk, ak, aj);
                            }
                        }
                    }
                }
            }
        }
        summ = block_sum(s_chi, red);
        if (o.icalcm == 1) {
            // ---- support classes: runs of globals with the same support of G ----
            if (tid == 0) {
                int ncls = 0, tot = 0, ok = 1;
                for (int a = 0; a < nalg; a++) {
                    bool same = a > 0;
                    for (int w = 0; w < NWM && same; w++) same = mask[a * NWM + w] == mask[(a - 1) * NWM + w];
                    if (!same) {
                        int cnt = 0;
                        for (int w = 0; w < NWM; w++) cnt += __popcll(mask[a * NWM + w]);
                        cls0[ncls] = a;
                        sptr[ncls] = tot;
                        tot += cnt;
                        ncls++;
                    }
                    clsof[a] = ncls - 1;
                }
                cls0[ncls] = nalg;
                sptr[ncls] = tot;
                // G values: one row of |S| entries per global
                int gt = 0;
                for (int a = 0; a < nalg; a++) {
                    gptr[a] = gt;
                    int q = clsof[a];
                    gt += sptr[q + 1] - sptr[q];
                }
                gptr[nalg] = gt;
                if (tot > P.max_sup || gt > P.max_sup) ok = 0;
                // work units of the Schur update: (global a, class Q <= class(a))
                int u = 0;
                for (int q = 0; q < ncls; q++) { upre[q] = u; u += nalg - cls0[q]; }
                upre[ncls] = u;
                sh[5] = ncls;
                sh[6] = ok;
            }
            __syncthreads();
            const int ncls = sh[5];
            if (!sh[6]) {
                if (tid == 0) { out.status[r] = 9; out.chi2[r] = 0.0; out.ndf[r] = 0; }
                __syncthreads();
                continue;
            }
            // support lists (ascending local indices) and zeroed G rows
            for (int q = tid; q < ncls; q += BNT) {
                int a = cls0[q], s = sptr[q];
                for (int w = 0; w < NWM; w++) {
                    unsigned long long m = mask[a * NWM + w];
                    while (m) {
                        int b = __ffsll((long long)m) - 1;
                        sidx[s++] = w * 64 + b + 1;
                        m &= m - 1;
                    }
                }
            }
            for (int t = tid; t < gptr[nalg]; t += BNT) gval[t] = 0.0;
            __syncthreads();
            // G = sum over equations of sqrt(dw) w g l^T, equations in order (4206-4219)
            for (int i = 0; i < neq; i++) {
                const int64_t sl = R.eq_sl[e0 + i];
                const uint32_t en = R.eq_n[e0 + i];
                const int nl = en & 0xffff, ng = en >> 16;
                const int64_t sg = sl + nl + 1;
                const double wght = lw[i];
                for (int t = tid; t < ng * nl; t += BNT) {
                    int j = t / nl + 1, k = t % nl + 1;
                    const int pj = R.widx[sg + j];
                    int a = gpos[pj];
                    if (a < 0) continue;
                    const int ik = R.widx[sl + k];
                    // word pairs of one equation that share the element (a label or a local
                    // index repeated in the equation) are summed by the thread of the first
                    // such pair, in the reference's (j, k) order
                    bool own = true, rep = false;
                    for (int x = 1; x <= ng; x++)
                        if (x != j && R.widx[sg + x] == pj) { rep = true; own = own && x > j; }
                    for (int x = 1; x <= nl; x++)
                        if (x != k && R.widx[sl + x] == ik) { rep = true; own = own && x > k; }
                    // zero derivatives add nothing and are not part of the support masks
                    if (!own || (!rep && bval<T>(R, sg + j) == 0.0)) continue;
                    const int kk = ik - 1;
                    int pos = 0;
                    for (int w = 0; w < (kk >> 6); w++) pos += __popcll(mask[a * NWM + w]);
                    pos += __popcll(mask[a * NWM + (kk >> 6)] & ((1ull << (kk & 63)) - 1ull));
                    double *ge = &gval[gptr[a] + pos];
                    if (!rep) {
                        *ge = *ge + dw2 * wght * bval<T>(R, sg + j) * bval<T>(R, sl + k);
                        continue;
                    }
                    // the element exists once some pair of the run has a non-zero derivative
                    double v = 0.0;
                    bool any = false;
                    for (int j2 = j; j2 <= ng; j2++) {
                        if (R.widx[sg + j2] != pj || bval<T>(R, sg + j2) == 0.0) continue;
                        for (int k2 = k; k2 <= nl; k2++) {
                            if (R.widx[sl + k2] != ik) continue;
                            if (!any) { v = *ge; any = true; }
                            v = v + dw2 * wght * bval<T>(R, sg + j2) * bval<T>(R, sl + k2);
                        }
                    }
                    if (any) *ge = v;
                }
                __syncthreads();
            }
            // ---- complete inverse of the local matrix (inv = 2) into vinv ----
            // band part: columns from the right with the dbcinv recursion
            //   vs(i,j) = [i==j] W(1,i) - sum_{k=j+1}^{j+mbnd} vs(max(i,k),min(i,k)) W(1+k-j, j)
            // the mbnd columns to the right of j live in the LDS ring (slot = column mod mp1)
            const int n = nalc;
            const int nq = (mbdr * mbdr + mbdr) / 2;
            for (int j = nmb; j >= 1; j--) {
                double *colj = &ring[(j % mp1) * (NA + 1)];
                const int kmax = min(nmb, j + mp1 - 1);
                // off-diagonal entries of column j (rows i > j) from the columns to the right
                for (int i = j + 1 + tid; i <= nmb; i += BNT) {
                    double rxw = 0.0;
                    for (int k = j + 1; k <= kmax; k++) {
                        double vik = (k <= i) ? ring[(k % mp1) * (NA + 1) + i] : ring[(i % mp1) * (NA + 1) + k];
                        rxw = rxw - vik * vbnd[(j - 1) * mp1 + 1 + k - j];
                    }
                    colj[i] = rxw;
                }
                __syncthreads();
                if (tid == 0) {
                    double rxw = vbnd[(j - 1) * mp1 + 1];
                    for (int k = j + 1; k <= kmax; k++) rxw = rxw - colj[k] * vbnd[(j - 1) * mp1 + 1 + k - j];
                    colj[j] = rxw;
                }
                __syncthreads();
                // store with the border correction sum_{ib,jb} VBK(ib,jb) AUX_ib(i) AUX_jb(j)
                for (int i = j + tid; i <= nmb; i += BNT) {
                    double v = colj[i];
                    if (mbdr > 0) {
                        for (int ib = 1; ib <= mbdr; ib++)
                            for (int jb = 1; jb <= mbdr; jb++) {
                                int ij = ib > jb ? ib : jb;
                                ij = (ij * ij - ij) / 2 + (ib < jb ? ib : jb);
                                const double ai = mside == 1 ? aux[nb1 + (ib - 1) * n + i - 1] : aux[(ib - 1) * n + i];
                                const double aj = mside == 1 ? aux[nb1 + (jb - 1) * n + j - 1] : aux[(jb - 1) * n + j];
                                v = v + vbk[ij] * ai * aj;
                            }
                    }
                    vinv[bpk64(bo + i, bo + j)] = v;
                }
                __syncthreads();
            }
            // border rows: -sum_jb VBK(ib,jb) AUX_jb(i) against the band, VBK within the border
            for (int t = tid; t < mbdr * nmb; t += BNT) {
                const int ib = t / nmb + 1, i = t % nmb + 1;
                double v = 0.0;
                for (int jb = 1; jb <= mbdr; jb++) {
                    int ij = ib > jb ? ib : jb;
                    ij = (ij * ij - ij) / 2 + (ib < jb ? ib : jb);
                    const double aj = mside == 1 ? aux[nb1 + (jb - 1) * n + i - 1] : aux[(jb - 1) * n + i];
                    v = v - vbk[ij] * aj;
                }
                const int dbi = mside == 1 ? ib : nmb + ib;
                vinv[bsym(dbi, bo + i)] = v;
            }
            for (int q = tid; q < nq; q += BNT) {
                int ib = 1;
                while ((ib * ib + ib) / 2 < q + 1) ib++;
                const int jb = q + 1 - (ib * ib - ib) / 2;
                const int dd = mside == 1 ? 0 : nmb;
                vinv[bpk64(dd + ib, dd + jb)] = vbk[q + 1];
            }
            __threadfence_block();
            __syncthreads();
            // ---- Schur complement: element (a >= b) -= sum_l (sum_{k in S(a)} G_ak Vinv(k,l)) G_bl ----
            const int nunit = upre[ncls];
            for (int u = tid; u < nunit; u += BNT) {
                int lo = 0, hi = ncls - 1;
                while (lo < hi) {
                    int mid = (lo + hi + 1) >> 1;
                    if (upre[mid] <= u) lo = mid; else hi = mid - 1;
                }
                const int Q = lo;
                const int a = cls0[Q] + (u - upre[Q]);
                const int ca = clsof[a];
                const int sa0 = sptr[ca], na_ = sptr[ca + 1] - sa0;
                const int sq0 = sptr[Q], nq_ = sptr[Q + 1] - sq0;
                const int bend = min(cls0[Q + 1] - 1, a);
                for (int l0 = 0; l0 < nq_; l0 += QCH) {
                    double m[QCH];
#pragma unroll
                    for (int l = 0; l < QCH; l++) m[l] = 0.0;
                    for (int s = 0; s < na_; s++) {
                        const int k = sidx[sa0 + s];
                        const double gak = gval[gptr[a] + s];
#pragma unroll
                        for (int l = 0; l < QCH; l++)
                            if (l0 + l < nq_) m[l] = m[l] + gak * vinv[bsym(k, sidx[sq0 + l0 + l])];
                    }
                    for (int b = cls0[Q]; b <= bend; b++) {
                        double c = 0.0;
#pragma unroll
                        for (int l = 0; l < QCH; l++)
                            if (l0 + l < nq_) c = c + m[l] * gval[gptr[b] + l0 + l];
                        band_mat_add(M, gvl[a], gvl[b], -c, segof, seggr, pblk, a, b);
                    }
                }
            }
        }
        if (tid == 0) { out.status[r] = 0; out.chi2[r] = summ * dw1; out.ndf[r] = ndf; }
        __syncthreads();
    }
}

hipError_t launch_records_band(const DevRecords &R, const DevModel &M, const PassOpts &o, const RecOut &out,
                               const BandPlan &P, const int64_t *ids, int64_t nb, const int32_t *binfo,
                               double *scratch, int grid, hipStream_t s)
{
    if (nb <= 0) return hipSuccess;
    if (P.bytes > 65536) {
        hipError_t e = R.vf ? hipFuncSetAttribute((const void *)k_records_band<float>,
                                                  hipFuncAttributeMaxDynamicSharedMemorySize, P.bytes)
                            : hipFuncSetAttribute((const void *)k_records_band<double>,
                                                  hipFuncAttributeMaxDynamicSharedMemorySize, P.bytes);
        if (e != hipSuccess) return e;
    }
    if (R.vf)
        hipLaunchKernelGGL(k_records_band<float>, dim3(grid), dim3(BNT), P.bytes, s, R, M, o, out, P, ids, nb, binfo,
                           scratch);
    else
        hipLaunchKernelGGL(k_records_band<double>, dim3(grid), dim3(BNT), P.bytes, s, R, M, o, out, P, ids, nb, binfo,
                           scratch);
    return hipGetLastError();
}

int band_lds_bytes(BandPlan &P)
{
    const size_t NA = P.nalc_max;
    size_t d = (size_t)P.max_gl + 4 * (size_t)P.max_eq + NA * P.mp1_max + 1 + (size_t)P.nbdr_max * NA + 1 +
               ((size_t)P.nbdr_max + 1) * NA + 1 + 2 * (NA + 1) + ((size_t)P.nbdr_max * P.nbdr_max + P.nbdr_max) / 2 + 1 +
               2 * ((size_t)P.nbdr_max + 1) + (size_t)P.mp1_max * (NA + 1) + P.max_sup + 8;
    size_t u64 = (size_t)P.max_var * NWM;
    size_t i32 = 2 * (size_t)P.max_gl + 4 * (size_t)P.max_var + 4 * ((size_t)P.max_var + 1) + P.max_sup + P.max_pair +
                 P.nbdr_max + 1 + 32;
    size_t b = d * 8 + u64 * 8 + i32 * 4;
    P.bytes = (int)b;
    return P.bytes;
}

} // namespace mpgpu
