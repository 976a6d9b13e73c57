// Record-loop kernel: one workgroup of RT threads (RW wavefronts) per Mille record.
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
// chi^2 and weight sums, which are wavefront reductions.  The local fit (a few dozen elements)
// runs in the first wavefront; loading, the equation lists, G and the update of the global
// system use the whole workgroup, which also keeps several wavefronts per record in flight to
// hide LDS and memory latency.  Matrix contributions go either into pede's packed full matrix
// (FP64 atomics) or, for group-block storage, into a per-record update buffer that
// k_block_gather sums per matrix block.
#include "mpgpu_device.h"
#include <float.h>
#include <algorithm>
#include <cstdio>
#include <cstdlib>

namespace mpgpu {

#define WAVE 64
typedef double d4 __attribute__((ext_vector_type(4)));
#define RT 256                  // threads per record workgroup
#define RW (RT / WAVE)          // wavefronts per record workgroup
// Lanes of one wavefront exchanging data through LDS: a wavefront executes its LDS operations
// in order, so a wavefront-scope barrier (no wait for the wave's outstanding global stores,
// unlike a workgroup barrier) orders them.
#define WSYNC()                                                    \
    do {                                                           \
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");     \
        __builtin_amdgcn_wave_barrier();                           \
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");     \
    } while (0)

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

// sqminv (src/mpnum.f90:97-213) for n = N <= 5 local parameters, every lane holding the whole
// packed matrix and right-hand side in registers (no cross-lane traffic): pivots in the
// reference's order (remaining candidates ascending, first largest diagonal above eps times its
// original size), the same elimination arithmetic as gj_lanes.  Reads cl/bl, writes the
// negated inverse and the solution back (lane 0) and returns the rank.
template <int N>
__device__ __forceinline__ int gj_regs(double *cl, double *bl, int lane)
{
    constexpr int NP = N * (N + 1) / 2;
    const double eps = 16.0 * DBL_EPSILON;
    double c[NP], b[N], dg[N];
#pragma unroll
    for (int t = 0; t < NP; t++) c[t] = cl[t];
#pragma unroll
    for (int j = 0; j < N; j++) {
        b[j] = bl[j];
        dg[j] = fabs(c[j * (j + 1) / 2 + j]);
    }
    unsigned rem = (1u << N) - 1u;   // candidates not yet pivoted
    int nrank = 0;
#pragma unroll
    for (int step = 0; step < N; step++) {
        int k = -1;
        double vkk = 0.0;
#pragma unroll
        for (int j = 0; j < N; j++) {
            const double vjj = c[j * (j + 1) / 2 + j];
            if (((rem >> j) & 1u) && fabs(vjj) > fmax(fabs(vkk), eps * dg[j])) { vkk = vjj; k = j; }
        }
        if (k < 0) {
            // no acceptable pivot left: the remaining rows and columns are set to zero
#pragma unroll
            for (int J = 0; J < N; J++)
#pragma unroll
                for (int K = 0; K <= J; K++)
                    if (((rem >> J) & 1u) && ((rem >> K) & 1u)) c[J * (J + 1) / 2 + K] = 0.0;
#pragma unroll
            for (int j = 0; j < N; j++)
                if ((rem >> j) & 1u) b[j] = 0.0;
            break;
        }
        rem &= ~(1u << k);
        nrank++;
        double co[N];   // column k of the current matrix
        double ckk = 0.0, bkk = 0.0;
#pragma unroll
        for (int j = 0; j < N; j++) {
            double v = 0.0;
#pragma unroll
            for (int q = 0; q < N; q++)
                if (q == k) v = j >= q ? c[j * (j + 1) / 2 + q] : c[q * (q + 1) / 2 + j];
            co[j] = v;
            if (j == k) { ckk = v; bkk = b[j]; }
        }
        const double vinv = 1.0 / ckk;
        const double bk = bkk * vinv;
#pragma unroll
        for (int J = 0; J < N; J++)
#pragma unroll
            for (int K = 0; K <= J; K++) {
                double &e = c[J * (J + 1) / 2 + K];
                if (J == k && K == k) e = -vinv;
                else if (J == k) e = vinv * co[K];
                else if (K == k) e = vinv * co[J];
                else e = e - (vinv * co[K]) * co[J];
            }
#pragma unroll
        for (int j = 0; j < N; j++) b[j] = (j == k) ? bk : b[j] - bk * co[j];
    }
    if (lane == 0) {
#pragma unroll
        for (int t = 0; t < NP; t++) cl[t] = -c[t];
#pragma unroll
        for (int j = 0; j < N; j++) bl[j] = b[j];
    }
    return nrank;
}

// Value of a lane-held double at a uniform lane
__device__ __forceinline__ double rdlane(double v, int l)
{
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}

// sqminv (src/mpnum.f90:97-213) for a packed symmetric n x n matrix held one element per lane
// (lane t < npk = n(n+1)/2) with the right-hand side in lanes npk .. npk+n-1 (npk + n <= 64):
// pivots in the reference's order (remaining candidates ascending, first largest diagonal
// above eps times its original size), elimination with lane shuffles.  Writes the negated
// result to cl[npk] and the solution to bl[n] (the form the LDS variant leaves) and returns
// the rank.
__device__ int gj_lanes(double acc, int npk, int n, int lane, double *cl, double *bl)
{
    const double eps = 16.0 * DBL_EPSILON;
    int J = 1, K = 1;
    if (lane < npk) {
        J = (int)((sqrt(8.0 * lane + 1.0) - 1.0) * 0.5) + 1;
        while ((J * J - J) / 2 > lane) J--;
        while ((J * J + J) / 2 <= lane) J++;
        K = lane - (J * J - J) / 2 + 1;
    }
    const int L1 = lane < n ? lane + 1 : 1;
    double cv = lane < npk ? acc : 0.0;
    const double cv0 = cv;
    double bv = __shfl(acc, min(npk + lane, WAVE - 1), WAVE);
    unsigned long long rem = n >= 64 ? ~0ull : (1ull << n) - 1ull;   // candidates not yet pivoted
    int nrank = 0;
    for (int step = 1; step <= n; step++) {
        int k = 0;
        double vkk = 0.0;
        for (int j = 1; j <= n; j++) {
            if (!((rem >> (j - 1)) & 1ull)) continue;
            const int d = pk(j, j);
            const double vjj = rdlane(cv, d);
            const double m = fmax(fabs(vkk), eps * fabs(rdlane(cv0, d)));
            if (fabs(vjj) > m) { vkk = vjj; k = j; }
        }
        if (k == 0) {
            // no acceptable pivot left: the remaining rows and columns are set to zero
            if (lane < npk && ((rem >> (J - 1)) & 1ull) && ((rem >> (K - 1)) & 1ull)) cv = 0.0;
            if (lane < n && ((rem >> lane) & 1ull)) bv = 0.0;
            break;
        }
        rem &= ~(1ull << (k - 1));
        nrank++;
        const double vinv = 1.0 / rdlane(cv, pk(k, k));
        const double coJ = __shfl(cv, J >= k ? pk(J, k) : pk(k, J), WAVE);
        const double coK = __shfl(cv, K >= k ? pk(K, k) : pk(k, K), WAVE);
        const double coL = __shfl(cv, L1 >= k ? pk(L1, k) : pk(k, L1), WAVE);
        const double bk = rdlane(bv, k - 1) * vinv;
        double nv;
        if (J == k && K == k) nv = -vinv;
        else if (J == k) nv = vinv * coK;
        else if (K == k) nv = vinv * coJ;
        else nv = cv - (vinv * coK) * coJ;
        if (lane < npk) cv = nv;
        if (lane < n) bv = (L1 == k) ? bk : bv - bk * coL;
    }
    if (lane < npk) cl[lane] = -cv;
    if (lane < n) bl[lane] = bv;
    return nrank;
}

// a rejected record contributes nothing: its slot of the update buffer is cleared
__device__ __forceinline__ void zero_update(const DevRecords &R, const DevModel &M, const PassOpts &o, int64_t r, int tid)
{
    if (o.icalcm != 1 || !M.sparse || R.cbuf == nullptr) return;
    const int64_t off = R.rec_cb[r];
    if (off < 0) return;
    const int n = R.rec_cn[r];
    for (int t = tid; t < n; t += RT) R.cbuf[off + t] = 0.0;
}

// LDS layout of the record kernel (byte offsets); the kernel carves its dynamic LDS with it and
// the host sizes the allocation with it, so both always agree
struct RecLds {
    int gp, rm, re, lc, lw, cl, bl, dg, cs, co, G, Gh, qg, dsh, segm, lmask, rdesc, vsegm, lv;
    int gv, gg, gpos, gvl, ggl, segof, segst, segro, qn, qe, nxt, sh, wcnt, li, les, len;
    int bytes;
};
__host__ __device__ inline RecLds rec_lds(const LdsPlan &P, int tsz)
{
    RecLds L;
    int o = 0;
    auto d = [&](int n) { int p = o; o += n * 8; return p; };
    auto i4 = [&](int n) { int p = o; o += n * 4; return p; };
    const int nv = P.max_var, ne = P.max_eq, ng = P.max_gl, nl = P.nalc_max, QE = P.qe;
    L.gp = d(ng);
    L.rm = d(ne);
    L.re = d(ne);
    L.lc = d(ne);
    L.lw = d(ne);
    L.cl = d((nl * nl + nl) / 2);
    L.bl = d(nl);
    L.dg = d(nl);
    L.cs = d(nl);
    L.co = d(nl);
    L.G = d(nv * nl);
    L.Gh = d(nv * nl);
    L.qg = d(nv * QE);
    L.dsh = d(8);
    L.segm = d(nv);
    L.lmask = d(ne);
    o = (o + 15) / 16 * 16;
    L.rdesc = d(P.stage > 0 ? 2 * nv : 0);
    L.vsegm = d(P.stage > 0 ? nv : 0);
    L.lv = o;
    o += ((P.max_w > 0 ? P.max_w : 1) * tsz + 7) / 8 * 8;
    L.gv = i4(ng);
    L.gg = i4(ng);
    L.gpos = i4(ng);
    L.gvl = i4(nv);
    L.ggl = i4(nv);
    L.segof = i4(nv);
    L.segst = i4(nv + 1);
    L.segro = i4(nv + 1);
    L.qn = i4(nv);
    L.qe = i4(nv * QE);
    L.nxt = i4(nl + 1);
    L.sh = i4(16);
    L.wcnt = i4((ng + WAVE - 1) / WAVE);
    L.li = i4(P.max_w);
    L.les = i4(ne);
    L.len = i4(ne);
    L.bytes = o;
    return L;
}

template <typename T, int MODE>
__global__ __launch_bounds__(RT) __attribute__((amdgpu_waves_per_eu(4))) void k_records(DevRecords R, DevModel M, PassOpts o, RecOut out, LdsPlan P,
                                                int64_t rec_begin, int64_t rec_end, MonCtx mon)
{
    extern __shared__ double lds[];
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wv = tid >> 6;
    const double chuber = 1.345, cauchy = 2.3849;
    const int QE = P.qe;                     // equations per record variable on the merged update path

    // ---- carve LDS ----
    const RecLds L = rec_lds(P, (int)sizeof(T));
    char *lb = (char *)lds;
    double *gp = (double *)(lb + L.gp);      // [max_gl] parameter value per distinct label
    double *rm = (double *)(lb + L.rm);      // [max_eq] corrected measurement / residual
    double *re = (double *)(lb + L.re);      // [max_eq] sigma
    double *lc = (double *)(lb + L.lc);      // [max_eq] local-fit prediction
    double *lw = (double *)(lb + L.lw);      // [max_eq] weight of the last local-fit pass
    double *cl = (double *)(lb + L.cl);      // [npk_max] local matrix -> inverse
    double *bl = (double *)(lb + L.bl);      // [nalc_max] local vector -> solution
    double *dg = (double *)(lb + L.dg);      // [nalc_max] |diag| for pivot test
    double *cs = (double *)(lb + L.cs);      // [nalc_max] scaled pivot column
    double *co = (double *)(lb + L.co);      // [nalc_max] original pivot column
    double *G = (double *)(lb + L.G);        // [max_var * nalc_max]; during the local fit the
    double *LD = G;                          //   dense local derivatives [neq][nalc] when they fit
    double *Gh = (double *)(lb + L.Gh);      // [max_var * nalc_max] G Gamma^-1
    double *qg = (double *)(lb + L.qg);      // [max_var*QE] derivative of a variable in its equations
    double *dsh = (double *)(lb + L.dsh);    // [8] double scalars
    unsigned long long *segm = (unsigned long long *)(lb + L.segm);   // [max_var] equation mask of each segment
    unsigned long long *lmask = (unsigned long long *)(lb + L.lmask); // [max_eq] local parameters of each equation
    T *lv = (T *)(lb + L.lv);                // [max_w] the record's word values
    int32_t *gv = (int32_t *)(lb + L.gv);    // [max_gl] var index
    int32_t *gg = (int32_t *)(lb + L.gg);    // [max_gl] parameter group (variables)
    int32_t *gpos = (int32_t *)(lb + L.gpos);   // [max_gl] position among the record's variables
    int32_t *gvl = (int32_t *)(lb + L.gvl);     // [max_var] sorted variable list
    int32_t *ggl = (int32_t *)(lb + L.ggl);     // [max_var] their parameter groups
    int32_t *segof = (int32_t *)(lb + L.segof); // [max_var] segment of each record variable
    int32_t *segst = (int32_t *)(lb + L.segst); // [max_var+1] first record variable of each segment
    int32_t *segro = (int32_t *)(lb + L.segro); // [max_var+1] first update value of each segment
    // [max_var] per record variable a: update-buffer offset of its row (column 0), the row's
    // column stride (its segment's size), its segment; and the segment's equation mask
    int4 *rdesc = (int4 *)(lb + L.rdesc);
    unsigned long long *vsegm = (unsigned long long *)(lb + L.vsegm);
    int32_t *qn = (int32_t *)(lb + L.qn);       // [max_var] equations of each record variable
    int32_t *qe = (int32_t *)(lb + L.qe);       // [max_var*QE] their indices (ascending)
    int32_t *nxt = (int32_t *)(lb + L.nxt);     // [nalc_max+1] pivot list
    int32_t *sh = (int32_t *)(lb + L.sh);       // [16] int scalars
    int32_t *wcnt = (int32_t *)(lb + L.wcnt);   // [max_gl/64] variables among each 64 labels
    int32_t *li = (int32_t *)(lb + L.li);       // [max_w] the record's word indices (local / global position)
    int32_t *les = (int32_t *)(lb + L.les);     // [max_eq] measurement word of each equation (record-local)
    uint32_t *len = (uint32_t *)(lb + L.len);   // [max_eq] local / global word counts
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
            if (tid == 0) { out.status[r] = 9; out.chi2[r] = 0.0; out.ndf[r] = 0; }
            zero_update(R, M, o, r, tid);
            continue;
        }
        // ---- the record's words and equation table into LDS (coalesced, once) ----
        for (int t = tid; t < nw; t += RT) {
            lv[t] = R.vf ? (T)R.vf[w0 + t] : (T)R.vd[w0 + t];
            li[t] = R.widx[w0 + t];
        }
        for (int i = tid; i < neq; i += RT) {
            les[i] = (int)(R.eq_sl[e0 + i] - w0);
            len[i] = R.eq_n[e0 + i];
        }
        if (tid == 0) { sh[0] = 0; /* nalc */ sh[7] = 0; /* next update segment */ }
        // ---- distinct labels: parameter value, variable index and group (loaded together with
        // the words; positions among the variables follow after the barrier) ----
        for (int k = tid; k < ngl; k += RT) {
            const int t = R.gl_tix[g0 + k];
            gp[k] = M.par[t];
            const int v = M.var[t];
            gv[k] = v;
            gg[k] = M.grp_tot[t];
            const unsigned long long m = __ballot(v > 0);
            if (lane == 0) wcnt[k >> 6] = __popcll(m);
        }
        __syncthreads();
        int nalg = 0;
        for (int c = 0; c < ((ngl + WAVE - 1) >> 6); c++) nalg += wcnt[c];
        if (nalg > P.max_var) {
            if (tid == 0) { out.status[r] = 9; out.chi2[r] = 0.0; out.ndf[r] = 0; }
            zero_update(R, M, o, r, tid);
            __syncthreads();
            continue;
        }
        for (int k = tid; k < ngl; k += RT) {
            const int v = gv[k];
            const unsigned long long m = __ballot(v > 0);
            int pos = __popcll(m & ((1ull << lane) - 1ull));
            for (int c = 0; c < (k >> 6); c++) pos += wcnt[c];
            gpos[k] = v > 0 ? pos : -1;
            if (v > 0) { gvl[pos] = v; ggl[pos] = gg[k]; }
        }
        // ---- first loop: subtract global contribution, count local parameters (3655-3693) ----
        {
            int nalc_t = 0;
            for (int i = tid; i < neq; i += RT) {
                const int sl = les[i];
                const uint32_t en = len[i];
                int nl = en & 0xffff, ng = en >> 16;
                const int sg = sl + nl + 1;
                double rmeas = lwv(sl);
                for (int j = 1; j <= ng; j++) rmeas = rmeas - lwv(sg + j) * gp[li[sg + j]];
                rm[i] = rmeas;
                // 'scaleerrors' (prepareRecords, src/mppar.f90:312-316): a separate instantiation,
                // since even this product slows the common case down
                if constexpr ((MODE & 2) != 0) re[i] = lwv(sg) * (ng > 0 ? o.esc[0] : o.esc[1]);
                else re[i] = lwv(sg);
                lc[i] = 0.0;
                for (int j = 1; j <= nl; j++) nalc_t = max(nalc_t, li[sl + j]);
            }
            nalc_t = wave_max(nalc_t);
            if (lane == 0) atomicMax(&sh[0], nalc_t);
        }
        __syncthreads();
        const int nalc = sh[0];
        if (nalc <= 0 || nalc > P.nalc_max) {
            if (tid == 0) { out.status[r] = nalc <= 0 ? 7 : 9; out.chi2[r] = 0.0; out.ndf[r] = 0; }
            zero_update(R, M, o, r, tid);
            __syncthreads();
            continue;
        }
        const int npk = (nalc * nalc + nalc) / 2;
        // dense local derivatives (row per equation) and the mask of the local parameters present
        const bool ldense = (int64_t)neq * nalc <= (int64_t)P.max_var * P.nalc_max;
        if (ldense)
            for (int i = tid; i < neq; i += RT) {
                const int sl = les[i];
                const int nl = len[i] & 0xffff;
                double *row = LD + (size_t)i * nalc;
                for (int k = 0; k < nalc; k++) row[k] = 0.0;
                unsigned long long msk = 0ull;
                for (int j = 1; j <= nl; j++) {
                    const int ij = li[sl + j];
                    if (ij >= 1) { row[ij - 1] = lwv(sl + j); msk |= 1ull << (ij - 1); }
                }
                lmask[i] = msk;
            }
        __syncthreads();
        // ---- local fit, chi^2 cuts and the final weights: wavefront 0 ----
        if (wv == 0) {
            // outlier-suppression iterations (3759-4094)
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
                WSYNC();
                // local normal equations, element-parallel, equations in order (3781-3836); systems
                // of up to 9 parameters keep the elements in registers for the inversion
                const bool regs5 = nalc <= 5;
                const bool regfit = !regs5 && npk + nalc <= WAVE;
                double acc_r = 0.0;
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
                    if (ldense) {
                        // equations without both parameters add an exact zero; the loads of eight
                        // equations are issued together
                        const unsigned long long need = (1ull << (J - 1)) | (K > 0 ? 1ull << (K - 1) : 0ull);
                        const int jo = J - 1, ko = (K > 0 ? K : J) - 1;
                        for (int i0 = 0; i0 < neq; i0 += 8) {
                            double tv[8];
                            bool tm[8];
#pragma unroll
                            for (int u = 0; u < 8; u++) {
                                const int i = i0 + u;
                                tm[u] = false;
                                tv[u] = 0.0;
                                if (i < neq) {
                                    const double dj = LD[i * nalc + jo], dk = LD[i * nalc + ko];
                                    const double x = K == 0 ? rm[i] : dj, y = K == 0 ? dj : dk;
                                    tm[u] = (lmask[i] & need) == need;
                                    tv[u] = lw[i] * x * y;
                                }
                            }
#pragma unroll
                            for (int u = 0; u < 8; u++) acc = acc + (tm[u] ? tv[u] : 0.0);
                        }
                    } else {
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
                    }
                    if (regfit) acc_r = acc;
                    else if (K == 0) bl[J - 1] = acc;
                    else cl[t] = acc;
                }
                if (regs5) {
                    WSYNC();
                    switch (nalc) {
                    case 1: nrank = gj_regs<1>(cl, bl, lane); break;
                    case 2: nrank = gj_regs<2>(cl, bl, lane); break;
                    case 3: nrank = gj_regs<3>(cl, bl, lane); break;
                    case 4: nrank = gj_regs<4>(cl, bl, lane); break;
                    default: nrank = gj_regs<5>(cl, bl, lane); break;
                    }
                    WSYNC();
                } else if (regfit) {
                    nrank = gj_lanes(acc_r, npk, nalc, lane, cl, bl);
                    WSYNC();
                } else {
                    WSYNC();
                    // ---- sqminv (src/mpnum.f90:97-213), element-parallel Gauss-Jordan ----
                    const double eps = 16.0 * DBL_EPSILON;
                    for (int i = lane; i < nalc; i += WAVE) {
                        dg[i] = fabs(cl[pk(i + 1, i + 1)]);
                        nxt[i + 1] = i + 2;
                    }
                    WSYNC();
                    if (lane == 0) { nxt[nalc] = -1; sh[1] = 1; /* next0 */ sh[2] = 0; /* nrank */ }
                    WSYNC();
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
                        WSYNC();
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
                            WSYNC();
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
                        WSYNC();
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
                        WSYNC();
                    }
                    for (int t = lane; t < npk; t += WAVE) cl[t] = -cl[t];
                    nrank = sh[2];
                    WSYNC();
                }
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
                    if constexpr ((MODE & 1) != 0) {
                        // residual monitoring, first local-fit iteration (3938-3980): variance of
                        // the fitted value g^T C g from the local inverse (element (j, k) addressed
                        // as pede does, by the order of the local words), pull, and the bin of the
                        // normalized residual pull * err / <err> (or of the pull) in the group of the
                        // measurement's first global parameter; bin width 0.1 as REAL(4) widened
                        if (iter == 1 && (len[i] >> 16) > 0) {
                            double dvar = 0.0;
                            for (int j = 1; j <= nl; j++) {
                                const int ij = li[sl + j];
                                const int jk = (ij * ij - ij) / 2;
                                const double dj = lwv(sl + j);
                                for (int k = 1; k < j; k++)
                                    dvar = dvar + cl[jk + li[sl + k] - 1] * dj * lwv(sl + k) * 2.0;
                                dvar = dvar + cl[jk + ij - 1] * dj * dj;
                            }
                            if (0.999999 / wght > dvar) {
                                const double pull = rmeas / sqrt(1.0 / wght - dvar);
                                const int t = R.gl_tix[g0 + li[sl + nl + 2]];
                                const int grp = mon.grp[t];
                                if (grp >= 0) {
                                    const double bsz = (double)0.1f;
                                    const double xv = (mon.mode == 1 ? pull * rerr / mon.res[grp] / bsz : pull / bsz) +
                                                      0.5 * (double)mon.bins;
                                    if (xv == xv) {
                                        const int bin = xv >= (double)mon.bins ? mon.bins : (xv < 1.0 ? 1 : (int)xv);
                                        atomicAdd(&mon.hist[(int64_t)grp * mon.bins + bin - 1], 1u);
                                    }
                                }
                            }
                        }
                    }
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
                WSYNC();
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
            if (status == 0) {
                // fourth loop, first part: chi^2 of the accepted fit and the weights of the global
                // update (4162-4183)
                double s_chi = 0.0;
                const bool down = (o.nloopn != 1 && o.lhuber != 0);
                for (int i = lane; i < neq; i += WAVE) {
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
                }
                summ = wave_sum(s_chi);
            }
            if (lane == 0) {
                sh[8] = status;
                sh[9] = ndf;
                sh[10] = ndf_add;
                dsh[0] = summ;
                dsh[1] = chi_add;
            }
        }
        __syncthreads();
        {
            const int status = sh[8];
            if (status != 0) {
                if (tid == 0) { out.status[r] = (int8_t)status; out.chi2[r] = dsh[1] * dw1; out.ndf[r] = sh[10]; }
                zero_update(R, M, o, r, tid);
                __syncthreads();
                continue;
            }
        }
        const int ndf = sh[9];
        const double summ = dsh[0];
        // ---- fourth loop: vector, counters, g g^T, G (4162-4229) ----
        // Every record variable collects the equations it occurs in (up to QE, ascending): the
        // vector, the counters and G get one update per variable.  Full storage: g g^T per
        // equation and the Schur complement per element with FP64 atomics.  Group-block
        // storage: the record's complete update (g g^T of the common equations minus the Schur
        // complement) is written block by block into the update buffer, to be summed per matrix
        // block by k_block_gather.  A variable in more than QE equations, or twice in one
        // equation, makes the record take the per-equation path for the vector and G (and, with
        // group-block storage, an equation scan for g g^T).
        const bool twophase = M.sparse != 0;   // the host always provides the update buffer then
        if (o.icalcm == 1)
            for (int t = tid; t < nalg * nalc; t += RT) G[t] = 0.0;
        for (int t = tid; t < nalg; t += RT) qn[t] = 0;
        if (tid == 0) sh[5] = 0;
        if (o.icalcm == 1 && twophase && wv == 0) {
            // segments: runs of record variables in the same parameter group
            int nseg = 0;
            for (int base = 0; base < nalg; base += WAVE) {
                int a = base + lane;
                int g = a < nalg ? ggl[a] : -1;
                int gprev = (a > 0 && a < nalg) ? ggl[a - 1] : -2;
                bool st = (a < nalg) && (a == 0 || g != gprev);
                unsigned long long m = __ballot(st);
                int s = nseg + __popcll(m & ((2ull << lane) - 1ull)) - 1;
                if (a < nalg && st) segst[s] = a;
                if (a < nalg) segof[a] = s;
                nseg += __popcll(m);
            }
            if (lane == 0) { segst[nseg] = nalg; sh[6] = nseg; }
        }
        __syncthreads();
        for (int i = tid; i < neq; i += RT) {
            const int sl = les[i];
            const uint32_t en = len[i];
            int nl = en & 0xffff, ng = en >> 16;
            const int sg = sl + nl + 1;
            const double wght = lw[i];
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
        __syncthreads();
        for (int a = tid; a < nalg; a += RT) {
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
        // the two-phase build stores each variable's vector value in the update buffer (formed
        // again where the row descriptors are written), for the block gather to sum in
        // processing order; otherwise, and on the per-equation path, it is added to the vector
        // directly
        const bool vgather = o.icalcm == 1 && twophase;
        if (!perq) {
            for (int a = tid; a < nalg; a += RT) {
                const int n = qn[a];
                double v = 0.0;
                for (int x = 0; x < n; x++) {
                    const int i = qe[a * QE + x];
                    v = v + dw1 * lw[i] * rm[i] * qg[a * QE + x];
                }
                if (!vgather) atomicAdd(&M.vec[gvl[a] - 1], v);
                atomicAdd(&M.cnt[gvl[a] - 1], (unsigned long long)n);
            }
        } else {
            for (int i = tid; i < neq; i += RT) {
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
            // G = sum over equations of sqrt(dw) w g l^T, equations in order (4206-4219): each
            // variable's row sums over its (ascending) equations; records off the list path add
            // equation by equation
            if (!perq) {
                for (int a = tid; a < nalg; a += RT) {
                    double *Ga = G + (size_t)a * nalc;
                    for (int x = 0; x < qn[a]; x++) {
                        const int i = qe[a * QE + x];
                        const double gj = qg[a * QE + x];
                        const double wght = lw[i];
                        const int sl = les[i];
                        const int nl = len[i] & 0xffff;
                        for (int k = 1; k <= nl; k++) {
                            const int ik = li[sl + k];
                            Ga[ik - 1] = Ga[ik - 1] + dw2 * wght * gj * lwv(sl + k);
                        }
                    }
                }
            } else {
                for (int i = 0; i < neq; i++) {
                    const int sl = les[i];
                    const uint32_t en = len[i];
                    int nl = en & 0xffff, ng = en >> 16;
                    const int sg = sl + nl + 1;
                    double wght = lw[i];
                    for (int t = tid; t < ng * nl; t += RT) {
                        int j = t / nl + 1, k = t % nl + 1;
                        int pj = li[sg + j];
                        int a = gpos[pj];
                        if (a < 0) continue;
                        const int ik = li[sl + k];
                        // word pairs of one equation that share the element (a label or a local
                        // index repeated in the equation) are summed by the thread of the first
                        // such pair, in the reference's (j, k) order
                        bool own = true, rep = false;
                        for (int x = 1; x <= ng; x++)
                            if (x != j && li[sg + x] == pj) { rep = true; own = own && x > j; }
                        for (int x = 1; x <= nl; x++)
                            if (x != k && li[sl + x] == ik) { rep = true; own = own && x > k; }
                        if (!own) continue;
                        double *ge = &G[a * nalc + ik - 1];
                        if (!rep) {
                            *ge = *ge + dw2 * wght * lwv(sg + j) * lwv(sl + k);
                            continue;
                        }
                        double v = *ge;
                        for (int j2 = j; j2 <= ng; j2++) {
                            if (li[sg + j2] != pj) continue;
                            for (int k2 = k; k2 <= nl; k2++)
                                if (li[sl + k2] == ik) v = v + dw2 * wght * lwv(sg + j2) * lwv(sl + k2);
                        }
                        *ge = v;
                    }
                    __syncthreads();
                }
            }
            __syncthreads();
            // Gh = G Gamma^-1 (row-wise, l ascending as dbavat), one element per thread
            for (int t = tid; t < nalg * nalc; t += RT) {
                const int a = t / nalc, k = t - a * nalc + 1;
                double c = 0.0;
                for (int l = 1; l <= nalc; l++) c = c + G[a * nalc + l - 1] * cl[l >= k ? pk(l, k) : pk(k, l)];
                Gh[t] = c;
            }
            __syncthreads();
            if (!twophase) {
                // full storage: element (a>=b) -= sum_k Gh[a,k] G[b,k]; packed (a, b) enumeration,
                // thread starts at element `tid`, then steps by RT
                const int ngg = (nalg * nalg + nalg) / 2;
                int a = (int)((sqrt(8.0 * tid + 1.0) - 1.0) * 0.5);
                while ((a * a + a) / 2 > tid) a--;
                while (((a + 1) * (a + 1) + (a + 1)) / 2 <= tid) a++;
                int b = tid - (a * a + a) / 2;
                for (int t = tid; t < ngg; t += RT) {
                    double c = 0.0;
                    for (int k = 0; k < nalc; k++) c = c + Gh[a * nalc + k] * G[b * nalc + k];
                    atomicAdd(&M.mat[pk64(gvl[a], gvl[b])], -c);
                    b += RT;
                    while (b > a) { b -= a + 1; a++; }
                }
            } else {
                // Group-block storage.  Layout of a record's update: for every segment sa (in
                // order) the columns b = 0 .. end of sa (all record variables up to and including
                // segment sa), each column holding the na values of the segment's rows
                // (value (ra, b) at b * na + ra), then one column with the rows' vector values
                // (padded to a multiple of four values, which keeps every region 32-byte aligned
                // when the columns are); the diagonal block is kept square, its upper part is
                // zero.  The block of segment pair (sa, sb) is thus contiguous
                // (columns of sb), which the block gather reads.  The values
                // C1(a,b) - sum_k Gh[a,k] G[b,k] (b <= a) of different segments come from 16 x 16
                // tiles of the record's lower triangle, one tile per wavefront step, the products
                // on the matrix cores (v_mfma_f64_16x16x4f64, nalc in steps of four); the
                // diagonal blocks are formed element by element; every value goes straight to its
                // place in the update buffer.
                const int nseg = sh[6];
                double *dst = R.cbuf + R.rec_cb[r];
                // equations touching each segment (bit i for equation i < 64; bit 63 also stands
                // for all equations >= 63): pairs without a common equation have no g g^T part
                // first update value of each segment (exclusive prefix of ncolumns * na); whether
                // two segments share an equation at all (else no cross-segment g g^T)
                if (wv == 0) {
                    unsigned long long mor = 0ull;
                    int bits = 0;
                    for (int q = lane; q < nseg; q += WAVE) {
                        unsigned long long m = 0ull;
                        for (int a = segst[q]; a < segst[q + 1]; a++) {
                            if (perq) { m = ~0ull; break; }
                            for (int x = 0; x < qn[a]; x++) m |= 1ull << min(qe[a * QE + x], 63);
                        }
                        segm[q] = m;
                        mor |= m;
                        bits += __popcll(m);
                    }
                    for (int o = 32; o > 0; o >>= 1) {
                        mor |= __shfl_xor(mor, o, WAVE);
                        bits += __shfl_xor(bits, o, WAVE);
                    }
                    if (lane == 0) sh[11] = (perq || bits > __popcll(mor)) ? 1 : 0;
                    int run = 0;
                    for (int base = 0; base < nseg; base += WAVE) {
                        const int q = base + lane;
                        const int na = q < nseg ? segst[q + 1] - segst[q] : 0;
                        int v = q < nseg ? segst[q + 1] * na + ((na + 3) & ~3) : 0;
                        int x = v;
                        for (int d = 1; d < WAVE; d <<= 1) {
                            const int y = __shfl_up(x, d, WAVE);
                            if (lane >= d) x += y;
                        }
                        if (q < nseg) segro[q] = run + x - v;
                        run += __shfl(x, WAVE - 1, WAVE);
                    }
                }
                __syncthreads();
                // row descriptors (update-buffer offset of the row, column stride, segment, first
                // row of the segment); the upper part of the diagonal blocks is zero
                const bool anycommon = sh[11] != 0;
                for (int a = tid; a < nalg; a += RT) {
                    const int q = segof[a], a0 = segst[q], na = segst[q + 1] - a0;
                    const int base = segro[q] + (a - a0);
                    rdesc[a] = make_int4(base, na, q, a0);
                    vsegm[a] = segm[q];
                    for (int b = a + 1; b < a0 + na; b++) dst[base + (int64_t)b * na] = 0.0;
                    {
                        // the segment's vector column: this variable's value (the per-equation
                        // path added its values directly and leaves a zero; a per-variable loop
                        // for it here costs the whole kernel registers)
                        double v = 0.0;
                        if (!perq)
                            for (int x = 0; x < qn[a]; x++) {
                                const int i = qe[a * QE + x];
                                v = v + dw1 * lw[i] * rm[i] * qg[a * QE + x];
                            }
                        dst[base + (int64_t)(a0 + na) * na] = v;
                    }
                }
                __syncthreads();
                // diagonal blocks (a, b <= a in one segment), one element per thread: g g^T of the
                // common equations minus sum_k Gh[a,k] G[b,k] (k ascending)
                const int gw = P.max_grp;
                for (int t = tid; t < nalg * gw; t += RT) {
                    const int a = t / gw, j = t - a * gw;
                    const int b = segst[segof[a]] + j;
                    if (b > a) continue;
                    double d = 0.0;
                    for (int k = 0; k < nalc; k++) d = d + Gh[a * nalc + k] * G[b * nalc + k];
                    const double c = dc1<T>(a, b, perq, QE, qn, qe, qg, lw, dw1, les, len, li, lv, neq, gpos);
                    const int4 da = rdesc[a];
                    dst[da.x + (int64_t)b * da.y] = c - d;
                }
                const int nt = (nalg + 15) >> 4;
                const int lr = lane & 15, lk = lane >> 4;
                // the product is formed transposed (rows of G as the first operand): lane l holds
                // row a = 16 ti + l%16 and four columns b, so the sixteen lanes of one store write
                // a column run of consecutive rows, contiguous within a segment
                int ti = 0, tj = wv;                  // tile tt = ti (ti + 1) / 2 + tj, tj <= ti
                while (tj > ti) { tj -= ti + 1; ti++; }
                for (int tt = wv; tt < nt * (nt + 1) / 2; tt += RW) {
                    const int a = 16 * ti + lr, rb = 16 * tj + lr;
                    const bool arow = a < nalg;
                    const int4 da = rdesc[arow ? a : nalg - 1];
                    d4 acc = {0.0, 0.0, 0.0, 0.0};
                    for (int kk = 0; kk < nalc; kk += 4) {
                        const int k = kk + lk;
                        const double gb = (rb < nalg && k < nalc) ? G[rb * nalc + k] : 0.0;
                        const double ha = (arow && k < nalc) ? Gh[a * nalc + k] : 0.0;
                        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(gb, ha, acc, 0, 0, 0);
                    }
#pragma unroll
                    for (int e = 0; e < 4; e++) {
                        const int b = 16 * tj + lk + 4 * e;   // result element e: (row a, column b)
                        const int bc = b < nalg ? b : nalg - 1;
                        if (arow && b < da.w) {               // columns before a's segment (diagonal blocks: above)
                            double c = 0.0;
                            if (anycommon && (vsegm[arow ? a : nalg - 1] & vsegm[bc]) != 0ull)
                                c = dc1<T>(a, b, perq, QE, qn, qe, qg, lw, dw1, les, len, li, lv, neq, gpos);
                            dst[da.x + (int64_t)b * da.y] = c - acc[e];
                        }
                    }
                    tj += RW;
                    while (tj > ti) { tj -= ti + 1; ti++; }
                }
            }
        }
        if (tid == 0) { out.status[r] = 0; out.chi2[r] = summ * dw1; out.ndf[r] = ndf; }
        __syncthreads();
    }
}

// Block gather of the two-phase matrix build: every stored block adds the values that the
// records of the batch wrote for it (in processing order).  One wavefront per block; lane e
// handles element e of the r x c block (row-major in the matrix, column-major in the records'
// updates); diagonal blocks use their lower triangle.
__global__ __launch_bounds__(256) void k_block_gather(int64_t nblock, const int64_t *boff, const int32_t *brg,
                                                      const int32_t *bcol, const int32_t *gfirst, const int64_t *ptr,
                                                      const int64_t *src, const double *cbuf, double *mat, double *vec)
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
    // a diagonal block also sums the vector column that follows it in the records' updates
    // (elements nr*nc .. nr*nc + nr - 1, column nc)
    const int ne = nr * nc + (diag ? nr : 0);
    for (int e0 = 0; e0 < ne; e0 += 64) {
        const int e = e0 + lane;
        const bool act = e < ne, inm = e < nr * nc;
        const int rr = !act ? 0 : inm ? e / nc : e - nr * nc;
        const int cc = !act ? 0 : inm ? e - rr * nc : nc;
        const int se = cc * nr + rr;        // records store their blocks column by column
        // the contributions' offsets, 64 at a time, one per lane (a coalesced read), then eight
        // independent partial sums over them, each offset broadcast from its lane
        double acc[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        for (int64_t c0 = i0; c0 < i1; c0 += 64) {
            const int n = (int)min((int64_t)64, i1 - c0);
            const int64_t mine = src[c0 + (lane < n ? lane : n - 1)];
            const int lo = (int)mine, hi = (int)(mine >> 32);
            int q = 0;
            for (; q + 8 <= n; q += 8) {
                double v[8];
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    const int64_t o = ((int64_t)__builtin_amdgcn_readlane(hi, q + u) << 32) |
                                      (uint32_t)__builtin_amdgcn_readlane(lo, q + u);
                    v[u] = cbuf[o + se];
                }
#pragma unroll
                for (int u = 0; u < 8; u++) acc[u] += v[u];
            }
            for (; q < n; q++) {
                const int64_t o = ((int64_t)__builtin_amdgcn_readlane(hi, q) << 32) | (uint32_t)__builtin_amdgcn_readlane(lo, q);
                acc[0] += cbuf[o + se];
            }
        }
        const double sum = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
        if (act && inm && !(diag && cc > rr)) B[e] += sum;
        if (act && !inm) vec[gfirst[ga] - 1 + rr] += sum;
    }
}

// host launcher
template <typename T, int MODE>
static hipError_t launch_records_t(const DevRecords &R, const DevModel &M, const PassOpts &o, const RecOut &out,
                                   const LdsPlan &P, int64_t rec_begin, int64_t rec_end, int grid, hipStream_t s,
                                   const MonCtx &mon)
{
    if (P.bytes > 65536) {
        hipError_t e = hipFuncSetAttribute((const void *)k_records<T, MODE>, hipFuncAttributeMaxDynamicSharedMemorySize, P.bytes);
        if (e != hipSuccess) return e;
    }
    // as many workgroups as fit on the device at once; each loops over the records
    int per_cu = 0, dev = 0, ncu = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_records<T, MODE>, RT, P.bytes);
    if (e != hipSuccess) return e;
    if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
    if ((e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return e;
    grid = std::max(1, std::min<int>(grid, std::max(1, per_cu) * std::max(1, ncu)));
    if (std::getenv("MPGPU_VERBOSE"))
        std::fprintf(stderr, "k_records: %d threads x %d workgroups, LDS %d bytes, %d workgroups per CU\n", RT, grid,
                     P.bytes, per_cu);
    hipLaunchKernelGGL((k_records<T, MODE>), dim3(grid), dim3(RT), P.bytes, s, R, M, o, out, P, rec_begin, rec_end, mon);
    return hipGetLastError();
}

hipError_t launch_records(const DevRecords &R, const DevModel &M, const PassOpts &o, const RecOut &out,
                          const LdsPlan &P, int64_t rec_begin, int64_t rec_end, int grid, hipStream_t s,
                          const MonCtx &mon)
{
    if (rec_end <= rec_begin) return hipSuccess;
    // instantiation: bit 0 residual monitoring, bit 1 scaled measurement errors
    const int mode = (mon.mode != 0 && mon.hist != nullptr ? 1 : 0) | (o.esc[0] != 1.0 || o.esc[1] != 1.0 ? 2 : 0);
#define MPGPU_REC_CASE(T_, M_) \
    case M_: return launch_records_t<T_, M_>(R, M, o, out, P, rec_begin, rec_end, grid, s, mon);
    if (R.vf) {
        switch (mode) { MPGPU_REC_CASE(float, 0) MPGPU_REC_CASE(float, 1) MPGPU_REC_CASE(float, 2) MPGPU_REC_CASE(float, 3) }
    } else {
        switch (mode) { MPGPU_REC_CASE(double, 0) MPGPU_REC_CASE(double, 1) MPGPU_REC_CASE(double, 2) MPGPU_REC_CASE(double, 3) }
    }
#undef MPGPU_REC_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_block_gather(int64_t nblock, const int64_t *boff, const int32_t *brg, const int32_t *bcol,
                               const int32_t *gfirst, const int64_t *ptr, const int64_t *src, const double *cbuf,
                               double *mat, double *vec, hipStream_t s)
{
    if (nblock <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_block_gather, dim3((unsigned)((nblock + 3) / 4)), dim3(256), 0, s, nblock, boff, brg, bcol, gfirst,
                       ptr, src, cbuf, mat, vec);
    return hipGetLastError();
}

int lds_bytes(LdsPlan &P)
{
    // the larger payload type sizes the word stage
    P.bytes = rec_lds(P, (int)sizeof(double)).bytes;
    return P.bytes;
}

} // namespace mpgpu
