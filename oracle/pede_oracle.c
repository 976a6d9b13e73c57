/*
 * pede_oracle.c -- CPU restatement of Millepede-II pede's global normal-matrix build and
 * global solve.  TEST INFRASTRUCTURE ONLY: this file is the checker the GPU product is
 * compared against.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it; the product library never links or calls it.
 *
 * Every function names the reference routine (path:line, relative to the millepede-ii source
 * tree) whose arithmetic it restates.  Matrices use pede's symmetric packed storage:
 * element (i,j), i>=j, 1-based, lives at (i*i-i)/2+j (src/pede.f90:4204, src/mpnum.f90:97).
 *
 * Pinning (how each function is checked, see tests/test_oracle_*.py):
 *   orc_sqminv, orc_sqminl, orc_chdec2/chslv2, orc_dbavat, orc_chindl, orc_lltdec/fwd/bwd,
 *   orc_equdecs/equslvs, orc_precons/presols, orc_sqmibb (with the dbc* band routines),
 *   orc_minres, orc_minresqlp
 *       -> bit-for-bit (or 1e-13) against the reference Fortran routines compiled from
 *          src/mpnum.f90, src/Dbandmatrix.f90, src/minresModule.f90 and src/minresqlpModule.f90
 *          into oracle/_ref, plus numpy closed forms.
 *   orc_qldecb, orc_qlmlq, orc_qlbsub, orc_qlpssq (constraint elimination)
 *       -> bit-for-bit against the reference's src/mpqldec.f90 in oracle/_ref on layouts built
 *          by pede's grouping rules, plus closed forms (Q orthogonal, C^T = Q [0; L],
 *          elimination = KKT solution); orc_qldecb's aligned start of L is pinned by the
 *          closed form only (the reference's start differs, see orc_qldecb).
 *   orc_loopn (record loop)
 *       -> brute-force joint least-squares solve (globals + every record's locals) on tiny
 *          inputs (Schur identity, doc/draftManual.md:291-330), per-record explicit inverse.
 *   orc_read_mille -> round trip against the Python/C writers and the layout of
 *          src/mille.f90:208-219 and python/readMilleBinary.py:128-216.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>

#define IDX(i, j) ((((int64_t)(i)) * (i) - (i)) / 2 + (j)) /* 1-based packed (i>=j) */

/* ------------------------------------------------------------------------------------------ */
/* chindl: chi^2/ndf cut values, REAL(4) table (src/mpnum.f90:2078-2115)                       */
/* ------------------------------------------------------------------------------------------ */
static const float chindl_sn[3] = {0.47523f, 1.690140f, 2.782170f};
static const float chindl_table[3][30] = {
    {1.0000f, 1.1479f, 1.1753f, 1.1798f, 1.1775f, 1.1730f, 1.1680f, 1.1630f, 1.1581f, 1.1536f,
     1.1493f, 1.1454f, 1.1417f, 1.1383f, 1.1351f, 1.1321f, 1.1293f, 1.1266f, 1.1242f, 1.1218f,
     1.1196f, 1.1175f, 1.1155f, 1.1136f, 1.1119f, 1.1101f, 1.1085f, 1.1070f, 1.1055f, 1.1040f},
    {4.0000f, 3.0900f, 2.6750f, 2.4290f, 2.2628f, 2.1415f, 2.0481f, 1.9736f, 1.9124f, 1.8610f,
     1.8171f, 1.7791f, 1.7457f, 1.7161f, 1.6897f, 1.6658f, 1.6442f, 1.6246f, 1.6065f, 1.5899f,
     1.5745f, 1.5603f, 1.5470f, 1.5346f, 1.5230f, 1.5120f, 1.5017f, 1.4920f, 1.4829f, 1.4742f},
    {9.0000f, 5.9146f, 4.7184f, 4.0628f, 3.6410f, 3.3436f, 3.1209f, 2.9468f, 2.8063f, 2.6902f,
     2.5922f, 2.5082f, 2.4352f, 2.3711f, 2.3143f, 2.2635f, 2.2178f, 2.1764f, 2.1386f, 2.1040f,
     2.0722f, 2.0428f, 2.0155f, 1.9901f, 1.9665f, 1.9443f, 1.9235f, 1.9040f, 1.8855f, 1.8681f}};

float orc_chindl(int n, int nd)
{
    if (nd < 1) return 0.0f;
    int m = n < 1 ? 1 : (n > 3 ? 3 : n);
    if (nd <= 30) return chindl_table[m - 1][nd - 1];
    float t = chindl_sn[m - 1] + sqrtf((float)(nd + nd - 1));
    return t * t / (float)(nd + nd);
}

/* ------------------------------------------------------------------------------------------ */
/* sqminv: symmetric Gauss-Jordan inversion with diagonal pivoting (src/mpnum.f90:97-213)      */
/* v packed (n(n+1)/2), b(n) -> inverse and solution in place                                 */
/* ------------------------------------------------------------------------------------------ */
void orc_sqminv(double *v, double *b, int n, int *nrank)
{
    double *diag = (double *)malloc(sizeof(double) * (n + 1));
    int *next = (int *)malloc(sizeof(int) * (n + 1));
    double eps = 16.0 * DBL_EPSILON;
    int next0 = 1, l = 1;
    v--; b--; /* 1-based */
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
    for (int ij = 1; ij <= (n * n + n) / 2; ij++) v[ij] = -v[ij];
    free(diag);
    free(next);
}

/* ------------------------------------------------------------------------------------------ */
/* sqminl: Gauss-Jordan for large matrices, relative pivot eps 1e-10 (src/mpnum.f90:230-355)   */
/* ------------------------------------------------------------------------------------------ */
void orc_sqminl(double *v, double *b, int n, int *nrank)
{
    double *diag = (double *)malloc(sizeof(double) * (n + 1));
    double *vk = (double *)malloc(sizeof(double) * (n + 1));
    int *next = (int *)malloc(sizeof(int) * (n + 1));
    const double eps = 1.0e-10;
    int next0 = 1, l = 1;
    v--; b--;
    for (int i = 1; i <= n; i++) {
        int64_t i8 = i;
        next[i] = i + 1;
        diag[i] = fabs(v[(i8 * i8 + i8) / 2]);
    }
    next[n] = -1;
    *nrank = 0;
    for (int i = 1; i <= n; i++) {
        int k = 0;
        double vkk = 0.0;
        int j = next0, last = 0;
        while (j > 0) {
            int64_t j8 = j, jj = (j8 * j8 + j8) / 2;
            double m = fabs(vkk) > eps * diag[j] ? fabs(vkk) : eps * diag[j];
            if (fabs(v[jj]) > m) { vkk = v[jj]; k = j; l = last; }
            last = j;
            j = next[last];
        }
        if (k != 0) {
            int64_t k8 = k, kk = (k8 * k8 + k8) / 2, kkmk = kk - k8;
            if (l == 0) next0 = next[k]; else next[l] = next[k];
            next[k] = 0;
            (*nrank)++;
            vkk = 1.0 / vkk;
            v[kk] = -vkk;
            b[k] = b[k] * vkk;
            int64_t jk = kkmk;
            for (int jx = 1; jx <= n; jx++) {
                if (jx == k) { jk = kk; vk[jx] = 0.0; }
                else {
                    if (jx < k) jk = jk + 1; else jk = jk + jx - 1;
                    v[jk] = v[jk] * vkk;
                    vk[jx] = v[jk];
                }
            }
            for (int jx = n; jx >= 1; jx--) {
                if (jx == k) continue;
                int64_t j8 = jx, jl = j8 * (j8 - 1) / 2;
                double vjk = vk[jx] / vkk;
                b[jx] = b[jx] - b[k] * vjk;
                for (int lx = 1; lx <= jx; lx++) v[jl + lx] = v[jl + lx] - vk[lx] * vjk;
            }
        } else {
            for (int kx = 1; kx <= n; kx++) {
                int64_t k8 = kx, kk = (k8 * k8 - k8) / 2;
                if (next[kx] != 0) {
                    b[kx] = 0.0;
                    for (int jx = 1; jx <= kx; jx++)
                        if (next[jx] != 0) v[kk + jx] = 0.0;
                }
            }
            break;
        }
    }
    int64_t nt = ((int64_t)n * (n + 1)) / 2;
    for (int64_t jj = 1; jj <= nt; jj++) v[jj] = -v[jj];
    free(diag);
    free(vk);
    free(next);
}

/* ------------------------------------------------------------------------------------------ */
/* chdec2 / chslv2: root-free LDL^T from the last row (src/mpnum.f90:891-987)                  */
/* ------------------------------------------------------------------------------------------ */
void orc_chdec2(double *g, int n, int *nrank, double *evmax, double *evmin)
{
    g--;
    *nrank = 0;
    *evmax = 0.0;
    *evmin = 0.0;
    int64_t ii = ((int64_t)n * (n + 1)) / 2;
    for (int i = n; i >= 1; i--) {
        if (g[ii] > 0.0) {
            (*nrank)++;
            if (*nrank == 1) { *evmax = g[ii]; *evmin = g[ii]; }
            else {
                if (g[ii] > *evmax) *evmax = g[ii];
                if (g[ii] < *evmin) *evmin = g[ii];
            }
            g[ii] = 1.0 / g[ii];
        }
        ii = ii - i;
        for (int j = 1; j <= i - 1; j++) {
            double ratio = g[ii + j] * g[ii + i];
            if (ratio == 0.0) continue;
            int64_t jj = ((int64_t)(j - 1) * j) / 2;
            for (int k = 1; k <= j; k++) g[jj + k] = g[jj + k] - g[ii + k] * ratio;
        }
        for (int k = 1; k <= i - 1; k++) g[ii + k] = g[ii + k] * g[ii + i];
    }
}

void orc_chslv2(const double *g, double *x, int n)
{
    g--; x--;
    int64_t ii = ((int64_t)n * (n + 1)) / 2;
    for (int i = n; i >= 1; i--) {
        double dsum = x[i];
        int64_t kk = ii;
        for (int k = i + 1; k <= n; k++) { dsum = dsum - g[kk + i] * x[k]; kk = kk + k; }
        x[i] = dsum;
        ii = ii - i;
    }
    for (int i = 1; i <= n; i++) {
        double dsum = x[i] * g[ii + i];
        for (int k = 1; k <= i - 1; k++) dsum = dsum - g[k + ii] * x[k];
        x[i] = dsum;
        ii = ii + i;
    }
}

/* ------------------------------------------------------------------------------------------ */
/* dbavat (iopt=1): W += A V A^T, A (m x n) row-major = localGlobalMatrix (src/mpnum.f90:1389)  */
/* ------------------------------------------------------------------------------------------ */
void orc_dbavat(const double *v, const double *a, double *w, int n, int m)
{
    v--; a--; w--;
    int il = -n, ijs = 0;
    for (int i = 1; i <= m; i++) {
        ijs = ijs + i - 1;
        il = il + n;
        int lkl = 0;
        for (int k = 1; k <= n; k++) {
            double cik = 0.0;
            lkl = lkl + k - 1;
            int lk = lkl;
            for (int l = 1; l <= k; l++) { lk = lk + 1; cik = cik + a[il + l] * v[lk]; }
            for (int l = k + 1; l <= n; l++) { lk = lk + l - 1; cik = cik + a[il + l] * v[lk]; }
            int jk = k, ij = ijs;
            for (int j = 1; j <= i; j++) { ij = ij + 1; w[ij] = w[ij] + cik * a[jk]; jk = jk + n; }
        }
    }
}

/* ------------------------------------------------------------------------------------------ */
/* Variable-band Cholesky used as MINRES preconditioner (src/mpnum.f90:2155-2336)              */
/* india(n) 1-based pointers to the diagonal elements                                         */
/* ------------------------------------------------------------------------------------------ */
void orc_lltdec(int n, double *c, const int *india, int *nrkd, int iopt)
{
    c--; india--;
    double eps = 16.0 * DBL_EPSILON;
    *nrkd = 0;
    double diag = 0.0;
    if (c[india[1]] > 0.0) c[india[1]] = 1.0 / sqrt(c[india[1]]);
    else { c[india[1]] = 0.0; *nrkd = -1; }
    for (int k = 2; k <= n; k++) {
        int mk = k - india[k] + india[k - 1] + 1;
        for (int j = mk; j <= k; j++) {
            int mj = 1;
            if (j > 1) mj = j - india[j] + india[j - 1] + 1;
            int kj = india[k] - k + j;
            diag = c[india[j]];
            for (int i = (mj > mk ? mj : mk); i <= j - 1; i++)
                c[kj] = c[kj] - c[india[k] - k + i] * c[india[j] - j + i];
            if (j != k) c[kj] = c[kj] * diag;
        }
        if (c[india[k]] > eps * diag) c[india[k]] = 1.0 / sqrt(c[india[k]]);
        else {
            for (int j = mk; j <= k; j++) c[india[k] - k + j] = 0.0;
            if (iopt > 0 && diag > 0.0) c[india[k]] = 1.0 / sqrt(diag);
            else {
                if (*nrkd == 0) *nrkd = -k;
                else { if (*nrkd < 0) *nrkd = 1; (*nrkd)++; }
            }
        }
    }
}

void orc_lltfwd(int n, const double *c, const int *india, double *x)
{
    c--; india--; x--;
    x[1] = x[1] * c[india[1]];
    for (int k = 2; k <= n; k++) {
        for (int j = k - india[k] + india[k - 1] + 1; j <= k - 1; j++) x[k] = x[k] - c[india[k] - k + j] * x[j];
        x[k] = x[k] * c[india[k]];
    }
}

static void orc_lltfwds(int n, const double *c, const int *india, double *x, int i0, int ns)
{
    c--; india--; x--;
    int kmax = ns < n - i0 ? ns : n - i0;
    for (int k = 1; k <= kmax; k++) {
        int l = k + i0;
        if (l > 1) {
            int j0 = k - india[l] + india[l - 1] + 1;
            if (j0 < 1) j0 = 1;
            for (int j = j0; j <= k - 1; j++) x[k] = x[k] - c[india[l] - k + j] * x[j];
        }
        x[k] = x[k] * c[india[l]];
    }
}

void orc_lltbwd(int n, const double *c, const int *india, double *x)
{
    c--; india--; x--;
    for (int k = n; k >= 2; k--) {
        x[k] = x[k] * c[india[k]];
        for (int j = k - india[k] + india[k - 1] + 1; j <= k - 1; j++) x[j] = x[j] - c[india[k] - k + j] * x[k];
    }
    x[1] = x[1] * c[india[1]];
}

/* equdecs: band decomposition + constraint Schur block (src/mpnum.f90:2486-2606).
 * c layout: [band (india(n) words)] [K K^T packed m(m+1)/2] [constraint blocks]
 * india has n+m entries; l(4*b) = (first con, last con, param offset, ncols) per block (1-based). */
void orc_equdecs(int n, int m, int b, int ls, double *c, int *india, const int *l, int *nrkd, int *nrkd2)
{
    *nrkd = 0;
    *nrkd2 = 0;
    orc_lltdec(n, c, india, nrkd, ls);
    if (m > 0) {
        double *C = c - 1;
        const int *L = l - 1;
        int *IND = india - 1;
        int64_t ij = IND[n] + (m * m + m) / 2;
        for (int ib = 1; ib <= b; ib++) {
            int p1 = L[ib * 4 - 3], p2 = L[ib * 4 - 2], i0 = L[ib * 4 - 1], in = L[ib * 4];
            for (int j = p1; j <= p2; j++) { orc_lltfwds(n, c, india, &C[ij + 1], i0, in); ij += in; }
        }
        ij = IND[n] + (m * m + m) / 2;
        for (int ib = 1; ib <= b; ib++) {
            int p1 = L[ib * 4 - 3], p2 = L[ib * 4 - 2], i0 = L[ib * 4 - 1], in = L[ib * 4];
            for (int i = 1; i <= in; i++) {
                ij = ij + 1;
                for (int j = p1; j <= p2; j++) {
                    int64_t jk = IND[n] + (j * j - j) / 2 + p1;
                    for (int k = p1; k <= j; k++) {
                        C[jk] = C[jk] + C[ij + (int64_t)(j - p1) * in] * C[ij + (int64_t)(k - p1) * in];
                        jk++;
                    }
                }
            }
            int64_t jj = ij + (int64_t)(p2 - p1) * in;
            for (int jb = ib + 1; jb <= b; jb++) {
                int q1 = L[jb * 4 - 3], q2 = L[jb * 4 - 2], j0 = L[jb * 4 - 1], jn = L[jb * 4];
                int nc = in + i0 - j0 < jn ? in + i0 - j0 : jn;
                if (nc < 1) break;
                int64_t kk = ij + j0 - i0 - in;
                for (int i = 1; i <= nc; i++) {
                    kk = kk + 1;
                    for (int j = q1; j <= q2; j++) {
                        int64_t jk = IND[n] + (j * j - j) / 2 + p1;
                        for (int k = p1; k <= p2; k++) {
                            C[jk] = C[jk] + C[jj + (int64_t)(j - q1) * jn + i] * C[kk + (int64_t)(k - p1) * in];
                            jk++;
                        }
                    }
                }
                jj = jj + (int64_t)(q2 + 1 - q1) * jn;
            }
            ij = ij + (int64_t)(p2 - p1) * in;
        }
        IND[n + 1] = 1;
        for (int i = 2; i <= m; i++) IND[n + i] = IND[n + i - 1] + (i < m ? i : m);
        orc_lltdec(m, &C[IND[n] + 1], &IND[n + 1], nrkd2, 0);
    }
}

void orc_equslvs(int n, int m, int b, const double *c, const int *india, const int *l, double *x)
{
    orc_lltfwd(n, c, india, x);
    if (m > 0) {
        const double *C = c - 1;
        const int *L = l - 1, *IND = india - 1;
        double *X = x - 1;
        int64_t ij = IND[n] + (m * m + m) / 2;
        for (int ib = 1; ib <= b; ib++) {
            int p1 = L[ib * 4 - 3], p2 = L[ib * 4 - 2], i0 = L[ib * 4 - 1], in = L[ib * 4];
            for (int j = p1; j <= p2; j++)
                for (int i = 1; i <= in; i++) { ij++; X[n + j] = X[n + j] - C[ij] * X[i + i0]; }
        }
        orc_lltfwd(m, &C[IND[n] + 1], &IND[n + 1], &X[n + 1]);
        orc_lltbwd(m, &C[IND[n] + 1], &IND[n + 1], &X[n + 1]);
        for (int i = 1; i <= m; i++) X[n + i] = -X[n + i];
        ij = IND[n] + (m * m + m) / 2;
        for (int ib = 1; ib <= b; ib++) {
            int p1 = L[ib * 4 - 3], p2 = L[ib * 4 - 2], i0 = L[ib * 4 - 1], in = L[ib * 4];
            for (int j = p1; j <= p2; j++)
                for (int i = 1; i <= in; i++) { ij++; X[i + i0] = X[i + i0] - C[ij] * X[n + j]; }
        }
    }
    orc_lltbwd(n, c, india, x);
}

/* precons / presols: diagonal preconditioner with constraint blocks (src/mpnum.f90:2881-3069) */
void orc_precons(int p, int n, int b, const double *c, double *cu, double *a, const int *l, double *s, int *nrkd)
{
    const double *C = c - 1;
    double *CU = cu - 1, *A = a - 1, *S = s - 1;
    const int *L = l - 1;
    *nrkd = 0;
    for (int i = 1; i <= (p * p + p) / 2; i++) S[i] = 0.0;
    for (int i = 1; i <= n; i++) {
        double div = C[i];
        if (div > 0.0) CU[i] = 1.0 / sqrt(div);
        else { CU[i] = 0.0; (*nrkd)++; }
    }
    int64_t ij = 0;
    for (int ib = 1; ib <= b; ib++) {
        int p1 = L[ib * 4 - 3], p2 = L[ib * 4 - 2], i0 = L[ib * 4 - 1], np = L[ib * 4];
        for (int i = 1; i <= np; i++) {
            ij++;
            for (int j = p1; j <= p2; j++) {
                int jk = (j * j - j) / 2 + p1;
                A[ij + (int64_t)(j - p1) * np] = A[ij + (int64_t)(j - p1) * np] * CU[i + i0];
                for (int k = p1; k <= j; k++) {
                    S[jk] = S[jk] + A[ij + (int64_t)(j - p1) * np] * A[ij + (int64_t)(k - p1) * np];
                    jk++;
                }
            }
        }
        ij = ij + (int64_t)(p2 - p1) * np;
    }
    int ii = 0;
    for (int i = 1; i <= p; i++) {
        ii = ii + i;
        if (S[ii] != 0.0) S[ii] = 1.0 / S[ii];
        int jj = ii;
        for (int j = i + 1; j <= p; j++) {
            double ratio = S[i + jj] * S[ii];
            int kk = jj;
            for (int k = j; k <= p; k++) { S[kk + j] = S[kk + j] - S[kk + i] * ratio; kk = kk + k; }
            S[i + jj] = ratio;
            jj = jj + j;
        }
    }
}

void orc_presols(int p, int n, int b, const double *cu, const double *a, const int *l, const double *s,
                 double *x, const double *y)
{
    const double *CU = cu - 1, *A = a - 1, *S = s - 1, *Y = y - 1;
    const int *L = l - 1;
    double *X = x - 1;
    for (int i = 1; i <= n + p; i++) X[i] = Y[i];
    for (int i = 1; i <= n; i++) X[i] = X[i] * CU[i];
    int64_t ij = 0;
    for (int ib = 1; ib <= b; ib++) {
        int p1 = L[ib * 4 - 3], p2 = L[ib * 4 - 2], i0 = L[ib * 4 - 1], np = L[ib * 4];
        for (int j = p1; j <= p2; j++)
            for (int i = 1; i <= np; i++) { ij++; X[n + j] = X[n + j] - A[ij] * X[i + i0]; }
    }
    int jj = 0;
    for (int j = 1; j <= p; j++) {
        double dsum = X[n + j];
        for (int k = 1; k <= j - 1; k++) dsum = dsum - S[k + jj] * X[n + k];
        X[n + j] = dsum;
        jj = jj + j;
    }
    for (int j = p; j >= 1; j--) {
        double dsum = X[n + j] * S[jj];
        int kk = jj;
        for (int k = j + 1; k <= p; k++) { dsum = dsum + S[kk + j] * X[n + k]; kk = kk + k; }
        X[n + j] = -dsum;
        jj = jj - j;
    }
    ij = 0;
    for (int ib = 1; ib <= b; ib++) {
        int p1 = L[ib * 4 - 3], p2 = L[ib * 4 - 2], i0 = L[ib * 4 - 1], np = L[ib * 4];
        for (int j = p1; j <= p2; j++)
            for (int i = 1; i <= np; i++) { ij++; X[i + i0] = X[i + i0] - A[ij] * X[n + j]; }
    }
    for (int i = 1; i <= n; i++) X[i] = X[i] * CU[i];
}

/* ------------------------------------------------------------------------------------------ */
/* MINRES (Paige & Saunders), as in src/minresModule.f90:294-660 (checkA=.false.)              */
/* ------------------------------------------------------------------------------------------ */
typedef void (*orc_op)(void *ctx, int n, const double *x, double *y);

static double dotp(int n, const double *a, const double *b)
{
    double s = 0.0;
    for (int i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

void orc_minres(int n, orc_op aprod, orc_op msolve, void *ctx, const double *b, double shift, int precon,
                double *x, int itnlim, double rtol, int *istop_out, int *itn_out, double *out5 /* Anorm,Acond,rnorm,Arnorm,ynorm */)
{
    double *r1 = calloc(n, 8), *r2 = calloc(n, 8), *v = calloc(n, 8), *w = calloc(n, 8), *w1 = calloc(n, 8),
           *w2 = calloc(n, 8), *y = calloc(n, 8);
    double eps = DBL_EPSILON;
    int istop = 0, itn = 0;
    double Anorm = 0, Acond = 0, rnorm = 0, ynorm = 0, Arnorml = 0, gmin = 0, gmax = 0;
    double alfa, beta, beta1, cs, dbar, delta, denom, diag, epsa, epsln, epsr, epsx, gamma, gbar, oldb, oldeps,
        qrnorm, phi, phibar, rhs1, rhs2, rnorml, rootl, relArnorml, s, sn, tnorm2, ynorm2, z;
    for (int i = 0; i < n; i++) { x[i] = 0.0; y[i] = b[i]; r1[i] = b[i]; }
    if (precon) msolve(ctx, n, b, y);
    beta1 = dotp(n, b, y);
    if (beta1 < 0.0) { istop = 8; goto done; }
    if (beta1 == 0.0) { istop = 0; goto done; }
    beta1 = sqrt(beta1);
    aprod(ctx, n, y, w);
    Arnorml = sqrt(dotp(n, w, w));
    oldb = 0; beta = beta1; dbar = 0; epsln = 0; qrnorm = beta1; phibar = beta1; rhs1 = beta1; rhs2 = 0;
    tnorm2 = 0; ynorm2 = 0; cs = -1; sn = 0;
    for (int i = 0; i < n; i++) { w[i] = 0; w2[i] = 0; r2[i] = r1[i]; }
    for (;;) {
        itn++;
        s = 1.0 / beta;
        for (int i = 0; i < n; i++) v[i] = s * y[i];
        aprod(ctx, n, v, y);
        for (int i = 0; i < n; i++) y[i] = y[i] - shift * v[i];
        if (itn >= 2)
            for (int i = 0; i < n; i++) y[i] = y[i] - (beta / oldb) * r1[i];
        alfa = dotp(n, v, y);
        for (int i = 0; i < n; i++) y[i] = y[i] - (alfa / beta) * r2[i];
        for (int i = 0; i < n; i++) { r1[i] = r2[i]; r2[i] = y[i]; }
        if (precon) msolve(ctx, n, r2, y);
        oldb = beta;
        beta = dotp(n, r2, y);
        if (beta < 0.0) { istop = 6; goto done; }
        beta = sqrt(beta);
        tnorm2 = tnorm2 + alfa * alfa + oldb * oldb + beta * beta;
        if (itn == 1) {
            if (beta / beta1 <= 10.0 * eps) istop = -1;
            gmax = fabs(alfa);
            gmin = gmax;
        }
        oldeps = epsln;
        delta = cs * dbar + sn * alfa;
        gbar = sn * dbar - cs * alfa;
        epsln = sn * beta;
        dbar = -cs * beta;
        gamma = sqrt(gbar * gbar + beta * beta);
        cs = gbar / gamma;
        sn = beta / gamma;
        phi = cs * phibar;
        phibar = sn * phibar;
        denom = 1.0 / gamma;
        for (int i = 0; i < n; i++) {
            w1[i] = w2[i];
            w2[i] = w[i];
            w[i] = (v[i] - oldeps * w1[i] - delta * w2[i]) * denom;
            x[i] = x[i] + phi * w[i];
        }
        gmax = gmax > gamma ? gmax : gamma;
        gmin = gmin < gamma ? gmin : gamma;
        z = rhs1 / gamma;
        ynorm2 = z * z + ynorm2;
        rhs1 = rhs2 - delta * z;
        rhs2 = -epsln * z;
        Anorm = sqrt(tnorm2);
        ynorm = sqrt(ynorm2);
        epsa = Anorm * eps;
        epsx = Anorm * ynorm * eps;
        epsr = Anorm * ynorm * rtol;
        diag = gbar;
        if (diag == 0.0) diag = epsa;
        (void)diag;
        qrnorm = phibar;
        rnorml = rnorm;
        rnorm = qrnorm;
        rootl = sqrt(gbar * gbar + dbar * dbar);
        Arnorml = rnorml * rootl;
        relArnorml = rootl / Anorm;
        Acond = gmax / gmin;
        if (istop == 0) {
            if (itn >= itnlim) istop = 5;
            if (Acond >= 0.1 / eps) istop = 4;
            if (epsx >= beta1) istop = 3;
            if (qrnorm <= epsx || relArnorml <= eps) istop = 2;
            if (qrnorm <= epsr || relArnorml <= rtol) istop = 1;
        }
        if (istop != 0) break;
    }
done:
    if (out5) { out5[0] = Anorm; out5[1] = Acond; out5[2] = rnorm; out5[3] = Arnorml; out5[4] = ynorm; }
    *istop_out = istop;
    *itn_out = itn;
    free(r1); free(r2); free(v); free(w); free(w1); free(w2); free(y);
}

/* Dense packed symmetric product y = A x for full storage (avprd0 matsto=1 path,
 * src/pede.f90:5069-5232): every stored off-diagonal element serves both halves. */
/* ---------------------------------------------------------------------------------------
 * MINRES-QLP (src/minresqlpModule.f90:411-1330, Choi/Paige/Saunders), as pede calls it
 * (mminrsqlp, src/pede.f90:8341-8437): no shift unless given, symmetry checks of A and M on,
 * disable off.  maxxnorm, trancond and acondlim are clipped as in the reference
 * (1/eps, 10^floor(log10(1/eps)/2), 10^floor(log10(1/eps))).  acondlim <= 0 means "absent".
 * out6 = rnorm, Arnorm, xnorm, Anorm, Acond, QLPiter.
 * --------------------------------------------------------------------------------------- */
static double nrm2s(int n, const double *x)   /* scaled 2-norm, src/minresqlpBlasModule.f90:184-217 */
{
    if (n < 1) return 0.0;
    if (n == 1) return fabs(x[0]);
    double scale = 0.0, ssq = 1.0;
    for (int i = 0; i < n; i++) {
        if (x[i] != 0.0) {
            double a = fabs(x[i]);
            if (scale < a) { double r = scale / a; ssq = 1.0 + ssq * (r * r); scale = a; }
            else { double r = a / scale; ssq = ssq + r * r; }
        }
    }
    return scale * sqrt(ssq);
}
static double nrm2v(int n, double a, double b, double c)
{
    double v[3] = {a, b, c};
    return nrm2s(n, v);
}
static void symortho(double a, double b, double *c, double *s, double *r)  /* src/minresqlpModule.f90:1337-1380 */
{
    double aa = fabs(a), ab = fabs(b), t;
    if (ab <= DBL_MIN) { *s = 0.0; *r = aa; *c = (a == 0.0) ? 1.0 : a / aa; }
    else if (aa <= DBL_MIN) { *c = 0.0; *r = ab; *s = b / ab; }
    else if (ab > aa) { t = a / b; *s = (b / ab) / sqrt(1.0 + t * t); *c = *s * t; *r = b / *s; }
    else { t = b / a; *c = (a / aa) / sqrt(1.0 + t * t); *s = *c * t; *r = a / *c; }
}

void orc_minresqlp(int n, orc_op aprod, orc_op msolve, void *ctx, const double *b, double shift, int precon,
                   double *x, int itnlim, double rtol, double maxxnorm, double trancond, double acondlim,
                   int *istop_out, int *itn_out, double *out6)
{
    const double eps = DBL_EPSILON, one = 1.0, zero = 0.0;
    const double EPSINV = pow(10.0, floor(log10(one / eps)));
    const double NORMMAX = pow(10.0, floor(log10(one / eps) / 2));
    double maxxnorm_ = fmin(maxxnorm, one / eps);
    double trancond_ = fmin(trancond, NORMMAX);
    double Acondlim_ = acondlim > 0 ? fmin(acondlim, EPSINV) : EPSINV;
    double *r1 = calloc(n, 8), *r2 = calloc(n, 8), *v = calloc(n, 8), *w = calloc(n, 8), *wl = calloc(n, 8),
           *wl2 = calloc(n, 8), *xl2 = calloc(n, 8), *y = calloc(n, 8);
    double t1 = 0, t2 = 0, gminl = 0, gminl2 = 0, s, t, z, epsa;
    int lastiter = 0, istop = 0, itn = 0, QLPiter = 0;
    double beta1 = nrm2s(n, b);
    const double ieps = 0.1 / eps;
    double xnorm = 0, xl2norm = 0, Axnorm = 0, Anorm = 0, Acond = 1, pnorm = 0, relresl = 0, relAresl = 0;
    double Arnorm = 0, relAres = 0;
    for (int i = 0; i < n; i++) { x[i] = 0.0; xl2[i] = 0.0; y[i] = b[i]; r1[i] = b[i]; }
    if (precon) msolve(ctx, n, b, y);
    beta1 = dotp(n, b, y);
    if (beta1 < zero && nrm2s(n, y) > eps) istop = 11;
    if (beta1 == zero) istop = 3;
    beta1 = sqrt(beta1);
    if (precon) {
        msolve(ctx, n, y, r2);
        s = dotp(n, y, y);
        t = dotp(n, r1, r2);
        z = fabs(s - t);
        epsa = (fabs(s) + eps) * pow(eps, 0.33333);
        if (z > epsa) istop = 10;
    }
    aprod(ctx, n, y, w);
    aprod(ctx, n, w, r2);
    s = dotp(n, w, w);
    t = dotp(n, y, r2);
    z = fabs(s - t);
    epsa = (fabs(s) + eps) * pow(eps, 0.33333);
    if (z > epsa) istop = 9;

    double tau = 0, taul = 0, taul2 = 0, gmin = 0, beta = 0, betal = 0, betan = beta1, dbar = 0, dltan = 0, eta = 0, etal = 0,
           etal2 = 0, vepln = 0, veplnl = 0, veplnl2 = 0, eplnn = 0, epln = 0, gama = 0, gamal = 0, gamal2 = 0, gamal3 = 0,
           phi = beta1, cr1 = -1, sr1 = 0, cr2 = -1, sr2 = 0, cs = -1, sn = 0, ul4 = 0, ul3 = 0, ul2 = 0, ul = 0, u = 0;
    double gamal_QLP = 0, vepln_QLP = 0, gama_QLP = 0, ul_QLP = 0, u_QLP = 0;
    double alfa, dlta, dlta_QLP, dlta_tmp, gbar, gama_tmp, gamal_tmp, xnorm_tmp, xnorml, abs_gama, Anorml, Acondl, rnorml,
        rootl, Arnorml = 0, epsx;
    double rnorm = betan;
    double relres = rnorm / (Anorm * xnorm + beta1);
    for (int i = 0; i < n; i++) { r2[i] = b[i]; w[i] = 0.0; wl[i] = 0.0; }

    while (istop <= 0) {
        itn++;
        betal = beta;
        beta = betan;
        s = one / beta;
        for (int i = 0; i < n; i++) v[i] = s * y[i];
        aprod(ctx, n, v, y);
        if (shift != zero) for (int i = 0; i < n; i++) y[i] = y[i] - shift * v[i];
        if (itn >= 2) { double c = -beta / betal; for (int i = 0; i < n; i++) y[i] = y[i] + c * r1[i]; }
        alfa = dotp(n, v, y);
        { double c = -alfa / beta; for (int i = 0; i < n; i++) y[i] = y[i] + c * r2[i]; }
        memcpy(r1, r2, 8 * (size_t)n);
        memcpy(r2, y, 8 * (size_t)n);
        if (!precon) betan = nrm2s(n, y);
        else {
            msolve(ctx, n, r2, y);
            betan = dotp(n, r2, y);
            if (betan > zero) betan = sqrt(betan);
            else if (nrm2s(n, y) > eps) { istop = 11; break; }
        }
        pnorm = itn == 1 ? nrm2v(2, alfa, betan, 0) : nrm2v(3, beta, alfa, betan);

        dbar = dltan;
        dlta = cs * dbar + sn * alfa;
        epln = eplnn;
        gbar = sn * dbar - cs * alfa;
        eplnn = sn * betan;
        dltan = -cs * betan;
        dlta_QLP = dlta;

        gamal3 = gamal2;
        gamal2 = gamal;
        gamal = gama;
        symortho(gbar, betan, &cs, &sn, &gama);
        gama_tmp = gama;
        taul2 = taul;
        taul = tau;
        tau = cs * phi;
        phi = sn * phi;
        Axnorm = sqrt(Axnorm * Axnorm + tau * tau);

        if (itn > 2) {
            veplnl2 = veplnl;
            etal2 = etal;
            etal = eta;
            dlta_tmp = sr2 * vepln - cr2 * dlta;
            veplnl = cr2 * vepln + sr2 * dlta;
            dlta = dlta_tmp;
            eta = sr2 * gama;
            gama = -cr2 * gama;
        }
        if (itn > 1) {
            symortho(gamal, dlta, &cr1, &sr1, &gamal_tmp);
            gamal = gamal_tmp;
            vepln = sr1 * gama;
            gama = -cr1 * gama;
        }

        xnorml = xnorm;
        ul4 = ul3;
        ul3 = ul2;
        if (itn > 2) ul2 = (taul2 - etal2 * ul4 - veplnl2 * ul3) / gamal2;
        if (itn > 1) ul = (taul - etal * ul3 - veplnl * ul2) / gamal;
        xnorm_tmp = nrm2v(3, xl2norm, ul2, ul);
        if (fabs(gama) > eps) {
            u = (tau - eta * ul2 - vepln * ul) / gama;
            int likeLS = relAresl < relresl;
            if (likeLS && nrm2v(2, xnorm_tmp, u, 0) > maxxnorm_) { u = zero; istop = 12; }
        } else {
            u = zero;
            istop = 14;
        }
        xl2norm = nrm2v(2, xl2norm, ul2, 0);
        xnorm = nrm2v(3, xl2norm, ul, u);

        if (Acond < trancond_ && istop == 0 && QLPiter == 0) {    /* MINRES updates */
            memcpy(wl2, wl, 8 * (size_t)n);
            memcpy(wl, w, 8 * (size_t)n);
            if (gama_tmp > eps) {
                s = one / gama_tmp;
                for (int i = 0; i < n; i++) w[i] = (v[i] - epln * wl2[i] - dlta_QLP * wl[i]) * s;
            }
            if (xnorm < maxxnorm_) {
                for (int i = 0; i < n; i++) x[i] = x[i] + tau * w[i];
            } else {
                istop = 12;
                lastiter = 1;
            }
        } else {                                                     /* MINRES-QLP updates */
            QLPiter++;
            if (QLPiter == 1) {
                for (int i = 0; i < n; i++) xl2[i] = 0.0;
                if (itn > 1) {
                    if (itn > 3) for (int i = 0; i < n; i++) wl2[i] = gamal3 * wl2[i] + veplnl2 * wl[i] + etal * w[i];
                    if (itn > 2) for (int i = 0; i < n; i++) wl[i] = gamal_QLP * wl[i] + vepln_QLP * w[i];
                    for (int i = 0; i < n; i++) w[i] = gama_QLP * w[i];
                    for (int i = 0; i < n; i++) xl2[i] = x[i] - ul_QLP * wl[i] - u_QLP * w[i];
                }
            }
            if (itn == 1) {
                for (int i = 0; i < n; i++) { wl2[i] = wl[i]; wl[i] = sr1 * v[i]; w[i] = -cr1 * v[i]; }
            } else if (itn == 2) {
                for (int i = 0; i < n; i++) {
                    double wo = w[i];
                    wl2[i] = wl[i];
                    wl[i] = cr1 * wo + sr1 * v[i];
                    w[i] = sr1 * wo - cr1 * v[i];
                }
            } else {
                for (int i = 0; i < n; i++) {
                    double a2 = wl[i], a1 = w[i], vv;
                    /* wl2 = wl; wl = w; w = sr2*wl2 - cr2*v; wl2 = cr2*wl2 + sr2*v;
                       v = cr1*wl + sr1*w; w = sr1*wl - cr1*w; wl = v */
                    double nw = sr2 * a2 - cr2 * v[i];
                    double nwl2 = cr2 * a2 + sr2 * v[i];
                    vv = cr1 * a1 + sr1 * nw;
                    w[i] = sr1 * a1 - cr1 * nw;
                    wl2[i] = nwl2;
                    v[i] = vv;
                    wl[i] = vv;
                }
            }
            for (int i = 0; i < n; i++) {
                xl2[i] = xl2[i] + ul2 * wl2[i];
                x[i] = xl2[i] + ul * wl[i] + u * w[i];
            }
        }

        gamal_tmp = gamal;
        symortho(gamal_tmp, eplnn, &cr2, &sr2, &gamal);
        gamal_QLP = gamal_tmp;
        vepln_QLP = vepln;
        gama_QLP = gama;
        ul_QLP = ul;
        u_QLP = u;

        abs_gama = fabs(gama);
        Anorml = Anorm;
        Anorm = fmax(fmax(Anorm, pnorm), fmax(gamal, abs_gama));
        if (itn == 1) {
            gmin = gama;
            gminl = gmin;
        } else {
            gminl2 = gminl;
            gminl = gmin;
            gmin = fmin(fmin(gminl2, gamal), abs_gama);
        }
        Acondl = Acond;
        Acond = Anorm / gmin;
        rnorml = rnorm;
        relresl = relres;
        if (istop != 14) rnorm = phi;
        relres = rnorm / (Anorm * xnorm + beta1);
        rootl = nrm2v(2, gbar, dltan, 0);
        Arnorml = rnorml * rootl;
        relAresl = rootl / Anorm;

        epsx = Anorm * xnorm * eps;
        if (istop == 0 || istop == 14) {
            t1 = one + relres;
            t2 = one + relAresl;
        }
        if (t1 <= one) istop = 5;
        else if (t2 <= one) istop = 7;
        else if (relres <= rtol) istop = 4;
        else if (relAresl <= rtol) istop = 6;
        else if (epsx >= beta1) istop = 2;
        else if (xnorm >= maxxnorm_) istop = 12;
        else if (Acond >= Acondlim_ || Acond >= ieps) istop = 13;
        else if (itn >= itnlim) istop = 8;
        else if (betan < eps) istop = 1;

        if (istop != 0) {
            if (istop == 6 || istop == 7 || istop == 12 || istop == 13) lastiter = 1;
            if (lastiter) {
                itn = itn - 1;
                Acond = Acondl;
                rnorm = rnorml;
                relres = relresl;
            }
            aprod(ctx, n, x, r1);
            for (int i = 0; i < n; i++) r1[i] = b[i] - r1[i] + shift * x[i];
            aprod(ctx, n, r1, wl2);
            for (int i = 0; i < n; i++) wl2[i] = wl2[i] - shift * r1[i];
            Arnorm = nrm2s(n, wl2);
            if (rnorm > zero && Anorm > zero) relAres = Arnorm / (Anorm * rnorm);
            break;
        }
    }
    (void)Arnorml; (void)relAres; (void)xnorml; (void)Anorml;
    *istop_out = istop;
    *itn_out = itn;
    if (out6) {
        out6[0] = rnorm; out6[1] = Arnorm; out6[2] = xnorm; out6[3] = Anorm; out6[4] = Acond; out6[5] = QLPiter;
    }
    free(r1); free(r2); free(v); free(w); free(wl); free(wl2); free(xl2); free(y);
}

void orc_packed_matvec(int n, const double *a, const double *x, double *y)
{
    for (int i = 0; i < n; i++) y[i] = 0.0;
    for (int i = 1; i <= n; i++) {
        int64_t off = IDX(i, 0);
        double s = 0.0;
        for (int j = 1; j < i; j++) {
            double aij = a[off + j - 1];
            s += aij * x[j - 1];
            y[j - 1] += aij * x[i - 1];
        }
        y[i - 1] += s + a[off + i - 1] * x[i - 1];
    }
}


/* ------------------------------------------------------------------------------------------ */
/* Root-free band Cholesky and bordered-band local fit (src/Dbandmatrix.f90:459-638,          */
/* src/mpnum.f90:3118-3628).  w is column-major w(mp1,n): W(r,c) = w[(c-1)*mp1 + r-1].        */
/* ------------------------------------------------------------------------------------------ */
#define W_(r, c) w[((int64_t)(c) - 1) * mp1 + (r) - 1]
void orc_dbcdec(double *w, int mp1, int n, double *aux)
{
    aux--;
    for (int i = 1; i <= n; i++) aux[i] = 16.0 * W_(1, i);
    for (int i = 1; i <= n; i++) {
        if (W_(1, i) + aux[i] != aux[i]) W_(1, i) = 1.0 / W_(1, i);
        else W_(1, i) = 0.0;
        int jm = mp1 - 1 < n - i ? mp1 - 1 : n - i;
        for (int j = 1; j <= jm; j++) {
            double rxw = W_(j + 1, i) * W_(1, i);
            for (int k = 1; k <= jm + 1 - j; k++) W_(k, i + j) = W_(k, i + j) - W_(k + j, i) * rxw;
            W_(j + 1, i) = rxw;
        }
    }
}

void orc_dbcslv(const double *w, int mp1, int n, const double *b, double *x)
{
    b--; x--;
    for (int i = 1; i <= n; i++) x[i] = b[i];
    for (int i = 1; i <= n; i++) {
        int jm = mp1 - 1 < n - i ? mp1 - 1 : n - i;
        for (int j = 1; j <= jm; j++) x[j + i] = x[j + i] - W_(j + 1, i) * x[i];
    }
    for (int i = n; i >= 1; i--) {
        double rxw = x[i] * W_(1, i);
        int jm = mp1 - 1 < n - i ? mp1 - 1 : n - i;
        for (int j = 1; j <= jm; j++) rxw = rxw - W_(j + 1, i) * x[j + i];
        x[i] = rxw;
    }
}

/* band part (inv=1, dbcinb) or complete (inv=2, dbcinv) inverse into packed vs */
static void dbc_inverse(const double *w, int mp1, int n, double *vs, int full)
{
    vs--;
    for (int i = n; i >= 1; i--) {
        double rxw = W_(1, i);
        int jlo = full ? 1 : (i - mp1 + 1 > 1 ? i - mp1 + 1 : 1);
        for (int j = i; j >= jlo; j--) {
            int kmax = n < j + mp1 - 1 ? n : j + mp1 - 1;
            for (int k = j + 1; k <= kmax; k++) {
                int64_t mx = i > k ? i : k, mn = i < k ? i : k;
                rxw = rxw - vs[(mx * (mx - 1)) / 2 + mn] * W_(1 + k - j, j);
            }
            vs[((int64_t)i * i - i) / 2 + j] = rxw;
            rxw = 0.0;
        }
    }
}
void orc_dbcinb(const double *w, int mp1, int n, double *vs) { dbc_inverse(w, mp1, n, vs, 0); }
void orc_dbcinv(const double *w, int mp1, int n, double *vs) { dbc_inverse(w, mp1, n, vs, 1); }
#undef W_

/* sqmibb (border first, side 1) and sqmibb2 (border last, side 2).  v packed (n), b(n);
 * returns nrank, evdmin, evdmax. */
void orc_sqmibb(double *v, double *b, int n, int nbdr, int nbnd, int inv, int *nrank, double *evdmin,
                double *evdmax, int side)
{
    const int mp1 = nbnd + 1, nmb = n - nbdr;
    double *vbnd = (double *)calloc((size_t)n * (nbnd + 1) + 1, sizeof(double));
    double *vbdr = (double *)calloc((size_t)n * (nbdr > 0 ? nbdr : 1) + 1, sizeof(double));
    double *aux = (double *)calloc((size_t)n * (nbdr > 0 ? nbdr : 1) + n + 1, sizeof(double));
    double *vbk = (double *)calloc((size_t)(nbdr * nbdr + nbdr) / 2 + 1, sizeof(double));
    double *vzru = (double *)calloc((size_t)nbdr + 1, sizeof(double));
    double *V = v - 1, *B = b - 1;
    double *VBND = vbnd - 1, *VBDR = vbdr - 1, *AUX = aux - 1, *VBK = vbk - 1, *VZRU = vzru - 1;
    *nrank = 0;
    const int nb1 = side == 1 ? nbdr + 1 : nmb + 1;
    if (side == 1) {
        for (int i = nb1; i <= n; i++) {
            int ip = (i * (i + 1)) / 2, is = 0;
            for (int j = i; j <= (n < i + nbnd ? n : i + nbnd); j++) {
                ip = ip + is; is = j;
                VBND[(j - i + 1) + (i - nb1) * mp1] = V[ip];
            }
        }
        if (nbdr > 0) {
            int ioff = 0;
            for (int i = 1; i <= nbdr; i++) {
                int ip = (i * (i + 1)) / 2, is = 0;
                for (int j = i; j <= n; j++) { ip = ip + is; is = j; VBDR[ioff + j] = V[ip]; }
                ioff = ioff + n;
            }
        }
    } else {
        for (int i = 1; i <= nmb; i++) {
            int ip = (i * (i + 1)) / 2, is = 0;
            for (int j = i; j <= (nmb < i + nbnd ? nmb : i + nbnd); j++) {
                ip = ip + is; is = j;
                VBND[(j - i + 1) + (i - 1) * mp1] = V[ip];
            }
        }
        if (nbdr > 0) {
            int ioff = 0;
            for (int i = nb1; i <= n; i++) {
                int ip = (i * (i - 1)) / 2;
                for (int j = 1; j <= i; j++) VBDR[ioff + j] = V[ip + j];
                ioff = ioff + n;
            }
        }
    }
    orc_dbcdec(vbnd, mp1, nmb, aux);
    int ip = 1;
    *evdmin = 0.0;
    *evdmax = 0.0;
    for (int i = 1; i <= nmb; i++) {
        if (VBND[ip] <= 0.0) {
            for (int q = 1; q <= n; q++) B[q] = 0.0;
            for (int q = 1; q <= (n * n + n) / 2; q++) V[q] = 0.0;
            *nrank = 0;
            goto out;
        }
        (*nrank)++;
        if (*nrank == 1) { *evdmin = VBND[ip]; *evdmax = VBND[ip]; }
        else {
            if (VBND[ip] < *evdmin) *evdmin = VBND[ip];
            if (VBND[ip] > *evdmax) *evdmax = VBND[ip];
        }
        ip = ip + mp1;
    }
    if (nbdr == 0) {
        orc_dbcslv(vbnd, mp1, nmb, b, b);
        if (inv > 0) { if (inv > 1) orc_dbcinv(vbnd, mp1, nmb, v); else orc_dbcinb(vbnd, mp1, nmb, v); }
        goto out;
    }
    if (side == 1) {
        int ioff = nb1;
        for (int ib = 1; ib <= nbdr; ib++) {
            orc_dbcslv(vbnd, mp1, nmb, &VBDR[ioff], &AUX[ioff]);
            VZRU[ib] = B[ib];
            for (int i = 0; i <= nmb - 1; i++) VZRU[ib] = VZRU[ib] - B[nb1 + i] * AUX[ioff + i];
            ioff = ioff + n;
        }
        orc_dbcslv(vbnd, mp1, nmb, &B[nb1], &B[nb1]);
        int q = 0;
        ioff = nb1;
        for (int ib = 1; ib <= nbdr; ib++) {
            int joff = nb1;
            for (int jb = 1; jb <= ib; jb++) {
                q = q + 1;
                VBK[q] = V[q];
                for (int i = 0; i <= nmb - 1; i++) VBK[q] = VBK[q] - VBDR[ioff + i] * AUX[joff + i];
                joff = joff + n;
            }
            ioff = ioff + n;
        }
    } else {
        int ioff = 0;
        for (int ib = 1; ib <= nbdr; ib++) {
            orc_dbcslv(vbnd, mp1, nmb, &VBDR[ioff + 1], &AUX[ioff + 1]);
            VZRU[ib] = B[nmb + ib];
            for (int i = 1; i <= nmb; i++) VZRU[ib] = VZRU[ib] - B[i] * AUX[ioff + i];
            ioff = ioff + n;
        }
        orc_dbcslv(vbnd, mp1, nmb, b, b);
        int q = 0, koff = nmb;
        ioff = 0;
        for (int ib = 1; ib <= nbdr; ib++) {
            int joff = 0;
            for (int jb = 1; jb <= ib; jb++) {
                q = q + 1;
                VBK[q] = VBDR[koff + jb];
                for (int i = 1; i <= nmb; i++) VBK[q] = VBK[q] - VBDR[ioff + i] * AUX[joff + i];
                joff = joff + n;
            }
            ioff = ioff + n;
            koff = koff + n;
        }
    }
    {
        int nrankb = 0;
        orc_sqminv(vbk, vzru, nbdr, &nrankb);
        if (nrankb == nbdr) *nrank = *nrank + nbdr;
        else {
            for (int ib = 1; ib <= nbdr; ib++) VZRU[ib] = 0.0;
            for (int q = (nbdr * nbdr + nbdr) / 2; q >= 1; q--) VBK[q] = 0.0;
        }
    }
    if (side == 1) {
        int ioff = nb1;
        for (int ib = 1; ib <= nbdr; ib++) {
            B[ib] = VZRU[ib];
            for (int i = 0; i <= nmb - 1; i++) B[nb1 + i] = B[nb1 + i] - B[ib] * AUX[ioff + i];
            ioff = ioff + n;
        }
        if (inv > 0) {
            if (inv > 1) orc_dbcinv(vbnd, mp1, nmb, v); else orc_dbcinb(vbnd, mp1, nmb, v);
            int ip1 = (nmb * nmb + nmb) / 2, ip2 = (n * n + n) / 2;
            for (int i = nmb - 1; i >= 0; i--) {
                int j0 = 0;
                if (inv == 1) j0 = i - nbnd > 0 ? i - nbnd : 0;
                for (int j = i; j >= j0; j--) {
                    V[ip2] = V[ip1];
                    int io = nb1;
                    for (int ib = 1; ib <= nbdr; ib++) {
                        int jo = nb1;
                        for (int jb = 1; jb <= nbdr; jb++) {
                            int ij = ib > jb ? ib : jb;
                            ij = (ij * ij - ij) / 2 + (ib < jb ? ib : jb);
                            V[ip2] = V[ip2] + VBK[ij] * AUX[io + i] * AUX[jo + j];
                            jo = jo + n;
                        }
                        io = io + n;
                    }
                    ip1 = ip1 - 1;
                    ip2 = ip2 - 1;
                }
                ip1 = ip1 - j0;
                ip2 = ip2 - j0;
                for (int ib = nbdr; ib >= 1; ib--) {
                    V[ip2] = 0.0;
                    int jo = nb1;
                    for (int jb = 1; jb <= nbdr; jb++) {
                        int ij = ib > jb ? ib : jb;
                        ij = (ij * ij - ij) / 2 + (ib < jb ? ib : jb);
                        V[ip2] = V[ip2] - VBK[ij] * AUX[i + jo];
                        jo = jo + n;
                    }
                    ip2 = ip2 - 1;
                }
            }
            for (int q = (nbdr * nbdr + nbdr) / 2; q >= 1; q--) { V[ip2] = VBK[q]; ip2 = ip2 - 1; }
        }
    } else {
        int ioff = 0;
        for (int ib = 1; ib <= nbdr; ib++) {
            for (int i = 1; i <= nmb; i++) B[i] = B[i] - VZRU[ib] * AUX[ioff + i];
            ioff = ioff + n;
            B[nmb + ib] = VZRU[ib];
        }
        if (inv > 0) {
            if (inv > 1) orc_dbcinv(vbnd, mp1, nmb, v); else orc_dbcinb(vbnd, mp1, nmb, v);
            int ip1 = (nmb * nmb + nmb) / 2;
            for (int i = nmb - 1; i >= 0; i--) {
                int j0 = 0;
                if (inv == 1) j0 = i - nbnd > 0 ? i - nbnd : 0;
                for (int j = i; j >= j0; j--) {
                    int io = 1;
                    for (int ib = 1; ib <= nbdr; ib++) {
                        int jo = 1;
                        for (int jb = 1; jb <= nbdr; jb++) {
                            int ij = ib > jb ? ib : jb;
                            ij = (ij * ij - ij) / 2 + (ib < jb ? ib : jb);
                            V[ip1] = V[ip1] + VBK[ij] * AUX[io + i] * AUX[jo + j];
                            jo = jo + n;
                        }
                        io = io + n;
                    }
                    ip1 = ip1 - 1;
                }
                ip1 = ip1 - j0;
            }
            ip1 = (nmb * nmb + nmb) / 2;
            int q = 0;
            for (int ib = 1; ib <= nbdr; ib++) {
                for (int i = 1; i <= nmb; i++) {
                    ip1 = ip1 + 1;
                    V[ip1] = 0.0;
                    int jo = 0;
                    for (int jb = 1; jb <= nbdr; jb++) {
                        int ij = ib > jb ? ib : jb;
                        ij = (ij * ij - ij) / 2 + (ib < jb ? ib : jb);
                        V[ip1] = V[ip1] - VBK[ij] * AUX[i + jo];
                        jo = jo + n;
                    }
                }
                for (int jb = 1; jb <= ib; jb++) { ip1 = ip1 + 1; q = q + 1; V[ip1] = VBK[q]; }
            }
        }
    }
out:
    free(vbnd); free(vbdr); free(aux); free(vbk); free(vzru);
}

/* ------------------------------------------------------------------------------------------ */
/* Record loop: loopbf restated (src/pede.f90:3387-4442) with decodeNextMeasurement            */
/* (src/pede.f90:2428-2468), addsums (11666-11689) and mgupdt full-storage flush              */
/* (3336-3356).  Local fits take the full (sqminv) or the bordered-band (sqmibb/sqmibb2) path  */
/* by the cost model of 3838-3874.                                                             */
/* ------------------------------------------------------------------------------------------ */
typedef struct {
    int64_t nrec;
    const int64_t *rec_ptr; /* record r occupies words [rec_ptr[r], rec_ptr[r+1]) (word 1 = header) */
    const int32_t *inder;
    const double *glder;
    int32_t ntot;                /* number of labels (total parameters) */
    const int32_t *labels;       /* sorted ascending, ntot */
    const int32_t *var_index;    /* per total index (0-based array): >0 variable index (1-based) */
    const double *global_par;    /* per total index */
    int32_t nvar;
} orc_records;

typedef struct {
    int32_t icalcm, nloopn, lhuber, lfitbb;
    float chicut, chhuge, dwcut, cndlmx;
    int32_t nthreads;
} orc_pass_opts;

typedef struct {
    double chi2;        /* Sum(chi^2) as getsums (fvalue) */
    int64_t ndf;
    double wndf;
    int64_t nrej[6];
    int64_t naccept;
    int64_t nbandrec;   /* records whose local fit took the bordered-band path */
} orc_pass_stats;

static int label_to_index(const orc_records *R, int32_t label)
{
    int lo = 0, hi = R->ntot - 1;
    while (lo <= hi) {
        int mid = (lo + hi) >> 1;
        int32_t v = R->labels[mid];
        if (v == label) return mid;
        if (v < label) lo = mid + 1; else hi = mid - 1;
    }
    return -1;
}

typedef struct { double chi_frac; int64_t chi_int; int64_t ndf; double wndf; } orc_sums;

/* matrix sink: packed full matrix, or an open-addressing hash of (i,j) -> value used to
 * accumulate very large sparse systems on the CPU (keys are i<<32|j, 0 = empty slot) */
typedef struct {
    double *mat; uint64_t *keys; double *vals; int64_t cap; int64_t used; int overflow;
    /* group-block storage (the device layout): groups of variables (gfirst 1-based, ngroup+1),
     * per variable its group, block rows over groups with ascending column groups and the
     * offset of each row-major block; updates added atomically (several threads share it) */
    const int32_t *gfirst, *group_of, *bcol; const int64_t *brow, *boff; double *blk;
} orc_sink;

static void sink_add(orc_sink *S, int va, int vb, double v)
{
    if (S->mat) { S->mat[IDX(va, vb) - 1] += v; return; }
    if (S->blk) {
        if (va < vb) { int t = va; va = vb; vb = t; }
        const int ga = S->group_of[va], gb = S->group_of[vb];
        int64_t lo = S->brow[ga], hi = S->brow[ga + 1] - 1;
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (S->bcol[mid] < gb) lo = mid + 1; else hi = mid;
        }
        if (lo > hi || S->bcol[lo] != gb) { S->overflow = 1; return; }
        const int nc = S->gfirst[gb + 1] - S->gfirst[gb];
        double *p = S->blk + S->boff[lo] + (int64_t)(va - S->gfirst[ga]) * nc + (vb - S->gfirst[gb]);
#pragma omp atomic
        *p += v;
        return;
    }
    uint64_t key = ((uint64_t)(uint32_t)va << 32) | (uint32_t)vb;
    uint64_t h = key * 0x9e3779b97f4a7c15ull;
    int64_t mask = S->cap - 1;
    int64_t p = (int64_t)(h >> 17) & mask;
    for (;;) {
        uint64_t k = S->keys[p];
        if (k == key) { S->vals[p] += v; return; }
        if (k == 0) {
            if (S->used * 10 >= S->cap * 9) { S->overflow = 1; return; }
            S->keys[p] = key; S->vals[p] = v; S->used++; return;
        }
        p = (p + 1) & mask;
    }
}

static void addsums(orc_sums *s, double chi2, int ndf, double dw)
{
    double add = chi2 * dw;
    int64_t nadd = (int64_t)add;
    s->chi_int += nadd;
    s->chi_frac += add - (double)nadd;
    if (s->chi_frac > 16.0) { s->chi_frac -= 16.0; s->chi_int += 16; }
    s->ndf += ndf;
    s->wndf += (double)ndf * dw;
}

/* one record; returns 0 ok (accepted or rejected, see stats), 20 bad record structure */
static int loop_record(const orc_records *R, int64_t r, const orc_pass_opts *o, orc_sink *mat, double *vec,
                       int64_t *counter, orc_sums *sums, int64_t nrej[6], int64_t *naccept, int64_t *nband,
                       int *back /* nvar+1 scratch, zero on entry/exit */)
{
    const double chuber = 1.345, cauchy = 2.3849;
    int64_t base = R->rec_ptr[r] - 1; /* so that word w (1-based) is at base+w */
    int nw = (int)(R->rec_ptr[r + 1] - R->rec_ptr[r]);
    const int32_t *I = R->inder + base;
    const double *Din = R->glder + base;
    double *D = (double *)malloc(sizeof(double) * (nw + 1));
    for (int w = 1; w <= nw; w++) D[w] = Din[w];
    int *tix = (int *)malloc(sizeof(int) * (nw + 1)); /* total index per global word (0-based), -1 unknown */
    int cap = 64, neq = 0;
    int *eqs = (int *)malloc(sizeof(int) * 3 * cap);
    int nalc = 0, nalg = 0, rc = 0;
    int *gidx = (int *)malloc(sizeof(int) * (nw + 1));
    int lastGlobal = 1, endOfEntry = nw;
    /* ---- first loop: decode, subtract global contribution, collect globals ---- */
    for (;;) {
        int startLocal = 0, startGlobal = 0;
        if (lastGlobal >= endOfEntry) break;
        do { lastGlobal++; if (lastGlobal > nw) { rc = 20; goto out; } } while (I[lastGlobal] != 0);
        startLocal = lastGlobal;
        do { lastGlobal++; if (lastGlobal > nw) { rc = 20; goto out; } } while (I[lastGlobal] != 0);
        startGlobal = lastGlobal;
        if (startLocal + 1 == startGlobal && D[startGlobal] < 0.0) { /* special data */
            lastGlobal = lastGlobal + (int)lrint(-D[startGlobal]);
            continue;
        }
        while (lastGlobal < endOfEntry && I[lastGlobal + 1] != 0) lastGlobal++;
        if (neq == cap) { cap *= 2; eqs = (int *)realloc(eqs, sizeof(int) * 3 * cap); }
        eqs[3 * neq] = startLocal; eqs[3 * neq + 1] = startGlobal; eqs[3 * neq + 2] = lastGlobal;
        neq++;
        double rmeas = D[startLocal];
        for (int j = 1; j <= lastGlobal - startGlobal; j++) {
            int w = startGlobal + j;
            int ti = label_to_index(R, I[w]);
            if (ti < 0) { rc = 20; goto out; }
            tix[w] = ti;
            rmeas = rmeas - D[w] * R->global_par[ti];
            if (o->icalcm == 1) {
                int ij = R->var_index[ti];
                if (ij > 0 && back[ij] == 0) { nalg++; gidx[nalg] = ij; back[ij] = nalg; }
            }
        }
        D[startLocal] = rmeas;
        for (int j = 1; j <= startGlobal - startLocal - 1; j++) if (I[startLocal + j] > nalc) nalc = I[startLocal + j];
    }
    if (nalc <= 0) goto out;
    {
        double dw1 = 1.0, dw2 = 1.0; /* file weight (fileLevelWeight), 1 unless steered */
        int ngg = (nalg * nalg + nalg) / 2;
        double *upd = NULL, *glm = NULL;
        if (o->icalcm == 1) {
            /* sort1k of the record's variable indices, update back index (3699-3741) */
            for (int a = 2; a <= nalg; a++) { int t = gidx[a], b = a - 1; while (b >= 1 && gidx[b] > t) { gidx[b + 1] = gidx[b]; b--; } gidx[b + 1] = t; }
            for (int k = 1; k <= nalg; k++) back[gidx[k]] = k;
            upd = (double *)calloc((size_t)ngg + 1, sizeof(double));
            glm = (double *)calloc((size_t)nalg * nalc + 1, sizeof(double));
        }
        int nter = 1;
        if (o->nloopn != 1 && o->lhuber != 0) nter = o->lhuber;
        double *lcorr = (double *)calloc(neq + 1, sizeof(double));
        double *lwgt = (double *)calloc(neq + 1, sizeof(double));
        int npk = (nalc * nalc + nalc) / 2;
        double *clmat = (double *)calloc(npk + 1, sizeof(double));
        double *blvec = (double *)calloc(nalc + 1, sizeof(double));
        int *ibandh = (int *)calloc(2 * nalc + 2, sizeof(int));
        int mbnd = -1, mbdr = nalc, mside = -1, nrank = 0, ndf = 0, nan = 0;
        double summ = 0.0, suwt = 0.0;
        float resing = 0.0f, cndl10 = 0.0f;
        int rejected = 0;
        for (int iter = 1; iter <= nter; iter++) {
            for (int i = 1; i <= nalc; i++) blvec[i] = 0.0;
            for (int i = 1; i <= npk; i++) clmat[i] = 0.0;
            int nweig = 0;
            cndl10 = 0.0f;
            for (int e = 0; e < neq; e++) {
                int sL = eqs[3 * e], sG = eqs[3 * e + 1];
                double rmeas = D[sL], rerr = D[sG];
                double wght = 1.0 / (rerr * rerr);
                nweig++;
                double resid = rmeas - lcorr[e + 1];
                if (o->nloopn != 1 && iter != 1 && o->lhuber != 0) {
                    if (iter <= 3) {
                        if (fabs(resid) > chuber * rerr) wght = wght * chuber * rerr / fabs(resid);
                    } else {
                        double t = resid / rerr / cauchy;
                        wght = wght / (1.0 + t * t);
                    }
                }
                lwgt[e + 1] = wght;
                for (int j = 1; j <= sG - sL - 1; j++) {
                    int ij = I[sL + j];
                    blvec[ij] = blvec[ij] + wght * rmeas * D[sL + j];
                    for (int k = 1; k <= j; k++) {
                        int ik = I[sL + k];
                        int jk = (ij * ij - ij) / 2 + ik;
                        clmat[jk] = clmat[jk] + wght * D[sL + j] * D[sL + k];
                        if (iter == 1) {
                            int id = abs(ij - ik) + 1;
                            int im = ij < ik ? ij : ik;
                            if (im > ibandh[id]) ibandh[id] = im;
                            im = (nalc + 1 - ij) < (nalc + 1 - ik) ? (nalc + 1 - ij) : (nalc + 1 - ik);
                            if (im > ibandh[nalc + id]) ibandh[nalc + id] = im;
                        }
                    }
                }
            }
            if (iter == 1 && nalc > 5 && o->lfitbb > 0) {
                int kx = -1, kbdrx = 0;
                int64_t icmn = (int64_t)nalc * nalc * nalc;
                int kbdr = 0;
                for (int k = nalc; k >= 2; k--) {
                    int kbnd = k - 2;
                    if (ibandh[k] > kbdr) kbdr = ibandh[k];
                    int64_t t = kbnd + kbdr + 1;
                    int64_t icost = 6 * (int64_t)(nalc - kbdr) * t * t + 2 * (int64_t)kbdr * kbdr * kbdr;
                    if (icost < icmn) { icmn = icost; kx = k; kbdrx = kbdr; mside = 1; }
                }
                if (kx < 0) {
                    kbdr = 0;
                    for (int k = nalc; k >= 2; k--) {
                        int kbnd = k - 2;
                        if (ibandh[k + nalc] > kbdr) kbdr = ibandh[k + nalc];
                        int64_t t = kbnd + kbdr + 1;
                        int64_t icost = 6 * (int64_t)(nalc - kbdr) * t * t + 2 * (int64_t)kbdr * kbdr * kbdr;
                        if (icost < icmn) { icmn = icost; kx = k; kbdrx = kbdr; mside = 2; }
                    }
                }
                if (kx > 0) { mbnd = kx - 2; mbdr = kbdrx; }
            }
            if (mbnd >= 0) {
                /* bordered-band solution (3876-3897); band part of the inverse for pulls in the
                 * first iteration, complete inverse for the matrix update */
                int inv = 0;
                if (iter == 1) inv = 1;
                if (o->icalcm == 1) inv = 2;
                double evdmin = 0.0, evdmax = 0.0;
                orc_sqmibb(clmat + 1, blvec + 1, nalc, mbdr, mbnd, inv, &nrank, &evdmin, &evdmax, mside);
                if (evdmin > 0.0) cndl10 = log10f((float)(evdmax / evdmin));
                if (iter == 1) (*nband)++;
            } else
                orc_sqminv(clmat + 1, blvec + 1, nalc, &nrank);
            nan = 0;
            for (int k = 1; k <= nalc; k++) if (!(blvec[k] <= 0.0) && !(blvec[k] > 0.0)) nan++;
            summ = 0.0;
            suwt = 0.0;
            for (int e = 0; e < neq; e++) {
                int sL = eqs[3 * e], sG = eqs[3 * e + 1];
                double rmeas = D[sL], rerr = D[sG];
                double wght = 1.0 / (rerr * rerr);
                double rmloc = 0.0;
                for (int j = 1; j <= sG - sL - 1; j++) rmloc = rmloc + D[sL + j] * blvec[I[sL + j]];
                lcorr[e + 1] = rmloc;
                rmeas = rmeas - rmloc;
                double dchi2 = wght * rmeas * rmeas;
                double resid = rmeas;
                if (o->nloopn != 1 && iter != 1 && o->lhuber != 0) {
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
                suwt = suwt + rerr / down;
                if (iter == nter) D[sL] = rmeas;
                summ = summ + dchi2;
            }
            ndf = neq - nrank;
            resing = ((float)nweig - (float)suwt) / (float)nweig;
            if (nan > 0) { nrej[0]++; rejected = 1; break; }
            if (nrank != nalc) { nrej[1]++; rejected = 1; break; }
            if (cndl10 > o->cndlmx) { nrej[2]++; rejected = 1; break; }
            if (ndf <= 0) { nrej[3]++; rejected = 1; break; }
        }
        if (!rejected) {
            float chichi = orc_chindl(3, ndf) * (float)ndf;
            if (o->chicut >= 0.0f) {
                if (summ > (double)(o->chhuge * chichi)) { nrej[4]++; rejected = 1; }
                else if (o->chicut > 0.0f) {
                    float chlimt = o->chicut * chichi;
                    if (summ > (double)chlimt) { addsums(sums, (double)chlimt, ndf, dw1); nrej[5]++; rejected = 1; }
                }
            }
        }
        if (!rejected && o->lhuber > 1 && o->dwcut != 0.0f && resing > o->dwcut) {
            addsums(sums, summ, ndf, dw1);
            nrej[5]++;
            rejected = 1;
        }
        if (!rejected) {
            (*naccept)++;
            summ = 0.0;
            for (int e = 0; e < neq; e++) {
                int sL = eqs[3 * e], sG = eqs[3 * e + 1], lG = eqs[3 * e + 2];
                double rmeas = D[sL], rerr = D[sG];
                double wght = 1.0 / (rerr * rerr);
                double dchi2 = wght * rmeas * rmeas;
                if (o->nloopn != 1 && o->lhuber != 0) {
                    double resid = fabs(rmeas);
                    if (nter <= 3) {
                        if (resid > chuber * rerr) dchi2 = 2.0 * chuber * (resid / rerr - 0.5 * chuber);
                    } else {
                        double t = resid / rerr / cauchy;
                        dchi2 = log(1.0 + t * t) * cauchy * cauchy;
                    }
                    wght = lwgt[e + 1];
                }
                summ = summ + dchi2;
                for (int j = 1; j <= lG - sG; j++) {
                    int wj = sG + j;
                    int ivgbj = R->var_index[tix[wj]];
                    if (D[wj] == 0.0) continue;
                    if (ivgbj > 0) {
                        vec[ivgbj - 1] += dw1 * wght * rmeas * D[wj];
                        counter[ivgbj - 1] += 1;
                        if (o->icalcm == 1) {
                            int ije = back[ivgbj];
                            for (int k = 1; k <= j; k++) {
                                int ivgbk = R->var_index[tix[sG + k]];
                                if (ivgbk > 0) {
                                    int ike = back[ivgbk];
                                    int ia = ije > ike ? ije : ike, ib = ije < ike ? ije : ike;
                                    int ij = ib + (ia * ia - ia) / 2;
                                    upd[ij] = upd[ij] - dw1 * wght * D[wj] * D[sG + k];
                                }
                            }
                        }
                    }
                }
                if (o->icalcm != 1) continue;
                for (int j = 1; j <= lG - sG; j++) {
                    int ivgbj = R->var_index[tix[sG + j]];
                    if (ivgbj > 0) {
                        int ije = back[ivgbj];
                        for (int k = 1; k <= sG - sL - 1; k++) {
                            int ik = I[sL + k];
                            int jk = ik + (ije - 1) * nalc;
                            glm[jk] = glm[jk] + dw2 * wght * D[sG + j] * D[sL + k];
                        }
                    }
                }
            }
            addsums(sums, summ, ndf, dw1);
            if (o->icalcm == 1) {
                orc_dbavat(clmat + 1, glm + 1, upd + 1, nalc, nalg);
                /* mgupdt, full storage: globalMat(row offsets + col) -= sub */
                for (int a = 1; a <= nalg; a++) {
                    int va = gidx[a];
                    for (int b = 1; b <= a; b++) {
                        int vb = gidx[b];
                        sink_add(mat, va, vb, -upd[(a * a - a) / 2 + b]);
                    }
                }
            }
        }
        free(lcorr); free(lwgt); free(clmat); free(blvec); free(ibandh);
        if (upd) free(upd);
        if (glm) free(glm);
        if (o->icalcm == 1) for (int k = 1; k <= nalg; k++) back[gidx[k]] = 0;
        nalg = 0;
    }
out:
    if (rc == 20 && o->icalcm == 1) for (int k = 1; k <= nalg; k++) back[gidx[k]] = 0;
    free(D); free(tix); free(eqs); free(gidx);
    return rc == 2 ? 0 : rc;
}

/* One data pass (loopn, src/pede.f90:2479-2885, record part + pre-sigma diagonal 2711-2718).
 * mat: packed nvar(nvar+1)/2 (zeroed here when icalcm==1), vec/counter: nvar.
 * Records are processed in order on one thread, so sums are deterministic. */
int orc_loopn(const orc_records *R, const orc_pass_opts *o, const double *pre_weight, double *mat, double *vec,
              int64_t *counter, orc_pass_stats *st)
{
    int nvar = R->nvar;
    int64_t npk = ((int64_t)nvar * (nvar + 1)) / 2;
    for (int i = 0; i < nvar; i++) { vec[i] = 0.0; counter[i] = 0; }
    if (o->icalcm == 1 && mat) memset(mat, 0, sizeof(double) * npk);
    orc_sums sums = {0.0, 0, 0, 0.0};
    memset(st, 0, sizeof(*st));
    int *back = (int *)calloc(nvar + 2, sizeof(int));
    int rc = 0;
    orc_sink S = {mat, NULL, NULL, 0, 0, 0, NULL, NULL, NULL, NULL, NULL, NULL};
    for (int64_t r = 0; r < R->nrec; r++) {
        rc = loop_record(R, r, o, &S, vec, counter, &sums, st->nrej, &st->naccept, &st->nbandrec, back);
        if (rc) break;
    }
    free(back);
    if (o->icalcm == 1 && mat && pre_weight)
        for (int i = 1; i <= nvar; i++)
            if (pre_weight[i - 1] != 0.0) mat[IDX(i, i) - 1] += pre_weight[i - 1];
    st->chi2 = sums.chi_frac + (double)sums.chi_int;
    st->ndf = sums.ndf;
    st->wndf = sums.wndf;
    return rc;
}

/* Same pass accumulating the matrix into a caller-provided hash table (cap = power of two,
 * keys zeroed); returns 31 when the table overflows.  The pre-sigma diagonal is not added. */
int orc_loopn_hash(const orc_records *R, const orc_pass_opts *o, uint64_t *keys, double *vals, int64_t cap,
                   double *vec, int64_t *counter, orc_pass_stats *st, int64_t *used)
{
    int nvar = R->nvar;
    for (int i = 0; i < nvar; i++) { vec[i] = 0.0; counter[i] = 0; }
    orc_sums sums = {0.0, 0, 0, 0.0};
    memset(st, 0, sizeof(*st));
    int *back = (int *)calloc(nvar + 2, sizeof(int));
    int rc = 0;
    orc_sink S = {NULL, keys, vals, cap, 0, 0, NULL, NULL, NULL, NULL, NULL, NULL};
    for (int64_t r = 0; r < R->nrec; r++) {
        rc = loop_record(R, r, o, &S, vec, counter, &sums, st->nrej, &st->naccept, &st->nbandrec, back);
        if (rc) break;
        if (S.overflow) { rc = 31; break; }
    }
    free(back);
    st->chi2 = sums.chi_frac + (double)sums.chi_int;
    st->ndf = sums.ndf;
    st->wndf = sums.wndf;
    *used = S.used;
    return rc;
}

/* The same pass on nthreads OpenMP threads (records in dynamic chunks, as pede's loopbf loop),
 * the matrix accumulated into the group-block storage of the device (blk: nstored words,
 * zeroed here; gfirst/brow/bcol/boff as mpgpu_get_block_structure returns them), vector,
 * counters and sums per thread then added.  This is the CPU baseline of the timing, not a parity
 * reference: the order of the atomic additions varies between runs.  Returns 31 when an update
 * falls outside the block structure. */
int orc_loopn_blocks(const orc_records *R, const orc_pass_opts *o, int nthreads, int32_t ngroup, const int32_t *gfirst,
                     const int64_t *brow, const int32_t *bcol, const int64_t *boff, int64_t nstored, double *blk,
                     double *vec, int64_t *counter, orc_pass_stats *st)
{
    const int nvar = R->nvar;
    int32_t *group_of = (int32_t *)malloc(sizeof(int32_t) * (nvar + 2));
    for (int g = 0; g < ngroup; g++)
        for (int v = gfirst[g]; v < gfirst[g + 1]; v++) group_of[v] = g;
    memset(blk, 0, sizeof(double) * nstored);
    for (int i = 0; i < nvar; i++) { vec[i] = 0.0; counter[i] = 0; }
    memset(st, 0, sizeof(*st));
    int rc = 0, bad = 0;
    orc_sums tot = {0.0, 0, 0, 0.0};
#pragma omp parallel num_threads(nthreads)
    {
        double *tv = (double *)calloc(nvar + 1, sizeof(double));
        int64_t *tc = (int64_t *)calloc(nvar + 1, sizeof(int64_t));
        int *back = (int *)calloc(nvar + 2, sizeof(int));
        orc_sums sums = {0.0, 0, 0, 0.0};
        int64_t nrej[6] = {0, 0, 0, 0, 0, 0}, nacc = 0, nband = 0;
        orc_sink S = {NULL, NULL, NULL, 0, 0, 0, gfirst, group_of, bcol, brow, boff, blk};
        int lrc = 0;
#pragma omp for schedule(dynamic, 64)
        for (int64_t r = 0; r < R->nrec; r++) {
            if (lrc) continue;
            lrc = loop_record(R, r, o, &S, tv, tc, &sums, nrej, &nacc, &nband, back);
        }
#pragma omp critical
        {
            for (int i = 0; i < nvar; i++) { vec[i] += tv[i]; counter[i] += tc[i]; }
            for (int k = 0; k < 6; k++) st->nrej[k] += nrej[k];
            st->naccept += nacc;
            st->nbandrec += nband;
            tot.chi_frac += sums.chi_frac;
            tot.chi_int += sums.chi_int;
            tot.ndf += sums.ndf;
            tot.wndf += sums.wndf;
            if (lrc) rc = lrc;
            if (S.overflow) bad = 1;
        }
        free(tv); free(tc); free(back);
    }
    free(group_of);
    st->chi2 = tot.chi_frac + (double)tot.chi_int;
    st->ndf = tot.ndf;
    st->wndf = tot.wndf;
    return rc ? rc : (bad ? 31 : 0);
}

/* ------------------------------------------------------------------------------------------ */
/* Mille binary reader (C binary files: python/readMilleBinary.py:186-213; Fortran sequential: */
/* src/mille.f90:214-216, src/pede.f90:1827-1829).  Returns number of records or <0 on error.   */
/* Arrays are malloc'ed; values converted to double as pede does (src/pede.f90:1831-1834).     */
/* ------------------------------------------------------------------------------------------ */
int64_t orc_read_mille(const char *path, int fortran, int64_t **rec_ptr_out, int32_t **inder_out,
                       double **glder_out)
{
    FILE *f = fopen(path, "rb");
    if (!f) return -1;
    int64_t cap_r = 1024, cap_w = 1 << 16, nrec = 0, nwords = 0;
    int64_t *rp = malloc(sizeof(int64_t) * (cap_r + 1));
    int32_t *in = malloc(sizeof(int32_t) * cap_w);
    double *gd = malloc(sizeof(double) * cap_w);
    rp[0] = 0;
    for (;;) {
        int32_t marker = 0, nr = 0;
        if (fortran) {
            if (fread(&marker, 4, 1, f) != 1) break;
            if (fread(&nr, 4, 1, f) != 1) { nrec = -2; break; }
        } else {
            if (fread(&nr, 4, 1, f) != 1) break;
        }
        int dbl = nr < 0;
        int32_t nw = (dbl ? -nr : nr) / 2;
        if (nwords + nw > cap_w) {
            while (nwords + nw > cap_w) cap_w *= 2;
            in = realloc(in, sizeof(int32_t) * cap_w);
            gd = realloc(gd, sizeof(double) * cap_w);
        }
        if (dbl) {
            if (fread(gd + nwords, 8, nw, f) != (size_t)nw) { nrec = -2; break; }
        } else {
            float *tmp = malloc(sizeof(float) * (nw > 0 ? nw : 1));
            if (fread(tmp, 4, nw, f) != (size_t)nw) { free(tmp); nrec = -2; break; }
            for (int i = 0; i < nw; i++) gd[nwords + i] = (double)tmp[i];
            free(tmp);
        }
        if (fread(in + nwords, 4, nw, f) != (size_t)nw) { nrec = -2; break; }
        if (fortran) {
            int32_t m2;
            if (fread(&m2, 4, 1, f) != 1 || m2 != marker) { nrec = -2; break; }
        }
        nwords += nw;
        nrec++;
        if (nrec >= cap_r) { cap_r *= 2; rp = realloc(rp, sizeof(int64_t) * (cap_r + 1)); }
        rp[nrec] = nwords;
    }
    fclose(f);
    if (nrec < 0) { free(rp); free(in); free(gd); return nrec; }
    *rec_ptr_out = rp;
    *inder_out = in;
    *glder_out = gd;
    return nrec;
}

void orc_free(void *p) { free(p); }

/* ------------------------------------------------------------------------------------------ */
/* QL decomposition of the constraint matrix (elimination of constraints), one parameter block */
/* covering all npar parameters (as pede uses it for MINRES, src/pede.f90:7147-7151).          */
/* Inputs in pede's layout (1-based values, Fortran column-major shapes):                       */
/*   a     block-compressed constraint columns (src/pede.f90:1423-1460, feasma)                 */
/*   bcon  3 x (nblk+1): first constraint, first and last parameter of each constraint block    */
/*   rcon  4 x ncon: group range (first, last), storage range (first, last) of each constraint  */
/*         (src/pede.f90:1077-1371, prpcon)                                                     */
/* Outputs: matV (reflectors, same size as a), ioffRow (ncon+1), ioffPar, irng (2 x ncon,       */
/* non-zero range of each reflector), vecVk (element at kn), matL (ncon x ncon, column-major).  */
/* aligned = 0 restates the reference's start of L: the stored part of column k at parameters   */
/* kn..iplast goes to rows iclast-ncol..iclast (src/mpqldec.f90:300-305), which is the right     */
/* row only when the constraint block's storage ends at the last parameter (iplast = npar);     */
/* aligned = 1 puts parameter p at row p - (npar - ncon), the rows the Householder updates of L */
/* use (line 366), so that C^T = Q [0; L] holds for every layout.                                */
/* ------------------------------------------------------------------------------------------ */
void orc_qldecb(int npar, int ncon, int nblk, const double *a, const int *bcon, const int *rcon, double *matV,
                int64_t *ioffRow, int *ioffPar, int *irng, double *vecVk, double *matL, int aligned)
{
#define QB(i, j) bcon[3 * ((j) - 1) + (i) - 1]
#define QR(i, j) rcon[4 * ((j) - 1) + (i) - 1]
#define QV(i) matV[(i) - 1]
#define QL(i) matL[(i) - 1]
#define QN(i) vecN[(i) - 1]
#define QIR(a_, k) irng[2 * ((k) - 1) + (a_) - 1]
    /* src/mpqldec.f90:272-309 (one parameter block: iboff 0, iblast npar, icoff 0, iclast ncon) */
    const int iblast = npar, icoff = 0, iclast = ncon, iboff = 0;
    double *vecN = (double *)calloc((size_t)npar + 1, sizeof(double));
    memset(vecVk, 0, sizeof(double) * ncon);
    memset(matL, 0, sizeof(double) * (size_t)ncon * ncon);
    ioffRow[0] = 0;
    for (int ibcon = 1; ibcon <= nblk; ibcon++)
        for (int i = QB(1, ibcon); i <= QB(1, ibcon + 1) - 1; i++) {
            QIR(1, i) = QR(1, i);
            QIR(2, i) = QR(2, i);
            ioffPar[i - 1] = QR(3, i) - 1;
            ioffRow[i] = ioffRow[i - 1] + QR(4, i) - ioffPar[i - 1];
        }
    memcpy(matV, a, sizeof(double) * (size_t)ioffRow[ncon]);
    for (int k = icoff + 1; k <= iclast; k++) {
        int kn = iblast + k - iclast;
        int64_t ioff1 = ioffRow[k - 1];
        int npb = (int)(ioffRow[k] - ioff1);
        int ncol = ioffPar[k - 1] + npb - kn;
        if (ncol >= 0) {
            int64_t ioff3 = (int64_t)(k - 1) * ncon;
            int row0 = aligned ? k - icoff : iclast - ncol - icoff;
            for (int t = 0; t <= ncol; t++) QL(ioff3 + row0 + t) = QV(ioff1 + npb - ncol + t);
        }
    }
    /* Householder procedure, src/mpqldec.f90:311-385 */
    int ibcon = nblk, k1 = QB(1, ibcon);
    for (int k = iclast; k >= icoff + 1; k--) {
        int kn = iblast + k - iclast;
        if (k < k1) { ibcon--; k1 = QB(1, ibcon); }
        int ipoff = ioffPar[k - 1];
        int ifirst = ipoff + 1;
        if (ifirst > kn) continue;
        int64_t ioff1 = ioffRow[k - 1];
        int npb = (int)(ioffRow[k] - ioff1);
        int iplast = ipoff + npb;
        int ilast = iplast < kn ? iplast : kn;
        int ncol = ilast - ipoff;
        int in0 = iblast + k1 - iclast;
        for (int t = in0; t <= kn; t++) QN(t) = 0.0;
        for (int t = 0; t < ncol; t++) QN(ifirst + t) = QV(ioff1 + 1 + t);
        double nrm = 0.0;
        for (int t = ifirst; t <= ilast; t++) nrm += QN(t) * QN(t);
        nrm = sqrt(nrm);
        if (nrm == 0.0) continue;
        if (QN(kn) >= 0.0) QN(kn) += nrm; else QN(kn) -= nrm;
        double s2 = 0.0;
        for (int t = ifirst; t <= ilast; t++) s2 += QN(t) * QN(t);
        if (ilast < kn) {
            nrm = sqrt(s2 + QN(kn) * QN(kn));
            for (int t = ifirst; t <= ilast; t++) QN(t) /= nrm;
            QN(kn) /= nrm;
        } else {
            nrm = sqrt(s2);
            for (int t = ifirst; t <= ilast; t++) QN(t) /= nrm;
        }
        int64_t ioff3 = (int64_t)(k1 - 2) * ncon;
        for (int i = k1; i <= k; i++) {
            ioff3 += ncon;
            if (QIR(1, k) > QIR(2, i)) continue;
            int64_t ioff2 = ioffRow[i - 1] + ioffPar[k - 1] - ioffPar[i - 1];
            double sp = 0.0;
            for (int t = 0; t < ncol; t++) sp += QN(ifirst + t) * QV(ioff2 + 1 + t);
            if (sp != 0.0) {
                for (int t = 0; t < ncol; t++) QV(ioff2 + 1 + t) -= 2.0 * QN(ifirst + t) * sp;
                int in1 = iblast + i - iclast;
                for (int t = 0; t <= k - i; t++) QL(ioff3 + i - icoff + t) -= 2.0 * QN(in1 + t) * sp;
                if (QIR(1, k) < QIR(1, i)) QIR(1, i) = QIR(1, k);
                if (QIR(2, k) > QIR(2, i)) QIR(2, i) = QIR(2, k);
            }
        }
        vecVk[icoff + k - 1] = QN(kn);
        ifirst = QIR(1, k);
        ilast = QIR(2, k) < kn - 1 ? QIR(2, k) : kn - 1;
        ncol = ilast - ifirst + 1;
        for (int t = 0; t < ncol; t++) QV(ioff1 + 1 + t) = QN(ifirst + t);
        for (int t = ncol > 0 ? ncol : 0; t < npb; t++) QV(ioff1 + 1 + t) = 0.0;
        QIR(1, k) = ifirst - iboff;
        QIR(2, k) = ilast - iboff;
    }
    free(vecN);
#undef QB
#undef QR
#undef QV
#undef QL
#undef QN
#undef QIR
}

/* Q x (t = 0) or Q^T x (t = 1) for m columns of x (npar each), Q = H_ncon ... H_1
   (src/mpqldec.f90:404-454, qlmlq, single block) */
void orc_qlmlq(int npar, int ncon, const double *matV, const int64_t *ioffRow, const int *irng, const double *vecVk,
               double *x, int m, int t)
{
    for (int j = 1; j <= ncon; j++) {
        int k = t ? ncon + 1 - j : j;
        int kn = npar + k - ncon;
        int ifirst = irng[2 * (k - 1)], ilast = irng[2 * (k - 1) + 1];
        const double *v = matV + ioffRow[k - 1];   /* v[p - ifirst] for p in [ifirst, ilast] */
        double vk = vecVk[k - 1];
        for (int c = 0; c < m; c++) {
            double *xc = x + (int64_t)c * npar;
            double sp = 0.0;
            for (int p = ifirst; p <= ilast; p++) sp += v[p - ifirst] * xc[p - 1];
            sp += vk * xc[kn - 1];
            for (int p = ifirst; p <= ilast; p++) xc[p - 1] -= 2.0 * v[p - ifirst] * sp;
            xc[kn - 1] -= 2.0 * vk * sp;
        }
    }
}

/* y from L^T y = d by backward substitution (src/mpqldec.f90:991-1014, qlbsub, single block) */
void orc_qlbsub(int ncon, const double *matL, const double *d, double *y)
{
    for (int k = ncon; k >= 1; k--) {
        double s = 0.0;
        for (int r = k + 1; r <= ncon; r++) s += matL[(int64_t)(k - 1) * ncon + r - 1] * y[r - 1];
        y[k - 1] = (d[k - 1] - s) / matL[(int64_t)(k - 1) * ncon + k - 1];
    }
}

/* Band part of the similarity transform Q^T A Q for the preconditioner (src/mpqldec.f90:711-945,
   qlpssq, single block), restated with its own arithmetic: the A*v products of the original
   matrix (aprod adds A[:, is..ie] x[is..ie], 1-based, to y), products v^T v'' above 16 eps,
   and the corrections to A*v carried as linear combinations of the reflectors.  B holds the
   band (width m including the diagonal) in pede's variable-band layout and is updated in place. */
typedef void (*orc_aprods)(void *ctx, int n, const double *x, int is, int ie, double *y);

void orc_qlpssq(int npar, int ncon, const double *matV, const int64_t *ioffRow, const int *irng, const double *vecVk,
                orc_aprods aprod, void *ctx, double *B, int m, int t)
{
    double *vecN = (double *)calloc((size_t)npar + 1, sizeof(double));
    double *vecAv = (double *)calloc((size_t)npar + 1, sizeof(double));
    double *vtvp = (double *)calloc((size_t)ncon * ncon + 1, sizeof(double));
    double *vtAvp = (double *)calloc((size_t)ncon * ncon + 1, sizeof(double));
    double *coef = (double *)calloc((size_t)ncon * ncon + 1, sizeof(double));
    int *irc = (int *)calloc(2 * (size_t)ncon + 2, sizeof(int));
    const int mbnd = m - 1 > 0 ? m - 1 : 0;
    const double eps = DBL_EPSILON;
#define PN(i) vecN[(i) - 1]
#define PA(i) vecAv[(i) - 1]
#define PV(i) matV[(i) - 1]
#define PIR(a_, k) irng[2 * ((k) - 1) + (a_) - 1]
#define PIC(a_, k) irc[2 * ((k) - 1) + (a_) - 1]
#define KOF(j) (t ? ncon + 1 - (j) : (j))
    /* initial A*v of every reflector, products, update with the initial A*v (lines 792-876) */
    for (int j = 1; j <= ncon; j++) {
        int k = KOF(j), kn = npar + k - ncon;
        int64_t ioff1 = (int64_t)(k - 1) * ncon;
        PIC(1, k) = ncon;
        PIC(2, k) = 1;
        int ifirst = PIR(1, k), ilast = PIR(2, k);
        memset(vecN, 0, sizeof(double) * npar);
        for (int p = ifirst; p <= ilast; p++) PN(p) = PV(ioffRow[k - 1] + 1 + p - ifirst);
        PN(kn) = vecVk[k - 1];
        memset(vecAv, 0, sizeof(double) * npar);
        aprod(ctx, npar, vecN, ifirst, ilast, vecAv);
        aprod(ctx, npar, vecN, kn, kn, vecAv);
        for (int j2 = j + 1; j2 <= ncon; j2++) {
            int k2 = KOF(j2), kn2 = npar + k2 - ncon;
            int64_t ioff2 = (int64_t)(k2 - 1) * ncon;
            int ifirst2 = PIR(1, k2), ilast2 = PIR(2, k2);
            double v2kn = 0.0;
            if (kn >= ifirst2 && kn <= ilast2) v2kn = PV(ioffRow[k2 - 1] + 1 + kn - ifirst2);
            int l1 = ifirst > ifirst2 ? ifirst : ifirst2, l2 = ilast < ilast2 ? ilast : ilast2;
            double s = PN(kn2) * vecVk[k2 - 1] + PN(kn) * v2kn;
            if (l1 <= l2) {
                double dd = 0.0;
                for (int p = l1; p <= l2; p++) dd += PN(p) * PV(ioffRow[k2 - 1] + 1 + p - ifirst2);
                s = s + dd;
            }
            if (fabs(s) > 16.0 * eps) { vtvp[ioff1 + k2 - 1] = s; vtvp[ioff2 + k - 1] = s; }
        }
        vtvp[ioff1 + k - 1] = 1.0;
        for (int j2 = 1; j2 <= j; j2++) {
            int k2 = KOF(j2), kn2 = npar + k2 - ncon;
            int ifirst2 = PIR(1, k2), ilast2 = PIR(2, k2);
            double dd = 0.0;
            for (int p = ifirst2; p <= ilast2; p++) dd += PA(p) * PV(ioffRow[k2 - 1] + 1 + p - ifirst2);
            vtAvp[ioff1 + k2 - 1] = vecVk[k2 - 1] * PA(kn2) + dd;
        }
        int64_t ib = 0;
        double vtAv = vtAvp[ioff1 + k - 1];
        for (int i = 1; i <= kn; i++)
            for (int l = (i - mbnd > 1 ? i - mbnd : 1); l <= i; l++, ib++)
                B[ib] = B[ib] + 2.0 * ((2.0 * PN(i) * vtAv - PA(i)) * PN(l) - PA(l) * PN(i));
        for (int i = kn + 1; i <= npar; i++)
            for (int l = (i - mbnd > 1 ? i - mbnd : 1); l <= i; l++, ib++) B[ib] = B[ib] - 2.0 * PA(i) * PN(l);
    }
    /* corrections of A*v as linear combinations of the reflectors (lines 878-935) */
    for (int j = 1; j <= ncon; j++) {
        int k = KOF(j), kn = npar + k - ncon;
        int64_t ioff1 = (int64_t)(k - 1) * ncon;
        int ifirst = PIR(1, k), ilast = PIR(2, k);
        memset(vecN, 0, sizeof(double) * npar);
        for (int p = ifirst; p <= ilast; p++) PN(p) = PV(ioffRow[k - 1] + 1 + p - ifirst);
        PN(kn) = vecVk[k - 1];
        int l1 = PIC(1, k), l2 = PIC(2, k);
        double dd = 0.0;
        for (int q = l1; q <= l2; q++) dd += coef[ioff1 + q - 1] * vtvp[ioff1 + q - 1];
        double vtAv = vtAvp[ioff1 + k - 1] + dd;
        memset(vecAv, 0, sizeof(double) * npar);
        for (int k2 = l1; k2 <= l2; k2++) {
            double cf = coef[ioff1 + k2 - 1];
            if (cf == 0.0) continue;
            int kn2 = npar + k2 - ncon, ifirst2 = PIR(1, k2), ilast2 = PIR(2, k2);
            for (int p = ifirst2; p <= ilast2; p++) PA(p) = PA(p) + cf * PV(ioffRow[k2 - 1] + 1 + p - ifirst2);
            PA(kn2) = PA(kn2) + cf * vecVk[k2 - 1];
        }
        int64_t ib = 0;
        for (int i = 1; i <= kn; i++)
            for (int l = (i - mbnd > 1 ? i - mbnd : 1); l <= i; l++, ib++)
                B[ib] = B[ib] + 2.0 * ((2.0 * PN(i) * vtAv - PA(i)) * PN(l) - PA(l) * PN(i));
        for (int i = kn + 1; i <= npar; i++)
            for (int l = (i - mbnd > 1 ? i - mbnd : 1); l <= i; l++, ib++) B[ib] = B[ib] - 2.0 * PA(i) * PN(l);
        for (int j2 = j + 1; j2 <= ncon; j2++) {
            int k2 = KOF(j2);
            int64_t ioff2 = (int64_t)(k2 - 1) * ncon;
            double s = vtvp[ioff1 + k2 - 1];
            int m1 = PIC(1, k2), m2 = PIC(2, k2);
            double vta = vtAvp[ioff2 + k - 1];
            if (m1 <= m2) {
                double d2 = 0.0;
                for (int q = m1; q <= m2; q++) d2 += coef[ioff2 + q - 1] * vtvp[ioff1 + q - 1];
                vta = vta + d2;
            }
            if (k < m1) m1 = k;
            if (k > m2) m2 = k;
            coef[ioff2 + k - 1] = coef[ioff2 + k - 1] + 2.0 * (2.0 * vtAv * s - vta);
            if (s != 0.0) {
                if (PIC(1, k) < m1) m1 = PIC(1, k);
                if (PIC(2, k) > m2) m2 = PIC(2, k);
                for (int q = m1; q <= m2; q++) coef[ioff2 + q - 1] = coef[ioff2 + q - 1] - 2.0 * coef[ioff1 + q - 1] * s;
            }
            PIC(1, k2) = m1;
            PIC(2, k2) = m2;
        }
    }
#undef PN
#undef PA
#undef PV
#undef PIR
#undef PIC
#undef KOF
    free(vecN); free(vecAv); free(vtvp); free(vtAvp); free(coef); free(irc);
}
