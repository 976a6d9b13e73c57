// Global-solve kernels: packed-matrix utilities, Gauss-Jordan inversion, symmetric
// matrix-vector products for packed and group-block storage, MINRES vector algebra and the
// MINRES preconditioners.
#include "mpgpu_device.h"
#include "mpgpu_internal.h"
#include <climits>
#include <mutex>
#include <float.h>

namespace mpgpu {

#define TPB 256

__device__ __forceinline__ int64_t pk64s(int64_t i, int64_t j) { return (i * i - i) / 2 + j - 1; }

// ---------------------------------------------------------------------------------------------
// element updates (mupdat, src/pede.f90:3134-3204); i,j 1-based in the active system
// ---------------------------------------------------------------------------------------------
__device__ int32_t find_blk(const int64_t *brow, const int32_t *bcol, int ga, int gb)
{
    int64_t lo = brow[ga], hi = brow[ga + 1] - 1;
    while (lo <= hi) {
        int64_t mid = (lo + hi) >> 1;
        int c = bcol[mid];
        if (c == gb) return (int32_t)mid;
        if (c < gb) lo = mid + 1; else hi = mid - 1;
    }
    return -1;
}

__global__ void k_mupdat(DevModel M, int32_t n, const int32_t *ii, const int32_t *jj, const double *vv)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int i = ii[t], j = jj[t];
    double v = vv[t];
    if (i <= 0 || j <= 0 || v == 0.0) return;
    int a = max(i, j), b = min(i, j);
    if (!M.sparse) { atomicAdd(&M.mat[pk64s(a, b)], v); return; }
    if (a > M.nvar) return; // Lagrange rows live in the constraint CSR for sparse storage
    int ga = M.group_of[a], gb = M.group_of[b];
    int32_t blk = find_blk(M.brow, M.bcol, ga, gb);
    if (blk < 0) return;
    int fa = M.gfirst[ga], fb = M.gfirst[gb];
    int nc = M.gfirst[gb + 1] - fb;
    atomicAdd(&M.mat[M.boff[blk] + (int64_t)(a - fa) * nc + (b - fb)], v);
}

__global__ void k_get_diag(DevModel M, int32_t n, double *d)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (i > n) return;
    if (!M.sparse) { d[i - 1] = M.mat[pk64s(i, i)]; return; }
    int g = M.group_of[i];
    int64_t blk = M.brow[g + 1] - 1; // diagonal block is last in its row
    int f = M.gfirst[g];
    int nc = M.gfirst[g + 1] - f;
    d[i - 1] = M.mat[M.boff[blk] + (int64_t)(i - f) * nc + (i - f)];
}

// band of the parameter block into pede's variable-band layout (matPreCond, india pointers,
// src/pede.f90:2638-2653): row i holds columns max(1,i-mbandw+1)..i ending at india(i)
__global__ void k_get_band(DevModel M, int32_t n, int32_t mbandw, const int32_t *india, double *c)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (i > n) return;
    int ia = india[i - 1];
    for (int j = max(1, i - mbandw + 1); j <= i; j++) {
        double v = 0.0;
        if (!M.sparse) v = M.mat[pk64s(i, j)];
        else {
            int ga = M.group_of[i], gb = M.group_of[j];
            int32_t blk = find_blk(M.brow, M.bcol, ga, gb);
            if (blk >= 0) {
                int fa = M.gfirst[ga], fb = M.gfirst[gb];
                int nc = M.gfirst[gb + 1] - fb;
                v = M.mat[M.boff[blk] + (int64_t)(i - fa) * nc + (j - fb)];
            }
        }
        c[ia - i + j - 1] = v;
    }
}

// ---------------------------------------------------------------------------------------------
// deterministic reductions
// ---------------------------------------------------------------------------------------------
__global__ void k_partial_dot(int64_t n, const double *a, const double *b, double *part)
{
    __shared__ double s[TPB];
    double acc = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB)
        acc += a[i] * (b ? b[i] : 1.0);
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int o = TPB / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

__global__ void k_final_sum(int n, const double *part, double *out)
{
    __shared__ double s[TPB];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += TPB) acc += part[i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int o = TPB / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = s[0];
}

// dot product in one launch: each workgroup's partial sum as k_partial_dot, then the last
// workgroup to finish (a counter in part[DOT_CNT], zero between calls) adds the partial sums in
// k_final_sum's order, so the result is bit-identical to the two-launch form
#define DOT_CNT 2047
__global__ __launch_bounds__(TPB) void k_dot_fused(int64_t n, const double *a, const double *b, double *part, double *out)
{
    __shared__ double s[TPB];
    __shared__ int last;
    double acc = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB)
        acc += a[i] * (b ? b[i] : 1.0);
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int o = TPB / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    unsigned int *cnt = (unsigned int *)(part + DOT_CNT);
    if (threadIdx.x == 0) {
        __hip_atomic_store(&part[blockIdx.x], s[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned int done = __hip_atomic_fetch_add(cnt, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        last = done == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    double acc2 = 0.0;
    for (int i = threadIdx.x; i < (int)gridDim.x; i += TPB)
        acc2 += __hip_atomic_load(&part[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s[threadIdx.x] = acc2;
    __syncthreads();
    for (int o = TPB / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        *out = s[0];
        __hip_atomic_store(cnt, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// pass statistics: chi2 (double), ndf (int64), status histogram [0..20]
__global__ void k_pass_stats(int64_t n, const double *chi2, const int32_t *ndf, const int8_t *status, double *part,
                             long long *ndf_sum, unsigned long long *hist)
{
    __shared__ double s[TPB];
    __shared__ unsigned long long h[24];
    if (threadIdx.x < 24) h[threadIdx.x] = 0;
    __syncthreads();
    double acc = 0.0;
    long long nd = 0;
    for (int64_t i = blockIdx.x * (int64_t)TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB) {
        acc += chi2[i];
        nd += ndf[i];
        int st = status[i];
        atomicAdd(&h[st < 0 || st > 23 ? 23 : st], 1ull);
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int o = TPB / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
    atomicAdd((unsigned long long *)ndf_sum, (unsigned long long)nd);
    __syncthreads();
    if (threadIdx.x < 24 && h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], h[threadIdx.x]);
}

// ---------------------------------------------------------------------------------------------
// Gauss-Jordan inversion of the packed matrix, restating sqminl (src/mpnum.f90:230-355):
// per pivot step one single-block kernel picks the pivot exactly as the sequential scan does
// (first unused j with |v_jj| > max(|v_kk|, eps*|diag_j|), i.e. the largest |v_jj| above its
// threshold, ties to the lowest index) and stages the scaled pivot column; a grid-wide kernel
// then applies the rank-1 elimination with the reference's operand order.
// state: used[n] (0 free / 1 used), info[0]=k, info[1]=nrank, info[2]=stop
// ---------------------------------------------------------------------------------------------
__global__ void k_gj_init(int n, const double *v, double *diag, int *used, int *info)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        diag[i] = fabs(v[pk64s(i + 1, i + 1)]);
        used[i] = 0;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) { info[0] = 0; info[1] = 0; info[2] = 0; }
}

__global__ __launch_bounds__(1024) void k_gj_pivot(int n, double *v, const double *diag, int *used, int *info,
                                                   double *vk, double *b, double eps)
{
    __shared__ double sv[1024];
    __shared__ int si[1024];
    if (info[2]) return;
    double best = -1.0;
    int bi = 0x7fffffff;
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        if (used[j]) continue;
        double a = fabs(v[pk64s(j + 1, j + 1)]);
        if (a > eps * diag[j] && (a > best || (a == best && j < bi))) { best = a; bi = j; }
    }
    sv[threadIdx.x] = best;
    si[threadIdx.x] = bi;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            double a = sv[threadIdx.x + o];
            int ai = si[threadIdx.x + o];
            if (a > sv[threadIdx.x] || (a == sv[threadIdx.x] && ai < si[threadIdx.x])) { sv[threadIdx.x] = a; si[threadIdx.x] = ai; }
        }
        __syncthreads();
    }
    int k = si[0];
    if (sv[0] < 0.0 || k == 0x7fffffff) { // no pivot: stop; clearing happens in k_gj_clear
        if (threadIdx.x == 0) { info[0] = 0; info[2] = 1; }
        return;
    }
    const int K = k + 1;
    const int64_t kk = pk64s(K, K);
    const double vkk = 1.0 / v[kk];
    // scaled column k: vk(j) = v(jk)*vkk, vk(k) = 0  (sqminl lines 299-311)
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        int J = j + 1;
        if (J == K) vk[j] = 0.0;
        else {
            int64_t jk = J < K ? pk64s(K, J) : pk64s(J, K);
            vk[j] = v[jk] * vkk;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        used[k] = 1;
        info[0] = K;
        info[1] += 1;
        b[k] = b[k] * vkk;
        vk[n] = vkk;
    }
}

// elimination (sqminl lines 312-319): v(jl) -= vk(l) * (vk(j)/vkk) for j,l != k;
// column k takes vk, v(kk) = -vkk; b(j) -= b(k)*(vk(j)/vkk)
__global__ void k_gj_update(int n, double *v, const double *vk, double *b, const int *info)
{
    if (info[2]) return;
    const int K = info[0];
    const double vkk = vk[n];
    const int64_t total = ((int64_t)n * (n + 1)) / 2;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        int64_t J = (int64_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5) + 1;
        while ((J * J - J) / 2 > t) J--;
        while ((J * J + J) / 2 <= t) J++;
        int64_t L = t - (J * J - J) / 2 + 1;
        if (J == K && L == K) v[t] = -vkk;
        else if (J == K) v[t] = vk[L - 1];
        else if (L == K) v[t] = vk[J - 1];
        else {
            double vjk = vk[J - 1] / vkk;
            v[t] = v[t] - vk[L - 1] * vjk;
        }
    }
    if (blockIdx.x == 0) {
        const double bk = b[K - 1];
        for (int j = threadIdx.x; j < n; j += blockDim.x) {
            if (j + 1 == K) continue;
            double vjk = vk[j] / vkk;
            b[j] = b[j] - bk * vjk;
        }
    }
}

// rank defect: clear rows/columns of unused pivots (sqminl lines 320-331), then negate all
__global__ void k_gj_finish(int n, double *v, double *b, const int *used, const int *info)
{
    const int64_t total = ((int64_t)n * (n + 1)) / 2;
    const bool clear = info[2] != 0;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        int64_t J = (int64_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5) + 1;
        while ((J * J - J) / 2 > t) J--;
        while ((J * J + J) / 2 <= t) J++;
        int64_t L = t - (J * J - J) / 2 + 1;
        double x = v[t];
        if (clear && !used[J - 1] && !used[L - 1]) x = 0.0;
        v[t] = -x;
    }
    if (clear && blockIdx.x == 0)
        for (int j = threadIdx.x; j < n; j += blockDim.x) if (!used[j]) b[j] = 0.0;
}

// ---------------------------------------------------------------------------------------------
// y = A x for the packed full matrix (avprd0 full storage, src/pede.f90:5069-5232):
// 64x64 tiles of the lower triangle; each stored element serves both halves.
// ---------------------------------------------------------------------------------------------
// Rows are taken in blocks of 64 (block b: rows 64b .. 64b+63, packed rows contiguous); a
// workgroup (b, s) stages the 64 x 64 tiles of columns 64 * (16 s .. 16 s + 15) of its row block
// through LDS (read once, coalesced along the packed rows) and forms from each tile its row
// products (summed over the workgroup's tiles into yr[s][row]) and its transposed products
// (z[b][col], the part of column col from the rows of block b).  A second kernel adds, in a
// fixed order, each row's partial sums over s and each column's over b: no atomics, the same
// result on every run.
#define ST 64
#define SSEG 16
__global__ __launch_bounds__(256) void k_psymv_tiles(int n, const double *a, const double *x, double *yr, double *z)
{
    const int nb = (n + ST - 1) / ST;
    const int b = nb - 1 - (int)blockIdx.x;         // the longest row blocks first
    const int sg = blockIdx.y;
    const int t0 = sg * SSEG, t1 = min(t0 + SSEG, b + 1);
    if (t0 > b) return;
    __shared__ double tl[ST][ST + 1];
    __shared__ double xr[ST], xc[ST];
    __shared__ double rp[4][ST];
    const int r0 = b * ST;
    const int tid = threadIdx.x;
    if (tid < ST) xr[tid] = r0 + tid < n ? x[r0 + tid] : 0.0;
    double rsum = 0.0;                  // thread (row = tid & 63, quarter = tid >> 6)
    const int row = tid & (ST - 1), qt = tid >> 6;
    double *zb = z + 32 * (int64_t)b * (b + 1);   // z block b: columns 0 .. 64 b + 63
    for (int tc = t0; tc < t1; tc++) {
        const int c0 = tc * ST;
        __syncthreads();
        // stage the tile: element (i, j) of rows r0.., columns c0.. (j <= i), zero elsewhere
        for (int e = tid; e < ST * ST; e += 256) {
            const int i = e >> 6, j = e & (ST - 1);
            const int gi = r0 + i, gj = c0 + j;
            tl[i][j] = (gi < n && gj <= gi) ? a[(int64_t)gi * (gi + 1) / 2 + gj] : 0.0;
        }
        if (tid < ST) xc[tid] = c0 + tid < n ? x[c0 + tid] : 0.0;
        __syncthreads();
        // row products: row `row`, columns 16 qt .. 16 qt + 15
        double sr = 0.0;
#pragma unroll
        for (int j = 16 * qt; j < 16 * qt + 16; j++) sr += tl[row][j] * xc[j];
        rsum += sr;
        // transposed products: column `row` of the tile, rows 16 qt .. 16 qt + 15, strictly below
        // the diagonal of the matrix
        double sc = 0.0;
#pragma unroll
        for (int i = 16 * qt; i < 16 * qt + 16; i++) sc += (r0 + i > c0 + row ? tl[i][row] : 0.0) * xr[i];
        rp[qt][row] = sc;
        __syncthreads();
        if (tid < ST) zb[c0 + tid] = rp[0][tid] + rp[1][tid] + rp[2][tid] + rp[3][tid];
    }
    __syncthreads();
    rp[qt][row] = rsum;
    __syncthreads();
    if (tid < ST && r0 + tid < n) yr[(int64_t)sg * n + r0 + tid] = rp[0][tid] + rp[1][tid] + rp[2][tid] + rp[3][tid];
}
// y_i = sum_s yr[s][i] (segments s covering row block i/64) + sum_{b >= i/64} z[b][i]
__global__ void k_psymv_sum(int n, const double *yr, const double *z, double *y)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int nb = (n + ST - 1) / ST;
    const int bi = i / ST;
    double s = 0.0;
    for (int sg = 0; sg * SSEG <= bi; sg++) s += yr[(int64_t)sg * n + i];
    for (int b = bi; b < nb; b++) s += z[32 * (int64_t)b * (b + 1) + i];
    y[i] = s;
}

int64_t packed_symv_scratch(int n)
{
    const int64_t nb = (n + ST - 1) / ST;
    const int64_t nseg = (nb + SSEG - 1) / SSEG;
    return nseg * n + 32 * nb * (nb + 1);
}

// ---------------------------------------------------------------------------------------------
// Allocation-free product for callers that cannot allocate device memory: the partial sums of
// the packed product live in a file-scope device array of fixed size (enough for matrices of
// NOALLOC_MAX_N rows), one copy per GPU, shared by every call on that GPU.  Its device address
// is looked up once per device, with that device current; kernels get that address, never the
// array's host-side name.  Each call fully overwrites the part of the array it reads, so no
// clearing is needed; calls that overlap on one GPU race for the array.
// ---------------------------------------------------------------------------------------------
#define NOALLOC_MAX_N 16384
#define NOALLOC_WORDS 2367488   // packed_symv_scratch(16384): 16 * 16384 + 32 * 256 * 257
static __device__ __align__(256) double scratch[NOALLOC_WORDS];

int noalloc_max_n() { return NOALLOC_MAX_N; }

static std::mutex noalloc_mu;
static double *noalloc_ptr[256];

// device address of this GPU's array (device must be current); nullptr on failure
static double *noalloc_scratch(int device)
{
    if (device < 0 || device >= 256) return nullptr;
    std::lock_guard<std::mutex> lk(noalloc_mu);
    if (!noalloc_ptr[device]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
        noalloc_ptr[device] = (double *)p;
    }
    return noalloc_ptr[device];
}

void l_packed_symv(int n, const double *a, const double *x, double *y, double *scratch, hipStream_t s);
// y = A x for the packed matrix of n <= NOALLOC_MAX_N rows, x and y device arrays
hipError_t noalloc_packed_symv(int device, int n, const double *a, const double *x, double *y, hipStream_t s)
{
    if (n > NOALLOC_MAX_N || packed_symv_scratch(n) > NOALLOC_WORDS) return hipErrorInvalidValue;
    double *sc = noalloc_scratch(device);
    if (!sc) return hipErrorInvalidSymbol;
    l_packed_symv(n, a, x, y, sc, s);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------------------
// y = A x for the group-block storage (avprd0 sparse storage, src/pede.f90:5069-5232).
// One thread per (block, row) computes the row part and pushes the transposed part.
// Diagonal blocks hold their lower triangle only.
// ---------------------------------------------------------------------------------------------
__global__ void k_block_symv(int32_t ngroup, const int64_t *brow, const int32_t *bcol, const int64_t *boff,
                             const int32_t *gfirst, const int32_t *brg, int64_t nblock, const double *a,
                             const double *x, double *y)
{
    int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    // t enumerates (block, row-in-block) with at most 8 rows per block slot
    int64_t blk = t >> 3;
    int rr = (int)(t & 7);
    if (blk >= nblock) return;
    int ga = brg[blk], gb = bcol[blk];
    int fa = gfirst[ga] - 1, fb = gfirst[gb] - 1;
    int nr = gfirst[ga + 1] - 1 - fa, nc = gfirst[gb + 1] - 1 - fb;
    const double *B = a + boff[blk];
    for (int r = rr; r < nr; r += 8) {
        double s = 0.0;
        const double xr = x[fa + r];
        int cmax = (ga == gb) ? r + 1 : nc;
        for (int c = 0; c < cmax; c++) {
            double v = B[(int64_t)r * nc + c];
            s += v * x[fb + c];
            if (ga != gb || c != r) atomicAdd(&y[fb + c], v * xr);
        }
        atomicAdd(&y[fa + r], s);
    }
}

// Symmetric product with the group-block matrix in two passes without atomics.
// Rows pass: one wavefront per block row ga owns y[rows of ga] = sum over the stored blocks
// (ga, gb <= ga) of B x[gb], the diagonal block used with both of its triangles.
// Columns pass: one wavefront per column group gb adds sum over the blocks (ga > gb) of
// B^T x[ga] to y[rows of gb] (column lists ccol_ptr / ccol_blk built with the structure).
// Lanes work in groups of 8 per block (lane = block slot * 8 + row or column), so that one
// wave instruction reads eight blocks' rows at once; partial sums are reduced over the slots.
__device__ __forceinline__ double sum_slots(double v)
{
    // reduce over the 8 block slots (lanes with equal lane % 8)
    for (int o = 32; o >= 8; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
// row R of block B (nr x nc, row-major; the diagonal block holds its lower triangle and is used
// with both) times x[fb .. fb+nc): loads of up to 8 columns issued together
__device__ __forceinline__ double bsymv_row(const double *B, int nc, int R, bool diag, const double *xb)
{
    double sv = 0.0;
    if (nc <= 8) {
        double v[8], xv[8];
#pragma unroll
        for (int c = 0; c < 8; c++) {
            const int cc = c < nc ? c : nc - 1;
            v[c] = (diag && cc > R) ? B[(int64_t)cc * nc + R] : B[(int64_t)R * nc + cc];
            xv[c] = xb[cc];
        }
#pragma unroll
        for (int c = 0; c < 8; c++)
            if (c < nc) sv += v[c] * xv[c];
        return sv;
    }
    for (int c = 0; c < nc; c++) sv += ((diag && c > R) ? B[(int64_t)c * nc + R] : B[(int64_t)R * nc + c]) * xb[c];
    return sv;
}
__global__ __launch_bounds__(256) void k_bsymv_rows(int32_t ngroup, const int64_t *brow, const int32_t *bcf,
                                                    const int64_t *boff, const int32_t *gfirst, const double *a,
                                                    const double *x, double *y)
{
    const int ga = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63, slot = lane >> 3, rl = lane & 7;
    if (ga >= ngroup) return;
    const int fa = gfirst[ga] - 1, nr = gfirst[ga + 1] - 1 - fa;
    const int64_t b0 = brow[ga], b1 = brow[ga + 1];
    for (int r0 = 0; r0 < nr; r0 += 8) {
        const int R = min(r0 + rl, nr - 1);   // lanes past the last row repeat it (result unused)
        double acc0 = 0.0, acc1 = 0.0;
        int64_t bk = b0 + slot;
        // two blocks per slot and step; the next step's block descriptors are read one step ahead
        int2 d0 = make_int2(0, 1), d1 = make_int2(0, 1);
        int64_t o0 = 0, o1 = 0;
        if (bk + 8 < b1) {
            d0 = *(const int2 *)(bcf + 2 * bk);
            d1 = *(const int2 *)(bcf + 2 * bk + 16);
            o0 = boff[bk];
            o1 = boff[bk + 8];
        }
        for (; bk + 8 < b1; bk += 16) {
            const int2 c0 = d0, c1 = d1;
            const int64_t p0 = o0, p1 = o1;
            if (bk + 24 < b1) {
                d0 = *(const int2 *)(bcf + 2 * bk + 32);
                d1 = *(const int2 *)(bcf + 2 * bk + 48);
                o0 = boff[bk + 16];
                o1 = boff[bk + 24];
            }
            acc0 += bsymv_row(a + p0, c0.y, R, c0.x == fa, x + c0.x);
            acc1 += bsymv_row(a + p1, c1.y, R, c1.x == fa, x + c1.x);
        }
        if (bk < b1) {
            const int fb0 = bcf[2 * bk], nc0 = bcf[2 * bk + 1];
            acc0 += bsymv_row(a + boff[bk], nc0, R, fb0 == fa, x + fb0);
        }
        const double acc = sum_slots(acc0 + acc1);
        if (slot == 0 && r0 + rl < nr) y[fa + r0 + rl] = acc;
    }
}
// Blocks of groups of at most 8 parameters: lane t of a slot takes column t of the slot's block
// (clamped to the last column) and loads its 8 rows (clamped to the last row), so one load of a
// slot reads one block row, nc consecutive doubles.
__device__ __forceinline__ void bsymv_ld8(double (&bv)[8], const double *B, int nr, int nc, int t)
{
    const double *p = B + min(t, nc - 1);
#pragma unroll
    for (int r = 0; r < 8; r++) bv[r] = p[min(r, nr - 1) * nc];
}
// Rows pass for groups of at most 8 parameters.  ys[r] collects row r of the lane's columns over
// the off-diagonal blocks of the block row; the values of a slot's next block are loaded while the
// current one is multiplied, and its descriptors one step earlier still.  The diagonal block (the
// last block of the row, lower triangle stored) is done by the whole wave: lane (slot s, t) takes
// the symmetric element (t, s) times x[fa+s].  Rows are summed over the eight lanes of a slot by a
// halving butterfly (lane t ends with row t) and then over the slots.  With tb, the transposed
// product of every off-diagonal block (column t: sum_r B[r][t] x[fa+r]) goes to tb[8 * block + t]
// from the same loads (the matrix is then read once; k_bsymv_rows_st does this with coalesced
// loads and is the one used).
__global__ __launch_bounds__(256) void k_bsymv_rows8(int32_t ngroup, const int64_t *brow, const int32_t *bcf,
                                                     const int64_t *boff, const int32_t *gfirst, const double *a,
                                                     const double *x, double *y, double *tb)
{
    const int ga = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63, slot = lane >> 3, t = lane & 7;
    if (ga >= ngroup) return;
    const int fa = gfirst[ga] - 1, nr = gfirst[ga + 1] - 1 - fa;
    const int64_t b0 = brow[ga], b1 = brow[ga + 1];
    int64_t be = b1;
    double dv = 0.0, dx = 0.0;
    if (b1 > b0 && bcf[2 * (b1 - 1)] == fa) {
        be = b1 - 1;
        const int tr = min(t, nr - 1), sr = min(slot, nr - 1);
        dv = a[boff[b1 - 1] + (tr >= sr ? tr * nr + sr : sr * nr + tr)];
        dx = (t < nr && slot < nr) ? x[fa + sr] : 0.0;
    }
    double ys[8], bv[8], xa[8], xb = 0.0;
    int ncb = 0;
#pragma unroll
    for (int r = 0; r < 8; r++) {
        ys[r] = bv[r] = 0.0;
        const double v = x[fa + min(r, nr - 1)];
        xa[r] = r < nr ? v : 0.0;
    }
    // two blocks per slot in flight: bv (block bk) is used while bm (bk + 8) and bn (bk + 16)
    // are loaded; the descriptors run one block further ahead
    int64_t bk = b0 + slot;
    int2 dn = make_int2(0, 1);
    int64_t on = 0;
    double bm[8], x1 = 0.0;
    int nc1 = 0;
#pragma unroll
    for (int r = 0; r < 8; r++) bm[r] = 0.0;
    if (bk < be) {
        const int2 d = *(const int2 *)(bcf + 2 * bk);
        bsymv_ld8(bv, a + boff[bk], nr, d.y, t);
        const double xv = x[d.x + min(t, d.y - 1)];
        xb = t < d.y ? xv : 0.0;
        ncb = d.y;
    }
    if (bk + 8 < be) {
        const int2 d = *(const int2 *)(bcf + 2 * (bk + 8));
        bsymv_ld8(bm, a + boff[bk + 8], nr, d.y, t);
        const double xv = x[d.x + min(t, d.y - 1)];
        x1 = t < d.y ? xv : 0.0;
        nc1 = d.y;
    }
    if (bk + 16 < be) {
        dn = *(const int2 *)(bcf + 2 * (bk + 16));
        on = boff[bk + 16];
    }
    for (; bk < be; bk += 8) {
        double bn[8], xn = 0.0;
        int ncn = 0;
#pragma unroll
        for (int r = 0; r < 8; r++) bn[r] = 0.0;
        if (bk + 16 < be) {
            bsymv_ld8(bn, a + on, nr, dn.y, t);
            const double xv = x[dn.x + min(t, dn.y - 1)];
            xn = t < dn.y ? xv : 0.0;
            ncn = dn.y;
        }
        if (bk + 24 < be) {
            dn = *(const int2 *)(bcf + 2 * (bk + 24));
            on = boff[bk + 24];
        }
        double tc = 0.0;
#pragma unroll
        for (int r = 0; r < 8; r++) {
            ys[r] += bv[r] * xb;
            tc += bv[r] * xa[r];
            bv[r] = bm[r];
            bm[r] = bn[r];
        }
        if (tb && t < ncb) tb[8 * bk + t] = tc;
        xb = x1;
        ncb = nc1;
        x1 = xn;
        nc1 = ncn;
    }
    const bool h2 = t & 4, h1 = t & 2, h0 = t & 1;
    double z[4], w[2];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const double snd = h2 ? ys[k] : ys[k + 4], kp = h2 ? ys[k + 4] : ys[k];
        z[k] = kp + __shfl_xor(snd, 4, 64);
    }
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const double snd = h1 ? z[k] : z[k + 2], kp = h1 ? z[k + 2] : z[k];
        w[k] = kp + __shfl_xor(snd, 2, 64);
    }
    double v;
    {
        const double snd = h0 ? w[0] : w[1], kp = h0 ? w[1] : w[0];
        v = kp + __shfl_xor(snd, 1, 64);
    }
    v += dv * dx;
    v += __shfl_xor(v, 8, 64);
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    if (slot == 0 && t < nr) y[fa + t] = v;
}
// Rows pass for groups of at most 8 parameters with the matrix read once, fully coalesced.  A
// wavefront takes one block row.  Per step it stages the row's next eight off-diagonal blocks
// (one contiguous range of the matrix, consecutive lanes reading consecutive doubles) into LDS,
// with the x values of their column groups beside them.  Lane (k, j) = (lane / 8, lane % 8) then
// forms row j of block k times x_b (summed over the blocks of the row in the lane) and column j of
// block k times x_a, the block's transposed product, which goes to tb[tbw * block + j] for the
// columns pass.  Block descriptors come in windows of 64 blocks (one per lane), the next window
// loaded one window ahead, so that a step's matrix loads depend on registers only; the next
// step's loads are in flight while a step computes.  The diagonal block (the last of the row,
// lower triangle stored) is done by the whole wave as lane (k, j) = element (j, k); the row sums
// are then added over k.
// (W: wavefronts per workgroup, chosen per matrix by the autotuner of the product)
template <int W>
__global__ __launch_bounds__(64 * W) void k_bsymv_rows_st(int32_t ngroup, const int64_t *brow, const int32_t *bcf,
                                                         const int64_t *boff, const int32_t *gfirst, const double *a,
                                                         const double *x, double *y, double *tb, int tbw)
{
    __shared__ double st[W][512];
    __shared__ double sx[W][64];
    __shared__ int sd[W][16];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int ga = blockIdx.x * W + w;
    if (ga >= ngroup) return;                        // whole wavefronts; no workgroup barrier below
    const int k = lane >> 3, j = lane & 7;
    const int fa = gfirst[ga] - 1, nr = gfirst[ga + 1] - 1 - fa;
    const int64_t b0 = brow[ga], b1 = brow[ga + 1];
    int64_t be = b1;
    double accy = 0.0;
    if (b1 > b0 && bcf[2 * (b1 - 1)] == fa) {
        be = b1 - 1;
        const int tr = min(j, nr - 1), sr = min(k, nr - 1);
        const double dv = a[boff[b1 - 1] + (tr >= sr ? tr * nr + sr : sr * nr + tr)];
        const double dx = x[fa + sr];
        if (j < nr && k < nr) accy = dv * dx;
    }
    double xa[8];
#pragma unroll
    for (int r = 0; r < 8; r++) {
        const double v = x[fa + min(r, nr - 1)];
        xa[r] = r < nr ? v : 0.0;
    }
    const int64_t rend = boff[be];                   // end of the row's off-diagonal blocks
    const int64_t nsteps = (be - b0 + 7) / 8;
    // window descriptors: block wb + lane's first column, columns and offset (past the end of
    // the row: no columns, offset rend)
    // (every load below is issued unconditionally, at a clamped index, so that the compiler's
    // wait counts stay exact and the loads of the step in flight are not waited for)
    auto win = [&](int64_t wb, int &f, int &n, int64_t &o) {
        const int64_t q = wb + lane, qc = min(q, be - 1);
        const int2 d = *(const int2 *)(bcf + 2 * qc);
        const int64_t oq = boff[qc];
        f = q < be ? d.x : 0;
        n = q < be ? d.y : 0;
        o = q < be ? oq : rend;
    };
    int wf = 0, wn = 0, nf = 0, nn = 0;
    int64_t wo = rend, no = rend, cw = 0;
    if (nsteps > 0) win(b0, wf, wn, wo);
    if (nsteps > 8) win(b0 + 64, nf, nn, no);
    // two steps in flight: the registers of step t + 2 are loaded while step t computes
    struct StepRegs {
        double v[8], x;
        int o, n, len;
    };
    auto load = [&](int64_t t, StepRegs &R) {
        if (t / 8 != cw) {                           // next window; prefetch the one after it
            wf = nf;
            wn = nn;
            wo = no;
            cw = t / 8;
            if ((cw + 1) * 64 < be - b0) win(b0 + (cw + 1) * 64, nf, nn, no);
        }
        const int i0 = (int)(t % 8) * 8;
        const int kn = (int)min((int64_t)8, be - (b0 + 8 * t));
        const int64_t s0 = __shfl(wo, i0, 64);
        const int64_t s1 = __shfl(wo + (int64_t)nr * wn, i0 + 7, 64);
        R.len = (int)(s1 - s0);
#pragma unroll
        for (int i = 0; i < 8; i++) R.v[i] = a[s0 + min(lane + 64 * i, R.len - 1)];
        const int64_t ol = __shfl(wo, i0 + j, 64);
        const int nl = __shfl(wn, i0 + j, 64);
        const int fk = __shfl(wf, i0 + k, 64), nk = __shfl(wn, i0 + k, 64);
        R.o = (int)(ol - s0);
        R.n = nl;
        const bool xv = k < kn && j < nk;
        const double xl = x[xv ? fk + j : 0];
        R.x = xv ? xl : 0.0;
    };
    auto step = [&](int64_t t, StepRegs &R) {
        const int64_t g0 = b0 + 8 * t;
        const int kn = (int)min((int64_t)8, be - g0);
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const int idx = lane + 64 * i;
            if (64 * i < R.len && idx < R.len) st[w][idx] = R.v[i];
        }
        sx[w][lane] = R.x;
        if (lane < 8) {
            sd[w][lane] = R.o;
            sd[w][8 + lane] = R.n;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (t + 2 < nsteps) load(t + 2, R);
        if (k < kn) {
            const int off = sd[w][k], nc = sd[w][8 + k];
            const double *B = &st[w][off];
            if (j < nr) {
                double sv = 0.0;
#pragma unroll
                for (int c = 0; c < 8; c++)
                    if (c < nc) sv += B[j * nc + c] * sx[w][8 * k + c];
                accy += sv;
            }
            if (j < nc) {
                double sv = 0.0;
#pragma unroll
                for (int r = 0; r < 8; r++)
                    if (r < nr) sv += B[r * nc + j] * xa[r];
                tb[(g0 + k) * tbw + j] = sv;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };
    StepRegs RA, RB;
    if (nsteps > 0) load(0, RA);
    if (nsteps > 1) load(1, RB);
    for (int64_t t = 0; t < nsteps; t += 2) {
        step(t, RA);
        if (t + 1 < nsteps) step(t + 1, RB);
    }
    accy += __shfl_xor(accy, 8, 64);
    accy += __shfl_xor(accy, 16, 64);
    accy += __shfl_xor(accy, 32, 64);
    if (k == 0 && j < nr) y[fa + j] = accy;
}
// Rows pass when every parameter group has the same size G <= 8 (blocks G x G; group g's first
// index is g G).  As k_bsymv_rows_st, with the block layout known: the row's off-diagonal blocks
// are one contiguous range, so a step's matrix loads need no descriptor, and the x values of a
// block's column group need only its column group (bcol).  Per step the wave issues, in this
// order, the column groups of step t + 3, the x values of step t + 2 and the matrix values of
// step t + 2, all unconditionally (clamped indices; the transposed products of lanes without a
// block go to a spare slot past the end of tb), so that every wait is for the step it needs.
template <int G, int W, bool V2>
__global__ __launch_bounds__(64 * W) void k_bsymv_rows_u(int32_t ngroup, const int64_t *__restrict__ brow,
                                                               const int32_t *__restrict__ bcol,
                                                               const int64_t *__restrict__ boff,
                                                               const double *__restrict__ a, const double *__restrict__ x,
                                                               double *__restrict__ y, double *__restrict__ tb,
                                                               int64_t tb_spare)
{
    constexpr int S = G * G;                   // values per block
    __shared__ __align__(16) double st[W][8 * S];
    __shared__ double sx[W][64];
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int ga = blockIdx.x * W + w;
    if (ga >= ngroup) return;                  // whole wavefronts; no workgroup barrier below
    const int k = lane >> 3, j = lane & 7;
    const int fa = ga * G;
    const int64_t b0 = brow[ga], b1 = brow[ga + 1];
    const int64_t be = b1 - 1;                 // the last block of a row is its diagonal block
    double accy = 0.0;
    {
        const int tr = min(j, G - 1), sr = min(k, G - 1);
        const double dv = a[boff[be] + (tr >= sr ? tr * G + sr : sr * G + tr)];
        const double dx = x[fa + sr];
        if (j < G && k < G) accy = dv * dx;
    }
    double xa[G];
#pragma unroll
    for (int r = 0; r < G; r++) xa[r] = x[fa + r];
    const int64_t s_row = boff[b0];
    const int64_t nsteps = (be - b0 + 7) / 8;
    // matrix values per lane and step; with V2 (even block size, so every step starts on 16
    // bytes) they are loaded two at a time
    constexpr int NV2 = (4 * S + 63) / 64;
    constexpr int NV = V2 ? 2 * NV2 : (8 * S + 63) / 64;
    struct StepRegs {
        double v[NV], x;
        int col;                                // column group of the lane's block k (step + 1 ahead)
    };
    // column group of block k of step t (clamped to the row's last off-diagonal block)
    auto ld_col = [&](int64_t t) -> int { return bcol[min(b0 + 8 * t + k, be - 1)]; };
    auto ld_x = [&](int64_t t, int col) -> double { return x[col * G + min(j, G - 1)]; };
    auto ld_v = [&](int64_t t, StepRegs &R) {
        const int len = (int)(min((int64_t)8, be - (b0 + 8 * t)) * S);
        const int64_t s0 = s_row + 8 * t * S;
        if constexpr (V2) {
            const double2 *a2 = reinterpret_cast<const double2 *>(a + s0);
#pragma unroll
            for (int i = 0; i < NV2; i++) {
                const double2 u = a2[min(lane + 64 * i, len / 2 - 1)];
                R.v[2 * i] = u.x;
                R.v[2 * i + 1] = u.y;
            }
        } else {
#pragma unroll
            for (int i = 0; i < NV; i++) R.v[i] = a[s0 + min(lane + 64 * i, len - 1)];
        }
    };
    if (nsteps == 0) {                          // a row with its diagonal block only
        accy += __shfl_xor(accy, 8, 64);
        accy += __shfl_xor(accy, 16, 64);
        accy += __shfl_xor(accy, 32, 64);
        if (k == 0 && j < G) y[fa + j] = accy;
        return;
    }
    // steps past the last are loaded at the last step's indices and computed as empty steps
    StepRegs RA, RB;
    RA.col = ld_col(0);
    RA.x = ld_x(0, RA.col);
    ld_v(0, RA);
    RB.col = ld_col(min((int64_t)1, nsteps - 1));
    RB.x = ld_x(1, RB.col);
    ld_v(min((int64_t)1, nsteps - 1), RB);
    int col2 = ld_col(min((int64_t)2, nsteps - 1)), col_next = 0;   // column groups of step t + 2
    auto step = [&](int64_t t, StepRegs &R) {
        const int64_t g0 = b0 + 8 * t;
        const int kn = (int)min((int64_t)8, be - g0);
        const int len = kn * S;
        // (no branch in the loop body: lanes past the step's values write the value of its last
        // entry to that entry again, which the loads gave them)
        if constexpr (V2) {
#pragma unroll
            for (int i = 0; i < NV2; i++) {
                const int q = min(lane + 64 * i, max(len / 2, 1) - 1);
                *reinterpret_cast<double2 *>(&st[w][2 * q]) = make_double2(R.v[2 * i], R.v[2 * i + 1]);
            }
        } else {
#pragma unroll
            for (int i = 0; i < NV; i++) st[w][min(lane + 64 * i, max(len, 1) - 1)] = R.v[i];
        }
        sx[w][lane] = R.x;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        {
            const int64_t t2 = min(t + 2, nsteps - 1), t3 = min(t + 3, nsteps - 1);
            col_next = ld_col(t3);
            R.col = col2;
            R.x = ld_x(t2, col2);
            ld_v(t2, R);
            col2 = col_next;
        }
        const double *B = &st[w][max(min(k, kn - 1), 0) * S];
        double sv = 0.0;
#pragma unroll
        for (int c = 0; c < G; c++) sv += B[min(j, G - 1) * G + c] * sx[w][8 * k + c];
        accy += (k < kn && j < G) ? sv : 0.0;
        double tv = 0.0;
#pragma unroll
        for (int r = 0; r < G; r++) tv += B[r * G + min(j, G - 1)] * xa[r];
        const bool ok = k < kn && j < G;
        tb[ok ? (g0 + k) * G + j : tb_spare + lane] = tv;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };
    for (int64_t t = 0; t < nsteps; t += 2) {
        step(t, RA);
        step(t + 1, RB);
    }
    accy += __shfl_xor(accy, 8, 64);
    accy += __shfl_xor(accy, 16, 64);
    accy += __shfl_xor(accy, 32, 64);
    if (k == 0 && j < G) y[fa + j] = accy;
}
// Columns pass from the transposed block products of the rows pass: column group gb sums
// tb[tbw * block + t] over its off-diagonal blocks in column-list order (slot s takes entries
// s, s+8, ...; one read of the block's values), then over the slots.
__global__ __launch_bounds__(256) void k_bsymv_colst(int32_t ngroup, const int64_t *ccol_ptr, const int32_t *cc_blk,
                                                     const int32_t *gfirst, const double *tb, int tbw, double *y)
{
    const int gb = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63, slot = lane >> 3, t = lane & 7;
    if (gb >= ngroup) return;
    const int fb = gfirst[gb] - 1, nc = gfirst[gb + 1] - 1 - fb;
    const int64_t q0 = ccol_ptr[gb], q1 = ccol_ptr[gb + 1];
    if (q1 <= q0) return;
    double a0 = 0.0, a1 = 0.0;
    int64_t q = q0 + slot;
    const int tt = min(t, nc - 1);
    for (; q + 8 < q1; q += 16) {
        const int64_t k0 = cc_blk[q], k1 = cc_blk[q + 8];
        a0 += tb[tbw * k0 + tt];
        a1 += tb[tbw * k1 + tt];
    }
    if (q < q1) a0 += tb[tbw * (int64_t)cc_blk[q] + tt];
    const double acc = sum_slots(a0 + a1);
    if (slot == 0 && t < nc) y[fb + t] += acc;
}
// Columns pass for groups of at most 8 parameters: the off-diagonal blocks (row group a, column
// group gb) of a column group, lane t of a slot taking column t of the slot's block and loading its
// rows as in the rows pass; lane t also loads x[fa+t] and the slot's lanes pass these to each other.
// Values of the next block are loaded while the current one is multiplied.
__global__ __launch_bounds__(256) void k_bsymv_cols8(int32_t ngroup, const int64_t *ccol_ptr, const int64_t *cc_off,
                                                     const int32_t *cc_fn, const int32_t *gfirst, const double *a,
                                                     const double *x, double *y)
{
    const int gb = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63, slot = lane >> 3, t = lane & 7;
    if (gb >= ngroup) return;
    const int fb = gfirst[gb] - 1, nc = gfirst[gb + 1] - 1 - fb;
    const int64_t q0 = ccol_ptr[gb], q1 = ccol_ptr[gb + 1];
    if (q1 <= q0) return;
    double bv[8], xr = 0.0, acc = 0.0;
#pragma unroll
    for (int r = 0; r < 8; r++) bv[r] = 0.0;
    int64_t q = q0 + slot;
    int2 dn = make_int2(0, 1);
    int64_t on = 0;
    if (q < q1) {
        const int2 d = *(const int2 *)(cc_fn + 2 * q);
        bsymv_ld8(bv, a + cc_off[q], d.y, nc, t);
        const double xv = x[d.x + min(t, d.y - 1)];
        xr = t < d.y ? xv : 0.0;
    }
    if (q + 8 < q1) {
        dn = *(const int2 *)(cc_fn + 2 * (q + 8));
        on = cc_off[q + 8];
    }
    for (; q < q1; q += 8) {
        double bn[8], xn = 0.0;
#pragma unroll
        for (int r = 0; r < 8; r++) bn[r] = 0.0;
        if (q + 8 < q1) {
            bsymv_ld8(bn, a + on, dn.y, nc, t);
            const double xv = x[dn.x + min(t, dn.y - 1)];
            xn = t < dn.y ? xv : 0.0;
        }
        if (q + 16 < q1) {
            dn = *(const int2 *)(cc_fn + 2 * (q + 16));
            on = cc_off[q + 16];
        }
#pragma unroll
        for (int r = 0; r < 8; r++) {
            acc += bv[r] * __shfl(xr, r, 8);
            bv[r] = bn[r];
        }
        xr = xn;
    }
    acc = sum_slots(acc);
    if (slot == 0 && t < nc) y[fb + t] += acc;
}
// column Cc of block B (nr x nc, row-major) times x[fa .. fa+nr)
__device__ __forceinline__ double bsymv_col(const double *B, int nr, int nc, int Cc, const double *xa)
{
    double sv = 0.0;
    if (nr <= 8) {
        double v[8], xv[8];
#pragma unroll
        for (int r = 0; r < 8; r++) {
            const int rr = r < nr ? r : nr - 1;
            v[r] = B[(int64_t)rr * nc + Cc];
            xv[r] = xa[rr];
        }
#pragma unroll
        for (int r = 0; r < 8; r++)
            if (r < nr) sv += v[r] * xv[r];
        return sv;
    }
    for (int r = 0; r < nr; r++) sv += B[(int64_t)r * nc + Cc] * xa[r];
    return sv;
}
__global__ __launch_bounds__(256) void k_bsymv_cols(int32_t ngroup, const int64_t *ccol_ptr, const int64_t *cc_off,
                                                    const int32_t *cc_fn, const int32_t *gfirst, const double *a,
                                                    const double *x, double *y)
{
    const int gb = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63, slot = lane >> 3, cl = lane & 7;
    if (gb >= ngroup) return;
    const int fb = gfirst[gb] - 1, nc = gfirst[gb + 1] - 1 - fb;
    const int64_t q0 = ccol_ptr[gb], q1 = ccol_ptr[gb + 1];
    if (q1 <= q0) return;
    for (int c0 = 0; c0 < nc; c0 += 8) {
        const int Cc = min(c0 + cl, nc - 1);   // lanes past the last column repeat it (result unused)
        double acc0 = 0.0, acc1 = 0.0;
        int64_t q = q0 + slot;
        // the next step's entries are read one step ahead
        int2 d0 = make_int2(0, 1), d1 = make_int2(0, 1);
        int64_t o0 = 0, o1 = 0;
        if (q + 8 < q1) {
            d0 = *(const int2 *)(cc_fn + 2 * q);
            d1 = *(const int2 *)(cc_fn + 2 * q + 16);
            o0 = cc_off[q];
            o1 = cc_off[q + 8];
        }
        for (; q + 8 < q1; q += 16) {
            const int2 c0 = d0, c1 = d1;
            const int64_t p0 = o0, p1 = o1;
            if (q + 24 < q1) {
                d0 = *(const int2 *)(cc_fn + 2 * q + 32);
                d1 = *(const int2 *)(cc_fn + 2 * q + 48);
                o0 = cc_off[q + 16];
                o1 = cc_off[q + 24];
            }
            acc0 += bsymv_col(a + p0, c0.y, nc, Cc, x + c0.x);
            acc1 += bsymv_col(a + p1, c1.y, nc, Cc, x + c1.x);
        }
        if (q < q1) acc0 += bsymv_col(a + cc_off[q], cc_fn[2 * q + 1], nc, Cc, x + cc_fn[2 * q]);
        const double acc = sum_slots(acc0 + acc1);
        if (slot == 0 && c0 + cl < nc) y[fb + c0 + cl] += acc;
    }
}

// constraint rows (Lagrange multipliers) in CSR over variable indices:
// y[nvar+j] += sum c_ji x_i ; y_i += c_ji x[nvar+j]
__global__ void k_cons_symv(int32_t ncon, int32_t nvar, const int64_t *cptr, const int32_t *cidx,
                            const double *cval, const double *x, double *y)
{
    int j = blockIdx.x;
    if (j >= ncon) return;
    __shared__ double s[TPB];
    double acc = 0.0;
    const double lam = x[nvar + j];
    for (int64_t p = cptr[j] + threadIdx.x; p < cptr[j + 1]; p += TPB) {
        int i = cidx[p] - 1;
        double c = cval[p];
        acc += c * x[i];
        atomicAdd(&y[i], c * lam);
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int o = TPB / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(&y[nvar + j], s[0]);
}

// ---------------------------------------------------------------------------------------------
// MINRES vector algebra (src/minresModule.f90:439-520)
// ---------------------------------------------------------------------------------------------
__global__ void k_scale_copy(int64_t n, double s, const double *a, double *out)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) out[i] = s * a[i];
}
// ---- MINRES with the scalar recurrence on the device (see MrState) ----
__global__ void k_mr_scale(int64_t n, const MrState *st, const double *y, double *v)
{
    if (st->istop != 0) return;
    const double sc = 1.0 / st->beta;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) v[i] = sc * y[i];
}
// y -= (beta/oldb) r (which = 0) or y -= (alfa/beta) r (which = 1)
__global__ void k_mr_axpy(int64_t n, const MrState *st, int which, const double *r, double *y)
{
    if (st->istop != 0) return;
    const double c = which == 0 ? st->beta / st->oldb : st->alfa / st->beta;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) y[i] = y[i] - c * r[i];
}
// one step of the scalar recurrence (src/minresModule.f90:468-600), the host restatement's
// operations in the same order, without contraction into fused multiply-adds
__global__ void k_mr_step(MrState *st)
{
#pragma clang fp contract(off)
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (st->istop != 0) { st->upd = 0; return; }
    const double eps = 2.220446049250313e-16;
    const int itn = ++st->itn;
    const double alfa = st->alfa;
    const double oldb = st->beta;
    double beta = st->bnew;
    st->oldb = oldb;
    if (beta < 0.0) { st->istop = 6; st->upd = 0; return; }
    beta = sqrt(beta);
    st->beta = beta;
    double tnorm2 = st->tnorm2 + alfa * alfa + oldb * oldb + beta * beta;
    int istop = st->istop;
    double gmax = st->gmax, gmin = st->gmin;
    if (itn == 1) {
        if (beta / st->beta1 <= 10.0 * eps) istop = -1;
        gmax = fabs(alfa);
        gmin = gmax;
    }
    const double cs0 = st->cs, sn0 = st->sn, dbar0 = st->dbar;
    const double oldeps = st->epsln;
    const double delta = cs0 * dbar0 + sn0 * alfa;
    const double gbar = sn0 * dbar0 - cs0 * alfa;
    const double epsln = sn0 * beta;
    const double dbar = -cs0 * beta;
    const double gamma = sqrt(gbar * gbar + beta * beta);
    const double cs = gbar / gamma;
    const double sn = beta / gamma;
    const double phi = cs * st->phibar;
    const double phibar = sn * st->phibar;
    const double denom = 1.0 / gamma;
    st->oldeps = oldeps;
    st->delta = delta;
    st->denom = denom;
    st->phi = phi;
    st->upd = 1;
    gmax = fmax(gmax, gamma);
    gmin = fmin(gmin, gamma);
    const double z = st->rhs1 / gamma;
    const double ynorm2 = z * z + st->ynorm2;
    st->rhs1 = st->rhs2 - delta * z;
    st->rhs2 = -epsln * z;
    const double Anorm = sqrt(tnorm2);
    const double ynorm = sqrt(ynorm2);
    const double epsx = Anorm * ynorm * eps;
    const double epsr = Anorm * ynorm * st->rtol;
    const double qrnorm = phibar;
    const double rnorml = st->rnorm;
    const double rnorm = qrnorm;
    const double rootl = sqrt(gbar * gbar + dbar * dbar);
    const double Arnorml = rnorml * rootl;
    const double relArnorml = rootl / Anorm;
    const double Acond = gmax / gmin;
    if (istop == 0) {
        if (itn >= st->itnlim) istop = 5;
        if (Acond >= 0.1 / eps) istop = 4;
        if (epsx >= st->beta1) istop = 3;
        if (qrnorm <= epsx || relArnorml <= eps) istop = 2;
        if (qrnorm <= epsr || relArnorml <= st->rtol) istop = 1;
    }
    st->tnorm2 = tnorm2;
    st->gmax = gmax;
    st->gmin = gmin;
    st->epsln = epsln;
    st->dbar = dbar;
    st->cs = cs;
    st->sn = sn;
    st->phibar = phibar;
    st->ynorm2 = ynorm2;
    st->Anorm = Anorm;
    st->ynorm = ynorm;
    st->qrnorm = qrnorm;
    st->rnorm = rnorm;
    st->Arnorml = Arnorml;
    st->Acond = Acond;
    st->istop = istop;
}
// w1 = w2; w2 = w; w = (v - oldeps*w1 - delta*w2)*denom; x = x + phi*w  (iterations that ran)
__global__ void k_mr_update(int64_t n, const MrState *st, const double *v, double *w, double *w1, double *w2, double *x)
{
    if (st->upd == 0) return;
    const double oldeps = st->oldeps, delta = st->delta, denom = st->denom, phi = st->phi;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double a1 = w2[i];
        double a2 = w[i];
        w1[i] = a1;
        w2[i] = a2;
        double wn = (v[i] - oldeps * a1 - delta * a2) * denom;
        w[i] = wn;
        x[i] = x[i] + phi * wn;
    }
}

// y = y - c*r
__global__ void k_axpy_neg(int64_t n, double c, const double *r, double *y)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) y[i] = y[i] - c * r[i];
}
// w1 = w2; w2 = w; w = (v - oldeps*w1 - delta*w2)*denom; x = x + phi*w
__global__ void k_minres_update(int64_t n, const double *v, double *w, double *w1, double *w2, double *x, double oldeps,
                                double delta, double denom, double phi)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double a1 = w2[i];
        double a2 = w[i];
        w1[i] = a1;
        w2[i] = a2;
        double wn = (v[i] - oldeps * a1 - delta * a2) * denom;
        w[i] = wn;
        x[i] = x[i] + phi * wn;
    }
}
// MINRES-QLP vector updates (src/minresqlpModule.f90:970-1034).
// MINRES step: wl2 = wl; wl = w; w = (v - epln*wl2 - dlta*wl)*s (if upd_w); x += tau*w (if upd_x)
__global__ void k_qlp_minres(int64_t n, const double *v, double *w, double *wl, double *wl2, double *x, double epln,
                             double dlta, double s, double tau, int upd_w, int upd_x)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double a2 = wl[i], a1 = w[i];
        wl2[i] = a2;
        wl[i] = a1;
        double wn = upd_w ? (v[i] - epln * a2 - dlta * a1) * s : a1;
        w[i] = wn;
        if (upd_x) x[i] = x[i] + tau * wn;
    }
}
// QLP step: on the first QLP iteration the MINRES directions are converted (w_{k-3..k-1})
// and xl2 rebuilt from x; then the right reflections rotate (wl2, wl, w, v) and
// xl2 += ul2*wl2, x = xl2 + ul*wl + u*w
__global__ void k_qlp_step(int64_t n, double *v, double *w, double *wl, double *wl2, double *x, double *xl2, QlpStep c)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double V = v[i], W = w[i], WL = wl[i], WL2 = wl2[i], X = x[i], XL2 = xl2[i];
        if (c.first) {
            XL2 = 0.0;
            if (c.itn > 1) {
                if (c.itn > 3) WL2 = c.gamal3 * WL2 + c.veplnl2 * WL + c.etal * W;
                if (c.itn > 2) WL = c.gamal_qlp * WL + c.vepln_qlp * W;
                W = c.gama_qlp * W;
                XL2 = X - c.ul_qlp * WL - c.u_qlp * W;
            }
        }
        if (c.itn == 1) {
            WL2 = WL;
            WL = c.sr1 * V;
            W = -c.cr1 * V;
        } else if (c.itn == 2) {
            WL2 = WL;
            double nwl = c.cr1 * W + c.sr1 * V;
            W = c.sr1 * W - c.cr1 * V;
            WL = nwl;
        } else {
            double a2 = WL, a1 = W;
            double nw = c.sr2 * a2 - c.cr2 * V;
            WL2 = c.cr2 * a2 + c.sr2 * V;
            V = c.cr1 * a1 + c.sr1 * nw;
            W = c.sr1 * a1 - c.cr1 * nw;
            WL = V;
        }
        XL2 = XL2 + c.ul2 * WL2;
        X = XL2 + c.ul * WL + c.u * W;
        v[i] = V; w[i] = W; wl[i] = WL; wl2[i] = WL2; x[i] = X; xl2[i] = XL2;
    }
}
// p = z + beta p
__global__ void k_xpby(int64_t n, const double *z, double beta, double *p)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = z[i] + beta * p[i];
}
__global__ void k_copy(int64_t n, const double *a, double *b)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) b[i] = a[i];
}
// r = b - A x  (A x given in ax)
__global__ void k_sub(int64_t n, const double *b, const double *ax, double *r)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) r[i] = b[i] - ax[i];
}

// ---------------------------------------------------------------------------------------------
// Diagonal preconditioner with constraint blocks: precons/presols (src/mpnum.f90:2881-3069)
// The constraint part (p constraints, CSR rows over variable indices, p small) is done by one
// workgroup; the Schur block S = K K^T (K = C D^-1/2) is factorised once (k_precons_s).
// ---------------------------------------------------------------------------------------------
__global__ void k_precons_cu(int32_t n, const double *diag, double *cu, int *nrkd)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double d = diag[i];
    if (d > 0.0) cu[i] = 1.0 / sqrt(d);
    else { cu[i] = 0.0; atomicAdd(nrkd, 1); }
}

// S(jk) = sum_i (c_ji cu_i)(c_ki cu_i), packed p(p+1)/2, then H D H^T (precons lines 2960-2975)
__global__ void k_precons_s(int32_t p, const int64_t *cptr, const int32_t *cidx, const double *cval, const double *cu,
                            int32_t nvar, double *s, double *scratch_row /* nvar */)
{
    // single workgroup: build S column by column using a dense scratch of constraint j
    for (int j = 0; j < p; j++) {
        for (int i = threadIdx.x; i < nvar; i += blockDim.x) scratch_row[i] = 0.0;
        __syncthreads();
        for (int64_t q = cptr[j] + threadIdx.x; q < cptr[j + 1]; q += blockDim.x) {
            int i = cidx[q] - 1;
            scratch_row[i] = cval[q] * cu[i];
        }
        __syncthreads();
        for (int k = 0; k <= j; k++) {
            __shared__ double red[TPB];
            double acc = 0.0;
            for (int64_t q = cptr[k] + threadIdx.x; q < cptr[k + 1]; q += blockDim.x) {
                int i = cidx[q] - 1;
                acc += scratch_row[i] * (cval[q] * cu[i]);
            }
            red[threadIdx.x] = acc;
            __syncthreads();
            for (int o = TPB / 2; o > 0; o >>= 1) {
                if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
                __syncthreads();
            }
            if (threadIdx.x == 0) s[(j * j + j) / 2 + k] = red[0];
            __syncthreads();
        }
    }
    if (threadIdx.x == 0) {
        double *S = s - 1;
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
}

// apply: x = M^-1 y (presols). n variables, p constraints.
__global__ void k_presols_a(int64_t n, const double *cu, const double *y, double *x)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        x[i] = y[i] * cu[i];
}
__global__ void k_presols_c(int32_t p, int32_t n, const int64_t *cptr, const int32_t *cidx, const double *cval,
                            const double *cu, const double *s, const double *y, double *x)
{
    // single workgroup; x[0..n) holds u = C^-1/2 y on entry
    __shared__ double red[TPB];
    __shared__ double lam[1024];
    for (int j = 0; j < p; j++) {
        double acc = 0.0;
        for (int64_t q = cptr[j] + threadIdx.x; q < cptr[j + 1]; q += blockDim.x) {
            int i = cidx[q] - 1;
            acc += (cval[q] * cu[i]) * x[i];
        }
        red[threadIdx.x] = acc;
        __syncthreads();
        for (int o = TPB / 2; o > 0; o >>= 1) {
            if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
            __syncthreads();
        }
        if (threadIdx.x == 0) lam[j] = y[n + j] - red[0];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double *S = s - 1;
        double *X = lam - 1;
        int jj = 0;
        for (int j = 1; j <= p; j++) {
            double dsum = X[j];
            for (int k = 1; k <= j - 1; k++) dsum = dsum - S[k + jj] * X[k];
            X[j] = dsum;
            jj = jj + j;
        }
        for (int j = p; j >= 1; j--) {
            double dsum = X[j] * S[jj];
            int kk = jj;
            for (int k = j + 1; k <= p; k++) { dsum = dsum + S[kk + j] * X[k]; kk = kk + k; }
            X[j] = -dsum;
            jj = jj - j;
        }
    }
    __syncthreads();
    for (int j = 0; j < p; j++) {
        for (int64_t q = cptr[j] + threadIdx.x; q < cptr[j + 1]; q += blockDim.x) {
            int i = cidx[q] - 1;
            x[i] = x[i] - (cval[q] * cu[i]) * lam[j];
        }
        __syncthreads();
    }
    for (int j = threadIdx.x; j < p; j += blockDim.x) x[n + j] = lam[j];
}
__global__ void k_presols_b(int64_t n, const double *cu, double *x)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        x[i] = x[i] * cu[i];
}

// ---------------------------------------------------------------------------------------------
// Variable-band Cholesky preconditioner (lltdec/lltfwd/lltbwd, src/mpnum.f90:2155-2336).
// The recurrences are sequential in the row index; they run on one lane.
// ---------------------------------------------------------------------------------------------
// lltdec (src/mpnum.f90:2155-2222): variable-band Cholesky C = L L^T with the
// inverse square roots of the pivots on the diagonal, restated operation by operation.  The
// recurrence is sequential in the row index; the band values are staged through LDS in tiles
// (coalesced loads and stores by the whole workgroup) and one thread runs the recurrence on
// LDS, with the last rows of the previous tile kept in a small ring.
#define LLT_TILE 6144     // band values per tile
#define LLT_MW 64         // widest row supported by the tiled path
__global__ __launch_bounds__(256) void k_lltdec(int n, double *c, const int *india, int *nrkd, int iopt, int mbw)
{
    __shared__ double buf[LLT_TILE];
    __shared__ int ind[LLT_TILE + 1];
    __shared__ double ring[LLT_MW * LLT_MW];   // row j < k0 in slot j % MW: element (j, i) at [slot*MW + MW-1 - j + i]
    const int tid = threadIdx.x;
    const double eps = 16.0 * DBL_EPSILON;
    const int rows = max(1, LLT_TILE / max(1, mbw));  // rows per tile (each row holds <= mbw values)
    int nr = 0;
    double diag = 0.0;
    for (int k0 = 1; k0 <= n; k0 += rows) {
        const int k1 = min(n, k0 + rows - 1);
        const int base = k0 > 1 ? india[k0 - 2] : 0;   // IND(k0-1)
        const int nel = india[k1 - 1] - base;
        for (int t = tid; t < nel; t += 256) buf[t] = c[base + t];
        for (int t = tid; t <= k1 - k0; t += 256) ind[t + 1] = india[k0 - 1 + t];
        if (tid == 0) ind[0] = base;
        __syncthreads();
        if (tid == 0) {
            // row pointers: element (j, i) = P[i]
            auto rowp = [&](int j) -> double * {
                if (j >= k0) return &buf[ind[j - k0 + 1] - j - base - 1];
                return &ring[(j % LLT_MW) * LLT_MW + LLT_MW - 1 - j];
            };
            auto indv = [&](int j) -> int {          // IND(j) for j >= 0
                if (j >= k0 - 1) return ind[j - k0 + 1];
                return j >= 1 ? india[j - 1] : 0;
            };
            for (int k = k0; k <= k1; k++) {
                double *Pk = rowp(k);
                if (k == 1) {
                    if (Pk[1] > 0.0) Pk[1] = 1.0 / sqrt(Pk[1]);
                    else { Pk[1] = 0.0; nr = -1; }
                    continue;
                }
                const int mk = k - indv(k) + indv(k - 1) + 1;
                int indj1 = indv(mk - 1);
                for (int j = mk; j <= k; j++) {
                    const int indj = indv(j);
                    const int mj = j > 1 ? j - indj + indj1 + 1 : 1;
                    indj1 = indj;
                    const double *Pj = rowp(j);
                    diag = Pj[j];
                    // accumulate in a register; the loads of the loop are independent
                    double ckj = Pk[j];
#pragma unroll 4
                    for (int i = max(mj, mk); i <= j - 1; i++) ckj = ckj - Pk[i] * Pj[i];
                    if (j != k) ckj = ckj * diag;
                    Pk[j] = ckj;
                }
                if (Pk[k] > eps * diag) Pk[k] = 1.0 / sqrt(Pk[k]);
                else {
                    for (int j = mk; j <= k; j++) Pk[j] = 0.0;
                    if (iopt > 0 && diag > 0.0) Pk[k] = 1.0 / sqrt(diag);
                    else {
                        if (nr == 0) nr = -k;
                        else { if (nr < 0) nr = 1; nr++; }
                    }
                }
            }
            // keep the last rows of the tile for the next one
            for (int k = max(k0, k1 - LLT_MW + 1); k <= k1; k++) {
                const int mk = k - indv(k) + indv(k - 1) + 1;
                const double *Pk = rowp(k);
                double *Rk = &ring[(k % LLT_MW) * LLT_MW + LLT_MW - 1 - k];
                for (int j = mk; j <= k; j++) Rk[j] = Pk[j];
            }
        }
        __syncthreads();
        for (int t = tid; t < nel; t += 256) c[base + t] = buf[t];
        __syncthreads();
    }
    if (tid == 0) *nrkd = nr;
}

// lltdec for rows of at most MW band values: the same recurrence and operation order, with the
// current row in registers (statically indexed, right-aligned: r[o] = element (k, k-MW+1+o)),
// the previous MW-1 finished rows in a right-aligned LDS ring, and every load of a row issued
// before its arithmetic.
template <int MW>
__global__ __launch_bounds__(256) void k_lltdec_w(int n, double *c, const int *india, int *nrkd, int iopt)
{
    __shared__ double buf[LLT_TILE];
    __shared__ int ind[LLT_TILE + 1];
    __shared__ double ring[MW * MW];     // row j at slot j % MW, element (j, j-MW+1+o) at [slot*MW + o]
    const int tid = threadIdx.x;
    const double eps = 16.0 * DBL_EPSILON;
    const int rows = LLT_TILE / MW;      // rows per tile (each row holds <= MW values)
    int nr = 0;
    for (int k0 = 1; k0 <= n; k0 += rows) {
        const int k1 = min(n, k0 + rows - 1);
        const int base = k0 > 1 ? india[k0 - 2] : 0;   // IND(k0-1)
        const int nel = india[k1 - 1] - base;
        for (int t = tid; t < nel; t += 256) buf[t] = c[base + t];
        for (int t = tid; t <= k1 - k0; t += 256) ind[t + 1] = india[k0 - 1 + t];
        if (tid == 0) ind[0] = base;
        __syncthreads();
        if (tid == 0) {
            for (int k = k0; k <= k1; k++) {
                const int indk = ind[k - k0 + 1], indk1 = ind[k - k0];
                const int mk = k - indk + indk1 + 1;          // first column of row k
                double *Pk = &buf[indk - k - base - 1];       // element (k, j) at Pk[j]
                double r[MW], pr[MW][MW], dj[MW];
                // every load is unconditional (in-bounds addresses) and selected afterwards, so that
                // no branch separates the loads of a row
#pragma unroll
                for (int o = 0; o < MW; o++) {
                    const int col = k - MW + 1 + o;
                    const double v = Pk[max(col, mk)];
                    r[o] = col >= mk ? v : 0.0;
                }
#pragma unroll
                for (int jo = 0; jo < MW - 1; jo++) {
                    const int j = k - MW + 1 + jo;
                    const int slot = (j + MW) % MW;
                    const bool valid = j >= mk;
                    const double d = ring[slot * MW + MW - 1];
                    dj[jo] = valid ? d : 0.0;
#pragma unroll
                    for (int io = 0; io < jo; io++) {
                        const double v = ring[slot * MW + MW - 1 - (jo - io)];
                        pr[jo][io] = valid ? v : 0.0;
                    }
                }
                if (k == 1) {
                    double v = r[MW - 1];
                    if (v > 0.0) v = 1.0 / sqrt(v);
                    else { v = 0.0; nr = -1; }
                    r[MW - 1] = v;
                } else {
                    // terms outside a row's band are exact zeros (r is zero below column mk, the
                    // ring rows below their own first column, rows before mk are read as zero),
                    // so the products run over the full window without selects in the chain
#pragma unroll
                    for (int jo = 0; jo < MW - 1; jo++) {
                        double ckj = r[jo];
#pragma unroll
                        for (int io = 0; io < jo; io++) ckj = ckj - r[io] * pr[jo][io];
                        r[jo] = ckj * dj[jo];
                    }
                    const double diag = r[MW - 1];
                    double ckk = r[MW - 1];
#pragma unroll
                    for (int io = 0; io < MW - 1; io++) ckk = ckk - r[io] * r[io];
                    if (ckk > eps * diag) r[MW - 1] = 1.0 / sqrt(ckk);
                    else {
#pragma unroll
                        for (int o = 0; o < MW; o++) r[o] = 0.0;
                        if (iopt > 0 && diag > 0.0) r[MW - 1] = 1.0 / sqrt(diag);
                        else {
                            if (nr == 0) nr = -k;
                            else { if (nr < 0) nr = 1; nr++; }
                        }
                    }
                }
                const int slot = k % MW;
#pragma unroll
                for (int o = 0; o < MW; o++) {
                    const int col = k - MW + 1 + o;
                    if (col >= mk) Pk[col] = r[o];   // the row's own values only
                    ring[slot * MW + o] = r[o];
                }
            }
        }
        __syncthreads();
        for (int t = tid; t < nel; t += 256) c[base + t] = buf[t];
        __syncthreads();
    }
    if (tid == 0) *nrkd = nr;
}


// One row of the band factorization (lltdec, src/mpnum.f90:2155-2230): row k (slot s = (k-1) % MW,
// first column mk) from its band values r (element (k, k-MW+1+o) at r[o], zero before mk) and the
// previous MW-1 factor rows in ring (row j in slot (j-1) % MW).  Returns 1 when the pivot is not
// above eps times the diagonal and the row is dropped (a rank defect), else 0.  The sequential and
// the chunk-parallel kernels both use it, so their results agree bit for bit.
template <int MW>
__device__ __forceinline__ int llt_row(int k, int mk, int s, double (&r)[MW], const double (&ring)[MW][MW], int iopt)
{
    const double eps = 16.0 * DBL_EPSILON;
    if (k == 1) {
        double v = r[MW - 1];
        int bad = 0;
        if (v > 0.0) v = 1.0 / sqrt(v);
        else { v = 0.0; bad = 1; }
        r[MW - 1] = v;
        return bad;
    }
#pragma unroll
    for (int jo = 0; jo < MW - 1; jo++) {
        const int sj = (s + 1 + jo) % MW;
        const bool valid = k - MW + 1 + jo >= mk;
        double ckj = r[jo];
#pragma unroll
        for (int io = 0; io < jo; io++) ckj = ckj - r[io] * (valid ? ring[sj][MW - 1 - (jo - io)] : 0.0);
        r[jo] = ckj * (valid ? ring[sj][MW - 1] : 0.0);
    }
    const double diag = r[MW - 1];
    double ckk = r[MW - 1];
#pragma unroll
    for (int io = 0; io < MW - 1; io++) ckk = ckk - r[io] * r[io];
    if (ckk > eps * diag) {
        r[MW - 1] = 1.0 / sqrt(ckk);
        return 0;
    }
#pragma unroll
    for (int o = 0; o < MW; o++) r[o] = 0.0;
    if (iopt > 0 && diag > 0.0) {
        r[MW - 1] = 1.0 / sqrt(diag);
        return 0;
    }
    return 1;
}
// rank-defect code of lltdec from the dropped rows: 0 none, -k for one (row k), else their number
__device__ __forceinline__ int llt_nr(int nr, int k)
{
    if (nr == 0) return -k;
    if (nr < 0) nr = 1;
    return nr + 1;
}

// The same recurrence with the tile's rows staged densely: the lanes copy row k's band (MW values,
// element (k, k-MW+1+o) at d[(k-k0)*MW + o], zero before the row's first column) and its first
// column into LDS, lane 0 reads and writes rows at fixed offsets, and the lanes scatter the
// in-band values back.  Terms outside a row's band are exact zeros, so the operations and their
// order are those of k_lltdec_w.
template <int MW>
__global__ __launch_bounds__(64) void k_lltdec_d(int n, double *c, const int *india, int *nrkd, int iopt)
{
    constexpr int ROWS = (LLT_TILE / MW / MW) * MW;   // rows per tile, a multiple of MW
    __shared__ double d[ROWS * MW];
    __shared__ int mkr[ROWS];
    const int tid = threadIdx.x;
    double ring[MW][MW];                  // slot s: row j with (j-1) % MW == s, element (j, j-MW+1+o) at [s][o]
#pragma unroll
    for (int q = 0; q < MW; q++)
#pragma unroll
        for (int o = 0; o < MW; o++) ring[q][o] = 0.0;
    int nr = 0;
    for (int k0 = 1; k0 <= n; k0 += ROWS) {
        const int k1 = min(n, k0 + ROWS - 1);
        const int ne = (k1 - k0 + 1) * MW;
        for (int t = tid; t < ne; t += 64) {
            const int r = t / MW, o = t - r * MW, k = k0 + r;
            const int indk = india[k - 1], indk1 = k > 1 ? india[k - 2] : 0;
            const int mk = k - indk + indk1 + 1, col = k - MW + 1 + o;
            d[t] = col >= mk ? c[indk - k + col - 1] : 0.0;
            if (o == 0) mkr[r] = mk;
        }
        __syncthreads();
        if (tid == 0) {
            for (int kg = k0; kg <= k1; kg += MW) {
#pragma unroll
                for (int s = 0; s < MW; s++) {
                    const int k = kg + s;                   // (k-1) % MW == s
                    if (k > k1) break;
                    const int mk = mkr[k - k0];
                    double *Pk = &d[(k - k0) * MW];
                    double r[MW];
#pragma unroll
                    for (int o = 0; o < MW; o++) r[o] = Pk[o];
                    if (llt_row<MW>(k, mk, s, r, ring, iopt)) nr = llt_nr(nr, k);
#pragma unroll
                    for (int o = 0; o < MW; o++) {
                        Pk[o] = r[o];
                        ring[s][o] = r[o];
                    }
                }
            }
        }
        __syncthreads();
        for (int t = tid; t < ne; t += 64) {
            const int r = t / MW, o = t - r * MW, k = k0 + r;
            const int col = k - MW + 1 + o;
            if (col >= mkr[r]) c[india[k - 1] - k + col - 1] = d[t];
        }
        __syncthreads();
    }
    if (tid == 0) *nrkd = nr;
}

__device__ void lltfwd_dev(int n, const double *c, const int *india, double *x)
{
    const double *C = c - 1;
    const int *IND = india - 1;
    double *X = x - 1;
    X[1] = X[1] * C[IND[1]];
    for (int k = 2; k <= n; k++) {
        double xk = X[k];
        for (int j = k - IND[k] + IND[k - 1] + 1; j <= k - 1; j++) xk = xk - C[IND[k] - k + j] * X[j];
        X[k] = xk * C[IND[k]];
    }
}
__device__ void lltbwd_dev(int n, const double *c, const int *india, double *x)
{
    const double *C = c - 1;
    const int *IND = india - 1;
    double *X = x - 1;
    for (int k = n; k >= 2; k--) {
        double xk = X[k] * C[IND[k]];
        X[k] = xk;
        for (int j = k - IND[k] + IND[k - 1] + 1; j <= k - 1; j++) X[j] = X[j] - C[IND[k] - k + j] * xk;
    }
    X[1] = X[1] * C[IND[1]];
}
__global__ void k_lltfwd(int n, const double *c, const int *india, double *x)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) lltfwd_dev(n, c, india, x);
}
__global__ void k_lltbwd(int n, const double *c, const int *india, double *x)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) lltbwd_dev(n, c, india, x);
}

// K rows: forward solve of each constraint column restricted to its block range (lltfwds,
// src/mpnum.f90:2275-2306); one thread per constraint.  kmat: per constraint `in` values.
__global__ void k_band_k(int32_t ncon, int n, const double *c, const int *india, const int *cb /* per constraint: i0, in */,
                         const int64_t *koff, double *kmat)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ncon) return;
    int i0 = cb[2 * j], ns = cb[2 * j + 1];
    double *X = kmat + koff[j] - 1;
    const double *C = c - 1;
    const int *IND = india - 1;
    int kmax = min(ns, n - i0);
    for (int k = 1; k <= kmax; k++) {
        int l = k + i0;
        double xk = X[k];
        if (l > 1) {
            int j0 = max(1, k - IND[l] + IND[l - 1] + 1);
            for (int jj = j0; jj <= k - 1; jj++) xk = xk - C[IND[l] - k + jj] * X[jj];
        }
        X[k] = xk * C[IND[l]];
    }
}

// The same forward substitutions with the last MW-1 solution values of a column kept in registers
// (the same operations in the same order): a row then waits for its own arithmetic only, not for
// the store and reload of the value before it.
template <int MW>
__global__ void k_band_k_r(int32_t ncon, int n, const double *c, const int *india, const int *cb, const int64_t *koff,
                           double *kmat)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ncon) return;
    int i0 = cb[2 * j], ns = cb[2 * j + 1];
    double *X = kmat + koff[j] - 1;
    const double *C = c - 1;
    const int *IND = india - 1;
    int kmax = min(ns, n - i0);
    double xr[MW];                        // X[jj] at slot (jj - 1) % MW
#pragma unroll
    for (int q = 0; q < MW; q++) xr[q] = 0.0;
    // row k's operands (its band row of the factor, from column k - MW + 1 on, and X[k]) are
    // loaded one row ahead, so that the loads do not wait for the arithmetic
    auto load_row = [&](int k, double (&cr)[MW], double &xk0, int &j0) {
        const int l = k + i0;
        const int il = IND[l], il1 = l > 1 ? IND[l - 1] : 0;
        j0 = l > 1 ? max(1, k - il + il1 + 1) : k;
#pragma unroll
        for (int d = 0; d < MW; d++) {
            const int jj = k - (MW - 1) + d;                 // d = MW - 1: the diagonal
            cr[d] = (jj >= j0 || d == MW - 1) ? C[il - k + jj] : 0.0;
        }
        xk0 = X[k];
    };
    double cn[MW], xn = 0.0;
    int jn = 1;
    if (kmax >= 1) load_row(1, cn, xn, jn);
    for (int k = 1; k <= kmax; k++) {
        double cr[MW];
#pragma unroll
        for (int d = 0; d < MW; d++) cr[d] = cn[d];
        double xk = xn;
        const int j0 = jn;
        if (k < kmax) load_row(k + 1, cn, xn, jn);
#pragma unroll
        for (int d = MW - 1; d >= 1; d--) {                  // jj = k - d, ascending
            const int jj = k - d;
            if (jj >= j0) xk = xk - cr[MW - 1 - d] * xr[(jj - 1) % MW];
        }
        xk = xk * cr[MW - 1];
        X[k] = xk;
        xr[(k - 1) % MW] = xk;
    }
}

// H = K K^T restricted to overlapping ranges (equdecs products), packed m(m+1)/2: one thread per
// pair (j >= k), its sum in index order
__global__ void k_band_hprod(int32_t m, const int *cb, const int64_t *koff, const double *kmat, double *h)
{
    const int np = (m * m + m) / 2;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < np; t += gridDim.x * blockDim.x) {
        int j = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
        while ((j * j + j) / 2 > t) j--;
        while (((j + 1) * (j + 1) + (j + 1)) / 2 <= t) j++;
        int k = t - (j * j + j) / 2;
        int ij0 = cb[2 * j], ijn = cb[2 * j + 1];
        int ik0 = cb[2 * k], ikn = cb[2 * k + 1];
        int lo = max(ij0, ik0), hi = min(ij0 + ijn, ik0 + ikn);
        double acc = 0.0;
        for (int q = lo; q < hi; q++) acc += kmat[koff[j] + (q - ij0)] * kmat[koff[k] + (q - ik0)];
        h[t] = acc;
    }
}

// H = K K^T restricted to overlapping ranges (equdecs products), packed m(m+1)/2, plus lltdec;
// with products, the pairs are summed here (one workgroup), else by k_band_hprod before
__global__ void k_band_h(int32_t m, const int *cb, const int64_t *koff, const double *kmat, double *h, const int *hind,
                         int *nrkd2, int products)
{
    // one workgroup; pairs (j>=k) by threads
    const int np = products ? (m * m + m) / 2 : 0;
    for (int t = threadIdx.x; t < np; t += blockDim.x) {
        int j = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
        while ((j * j + j) / 2 > t) j--;
        while (((j + 1) * (j + 1) + (j + 1)) / 2 <= t) j++;
        int k = t - (j * j + j) / 2;
        int ij0 = cb[2 * j], ijn = cb[2 * j + 1];
        int ik0 = cb[2 * k], ikn = cb[2 * k + 1];
        int lo = max(ij0, ik0), hi = min(ij0 + ijn, ik0 + ikn);
        double acc = 0.0;
        for (int q = lo; q < hi; q++) acc += kmat[koff[j] + (q - ij0)] * kmat[koff[k] + (q - ik0)];
        h[t] = acc;
    }
    __syncthreads();
    // the factorization below is one thread's recurrence: H staged in LDS when it fits
    extern __shared__ double hs[];
    const int nh = (m * m + m) / 2;
    const bool lds = nh <= 6144;
    if (lds)
        for (int t = threadIdx.x; t < nh; t += blockDim.x) hs[t] = h[t];
    __syncthreads();
    if (threadIdx.x == 0) {
        // lltdec of the dense m x m lower-packed H (india = i(i+1)/2)
        double *C = (lds ? hs : h) - 1;
        const int *IND = hind - 1;
        const double eps = 16.0 * DBL_EPSILON;
        int nr = 0;
        double diag = 0.0;
        if (C[IND[1]] > 0.0) C[IND[1]] = 1.0 / sqrt(C[IND[1]]);
        else { C[IND[1]] = 0.0; nr = -1; }
        for (int k = 2; k <= m; k++) {
            int mk = k - IND[k] + IND[k - 1] + 1;
            for (int j = mk; j <= k; j++) {
                int mj = 1;
                if (j > 1) mj = j - IND[j] + IND[j - 1] + 1;
                int kj = IND[k] - k + j;
                diag = C[IND[j]];
                for (int i = max(mj, mk); i <= j - 1; i++) C[kj] = C[kj] - C[IND[k] - k + i] * C[IND[j] - j + i];
                if (j != k) C[kj] = C[kj] * diag;
            }
            if (C[IND[k]] > eps * diag) C[IND[k]] = 1.0 / sqrt(C[IND[k]]);
            else {
                for (int j = mk; j <= k; j++) C[IND[k] - k + j] = 0.0;
                if (nr == 0) nr = -k;
                else { if (nr < 0) nr = 1; nr++; }
            }
        }
        *nrkd2 = nr;
    }
    __syncthreads();
    if (lds)
        for (int t = threadIdx.x; t < nh; t += blockDim.x) h[t] = hs[t];
}

// equslvs apply (src/mpnum.f90:2613-2678) as one single-lane kernel: the band recurrences are
// sequential; the constraint products are short.
__global__ void k_equslvs(int n, int m, const double *c, const int *india, const int *cb, const int64_t *koff,
                          const double *kmat, const double *h, const int *hind, double *x)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    lltfwd_dev(n, c, india, x);
    if (m > 0) {
        for (int j = 0; j < m; j++) {
            int i0 = cb[2 * j], in = cb[2 * j + 1];
            double s = x[n + j];
            for (int i = 1; i <= in; i++) if (i0 + i <= n) s = s - kmat[koff[j] + i - 1] * x[i0 + i - 1];
            x[n + j] = s;
        }
        lltfwd_dev(m, h, hind, x + n);
        lltbwd_dev(m, h, hind, x + n);
        for (int i = 0; i < m; i++) x[n + i] = -x[n + i];
        for (int j = 0; j < m; j++) {
            int i0 = cb[2 * j], in = cb[2 * j + 1];
            double y = x[n + j];
            for (int i = 1; i <= in; i++) if (i0 + i <= n) x[i0 + i - 1] = x[i0 + i - 1] - kmat[koff[j] + i - 1] * y;
        }
    }
    lltbwd_dev(n, c, india, x);
}

// ---------------------------------------------------------------------------------------------
// host launch helpers
// ---------------------------------------------------------------------------------------------
static inline int nblk(int64_t n, int t = TPB) { int64_t b = (n + t - 1) / t; return (int)(b < 1 ? 1 : (b > 65535 * 16 ? 65535 * 16 : b)); }

void l_mupdat(const DevModel &M, int32_t n, const int32_t *i, const int32_t *j, const double *v, hipStream_t s)
{ if (n > 0) hipLaunchKernelGGL(k_mupdat, dim3(nblk(n)), dim3(TPB), 0, s, M, n, i, j, v); }
void l_get_diag(const DevModel &M, int32_t n, double *d, hipStream_t s)
{ hipLaunchKernelGGL(k_get_diag, dim3(nblk(n)), dim3(TPB), 0, s, M, n, d); }
void l_get_band(const DevModel &M, int32_t n, int32_t mbandw, const int32_t *india, double *c, hipStream_t s)
{ hipLaunchKernelGGL(k_get_band, dim3(nblk(n)), dim3(TPB), 0, s, M, n, mbandw, india, c); }

// dot product into *out (device), deterministic; part: >= 2048 doubles, part[2047] zero between calls
void l_dot(int64_t n, const double *a, const double *b, double *part, double *out, hipStream_t s)
{
    int g = (int)((n + TPB - 1) / TPB);
    if (g > 1024) g = 1024;
    if (g < 1) g = 1;
    hipLaunchKernelGGL(k_dot_fused, dim3(g), dim3(TPB), 0, s, n, a, b, part, out);
}
void l_pass_stats(int64_t n, const double *chi2, const int32_t *ndf, const int8_t *st, double *part, long long *ndfsum,
                  unsigned long long *hist, double *chi_out, hipStream_t s)
{
    int g = (int)((n + TPB - 1) / TPB);
    if (g > 1024) g = 1024;
    if (g < 1) g = 1;
    hipLaunchKernelGGL(k_pass_stats, dim3(g), dim3(TPB), 0, s, n, chi2, ndf, st, part, ndfsum, hist);
    hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(TPB), 0, s, g, part, chi_out);
}

void l_gj_init(int n, const double *v, double *diag, int *used, int *info, hipStream_t s)
{ hipLaunchKernelGGL(k_gj_init, dim3(nblk(n)), dim3(TPB), 0, s, n, v, diag, used, info); }
void l_gj_step(int n, double *v, const double *diag, int *used, int *info, double *vk, double *b, hipStream_t s)
{
    hipLaunchKernelGGL(k_gj_pivot, dim3(1), dim3(1024), 0, s, n, v, diag, used, info, vk, b, 1.0e-10);
    int64_t tot = ((int64_t)n * (n + 1)) / 2;
    int g = nblk(tot);
    if (g > 8192) g = 8192;
    hipLaunchKernelGGL(k_gj_update, dim3(g), dim3(TPB), 0, s, n, v, vk, b, info);
}
void l_gj_finish(int n, double *v, double *b, const int *used, const int *info, hipStream_t s)
{
    int64_t tot = ((int64_t)n * (n + 1)) / 2;
    int g = nblk(tot);
    if (g > 8192) g = 8192;
    hipLaunchKernelGGL(k_gj_finish, dim3(g), dim3(TPB), 0, s, n, v, b, used, info);
}

__global__ void k_zero(int64_t n, double *y);
void l_packed_symv(int n, const double *a, const double *x, double *y, double *scratch, hipStream_t s)
{
    const int nb = (n + ST - 1) / ST;
    const int nseg = (nb + SSEG - 1) / SSEG;
    double *yr = scratch, *z = scratch + (int64_t)nseg * n;
    if (n <= 0) return;
    hipLaunchKernelGGL(k_psymv_tiles, dim3(nb, nseg), dim3(256), 0, s, n, a, x, yr, z);
    hipLaunchKernelGGL(k_psymv_sum, dim3((n + 255) / 256), dim3(256), 0, s, n, yr, z, y);
}
__global__ void k_zero(int64_t n, double *y)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) y[i] = 0.0;
}
// Product variants with the same arithmetic (same summation order, bit-identical results), for
// the autotuner: kind (0 the uniform-group rows pass, 1 the general staged rows pass, 2 the
// uniform-group rows pass with 16-byte loads, for even group sizes) * 16 + W wavefronts per
// workgroup
int block_symv_variants(int32_t ugrp, int32_t maxgrp, int *out)
{
    int n = 0;
    if (maxgrp > 8) return 0;
    for (int w : {2, 4, 8}) {
        if (ugrp >= 1 && ugrp <= 8) out[n++] = 0 * 16 + w;
        if (ugrp >= 2 && ugrp <= 8 && ugrp % 2 == 0) out[n++] = 2 * 16 + w;
        out[n++] = 1 * 16 + w;
    }
    return n;
}
template <int G, int W, bool V2>
static void launch_rows_u(int32_t ngroup, const int64_t *brow, const int32_t *bcol, const int64_t *boff, const double *a,
                          const double *x, double *y, double *tb, int64_t spare, hipStream_t s)
{
    hipLaunchKernelGGL((k_bsymv_rows_u<G, W, V2>), dim3((unsigned)((ngroup + W - 1) / W)), dim3(64 * W), 0, s, ngroup,
                       brow, bcol, boff, a, x, y, tb, spare);
}
template <int W>
static void launch_rows_u_g(int32_t g, bool v2, int32_t ngroup, const int64_t *brow, const int32_t *bcol,
                            const int64_t *boff, const double *a, const double *x, double *y, double *tb, int64_t spare,
                            hipStream_t s)
{
    switch (g) {
    case 1: launch_rows_u<1, W, false>(ngroup, brow, bcol, boff, a, x, y, tb, spare, s); break;
    case 2:
        if (v2) launch_rows_u<2, W, true>(ngroup, brow, bcol, boff, a, x, y, tb, spare, s);
        else launch_rows_u<2, W, false>(ngroup, brow, bcol, boff, a, x, y, tb, spare, s);
        break;
    case 3: launch_rows_u<3, W, false>(ngroup, brow, bcol, boff, a, x, y, tb, spare, s); break;
    case 4:
        if (v2) launch_rows_u<4, W, true>(ngroup, brow, bcol, boff, a, x, y, tb, spare, s);
        else launch_rows_u<4, W, false>(ngroup, brow, bcol, boff, a, x, y, tb, spare, s);
        break;
    case 5: launch_rows_u<5, W, false>(ngroup, brow, bcol, boff, a, x, y, tb, spare, s); break;
    case 6:
        if (v2) launch_rows_u<6, W, true>(ngroup, brow, bcol, boff, a, x, y, tb, spare, s);
        else launch_rows_u<6, W, false>(ngroup, brow, bcol, boff, a, x, y, tb, spare, s);
        break;
    case 7: launch_rows_u<7, W, false>(ngroup, brow, bcol, boff, a, x, y, tb, spare, s); break;
    default:
        if (v2) launch_rows_u<8, W, true>(ngroup, brow, bcol, boff, a, x, y, tb, spare, s);
        else launch_rows_u<8, W, false>(ngroup, brow, bcol, boff, a, x, y, tb, spare, s);
        break;
    }
}
template <int W>
static void launch_rows_st(int32_t ngroup, const int64_t *brow, const int32_t *bcf, const int64_t *boff,
                           const int32_t *gfirst, const double *a, const double *x, double *y, double *tb, int tbw,
                           hipStream_t s)
{
    hipLaunchKernelGGL(k_bsymv_rows_st<W>, dim3((unsigned)((ngroup + W - 1) / W)), dim3(64 * W), 0, s, ngroup, brow, bcf,
                       boff, gfirst, a, x, y, tb, tbw);
}
void l_block_symv(int32_t ngroup, const int64_t *brow, const int32_t *bcf, const int64_t *boff, const int32_t *gfirst,
                  const int64_t *ccol_ptr, const int64_t *cc_off, const int32_t *cc_fn, const double *a,
                  const double *x, double *y, int32_t nvar, int64_t nall, int32_t maxgrp, const int32_t *cc_blk,
                  double *tb, int32_t ugrp, const int32_t *bcol, int64_t nblock, int variant, hipStream_t s)
{
    // rows beyond the parameter groups (Lagrange multipliers) start at zero
    if (nall > nvar) hipLaunchKernelGGL(k_zero, dim3(nblk(nall - nvar)), dim3(TPB), 0, s, nall - nvar, y + nvar);
    if (maxgrp <= 8 && tb && cc_blk) {
        // the matrix is read once: rows pass with the transposed block products (tbw values per
        // block), then their sums over the column lists
        const bool uni = ugrp >= 1 && ugrp <= 8 && bcol;
        int kind = variant >= 0 ? variant / 16 : (uni ? 0 : 1), w = variant >= 0 ? variant % 16 : 4;
        if (kind == 2 && (!uni || ugrp % 2 != 0)) kind = 0;
        if (kind == 0 && !uni) kind = 1;
        const int tbw = kind != 1 ? ugrp : maxgrp;
        if (kind != 1) {
            const int64_t spare = (int64_t)ugrp * nblock;
            const bool v2 = kind == 2;
            if (w == 2) launch_rows_u_g<2>(ugrp, v2, ngroup, brow, bcol, boff, a, x, y, tb, spare, s);
            else if (w == 8) launch_rows_u_g<8>(ugrp, v2, ngroup, brow, bcol, boff, a, x, y, tb, spare, s);
            else launch_rows_u_g<4>(ugrp, v2, ngroup, brow, bcol, boff, a, x, y, tb, spare, s);
        } else {
            if (w == 2) launch_rows_st<2>(ngroup, brow, bcf, boff, gfirst, a, x, y, tb, tbw, s);
            else if (w == 8) launch_rows_st<8>(ngroup, brow, bcf, boff, gfirst, a, x, y, tb, tbw, s);
            else launch_rows_st<4>(ngroup, brow, bcf, boff, gfirst, a, x, y, tb, tbw, s);
        }
        hipLaunchKernelGGL(k_bsymv_colst, dim3((unsigned)((ngroup + 3) / 4)), dim3(256), 0, s, ngroup, ccol_ptr, cc_blk,
                           gfirst, tb, tbw, y);
        return;
    }
    if (maxgrp <= 8)
        hipLaunchKernelGGL(k_bsymv_rows8, dim3((unsigned)((ngroup + 3) / 4)), dim3(256), 0, s, ngroup, brow, bcf, boff,
                           gfirst, a, x, y, (double *)nullptr);
    else
        hipLaunchKernelGGL(k_bsymv_rows, dim3((unsigned)((ngroup + 3) / 4)), dim3(256), 0, s, ngroup, brow, bcf, boff,
                           gfirst, a, x, y);
    if (maxgrp <= 8)
        hipLaunchKernelGGL(k_bsymv_cols8, dim3((unsigned)((ngroup + 3) / 4)), dim3(256), 0, s, ngroup, ccol_ptr, cc_off,
                           cc_fn, gfirst, a, x, y);
    else
        hipLaunchKernelGGL(k_bsymv_cols, dim3((unsigned)((ngroup + 3) / 4)), dim3(256), 0, s, ngroup, ccol_ptr, cc_off,
                           cc_fn, gfirst, a, x, y);
}
void l_cons_symv(int32_t ncon, int32_t nvar, const int64_t *cptr, const int32_t *cidx, const double *cval, const double *x,
                 double *y, hipStream_t s)
{ if (ncon > 0) hipLaunchKernelGGL(k_cons_symv, dim3(ncon), dim3(TPB), 0, s, ncon, nvar, cptr, cidx, cval, x, y); }

void l_scale_copy(int64_t n, double sc, const double *a, double *o, hipStream_t s)
{ hipLaunchKernelGGL(k_scale_copy, dim3(nblk(n)), dim3(TPB), 0, s, n, sc, a, o); }
void l_mr_scale(int64_t n, const MrState *st, const double *y, double *v, hipStream_t s)
{ hipLaunchKernelGGL(k_mr_scale, dim3(nblk(n)), dim3(TPB), 0, s, n, st, y, v); }
void l_mr_axpy(int64_t n, const MrState *st, int which, const double *r, double *y, hipStream_t s)
{ hipLaunchKernelGGL(k_mr_axpy, dim3(nblk(n)), dim3(TPB), 0, s, n, st, which, r, y); }
void l_mr_step(MrState *st, hipStream_t s) { hipLaunchKernelGGL(k_mr_step, dim3(1), dim3(64), 0, s, st); }
void l_mr_update(int64_t n, const MrState *st, const double *v, double *w, double *w1, double *w2, double *x, hipStream_t s)
{ hipLaunchKernelGGL(k_mr_update, dim3(nblk(n)), dim3(TPB), 0, s, n, st, v, w, w1, w2, x); }
void l_axpy_neg(int64_t n, double c, const double *r, double *y, hipStream_t s)
{ hipLaunchKernelGGL(k_axpy_neg, dim3(nblk(n)), dim3(TPB), 0, s, n, c, r, y); }
void l_minres_update(int64_t n, const double *v, double *w, double *w1, double *w2, double *x, double oldeps, double delta,
                     double denom, double phi, hipStream_t s)
{ hipLaunchKernelGGL(k_minres_update, dim3(nblk(n)), dim3(TPB), 0, s, n, v, w, w1, w2, x, oldeps, delta, denom, phi); }
void l_qlp_minres(int64_t n, const double *v, double *w, double *wl, double *wl2, double *x, double epln, double dlta,
                  double sc, double tau, int upd_w, int upd_x, hipStream_t s)
{ hipLaunchKernelGGL(k_qlp_minres, dim3(nblk(n)), dim3(TPB), 0, s, n, v, w, wl, wl2, x, epln, dlta, sc, tau, upd_w, upd_x); }
void l_qlp_step(int64_t n, double *v, double *w, double *wl, double *wl2, double *x, double *xl2, const QlpStep &c,
                hipStream_t s)
{ hipLaunchKernelGGL(k_qlp_step, dim3(nblk(n)), dim3(TPB), 0, s, n, v, w, wl, wl2, x, xl2, c); }
void l_xpby(int64_t n, const double *z, double beta, double *p, hipStream_t s)
{ hipLaunchKernelGGL(k_xpby, dim3(nblk(n)), dim3(TPB), 0, s, n, z, beta, p); }
void l_copy(int64_t n, const double *a, double *b, hipStream_t s)
{ hipLaunchKernelGGL(k_copy, dim3(nblk(n)), dim3(TPB), 0, s, n, a, b); }
void l_sub(int64_t n, const double *b, const double *ax, double *r, hipStream_t s)
{ hipLaunchKernelGGL(k_sub, dim3(nblk(n)), dim3(TPB), 0, s, n, b, ax, r); }

void l_precons(int32_t n, int32_t p, const double *diag, double *cu, int *nrkd, const int64_t *cptr, const int32_t *cidx,
               const double *cval, double *sm, double *scratch, hipStream_t s)
{
    hipMemsetAsync(nrkd, 0, sizeof(int), s);
    hipLaunchKernelGGL(k_precons_cu, dim3(nblk(n)), dim3(TPB), 0, s, n, diag, cu, nrkd);
    if (p > 0) hipLaunchKernelGGL(k_precons_s, dim3(1), dim3(TPB), 0, s, p, cptr, cidx, cval, cu, n, sm, scratch);
}
void l_presols(int32_t n, int32_t p, const double *cu, const int64_t *cptr, const int32_t *cidx, const double *cval,
               const double *sm, const double *y, double *x, hipStream_t s)
{
    hipLaunchKernelGGL(k_presols_a, dim3(nblk(n)), dim3(TPB), 0, s, (int64_t)n, cu, y, x);
    if (p > 0) hipLaunchKernelGGL(k_presols_c, dim3(1), dim3(TPB), 0, s, p, n, cptr, cidx, cval, cu, sm, y, x);
    hipLaunchKernelGGL(k_presols_b, dim3(nblk(n)), dim3(TPB), 0, s, (int64_t)n, cu, x);
}

// Chunk-parallel band factorization with the result of the sequential one, bit for bit.  The
// rows are cut into chunks of ch rows, one thread each.  Phase 1 factors every chunk from zero
// rows before it (exact for the first chunk); the error this makes in the later chunks decays
// along them, since each row depends on the matrix and the MW-1 rows before it only.  Phase 2
// redoes every later chunk from the phase-1 rows before it and stops at the first MW-1
// consecutive rows equal bit for bit to phase 1 (and of the same dropped-or-not status): from
// there the phase-1 rows are the sequential ones.  A chunk that ends before that sets *unconv,
// and the sequential kernel runs instead.  L: n x MW dense rows, bad: the dropped rows.
template <int MW>
__global__ __launch_bounds__(64) void k_lltdec_chunks(int n, int ch, const double *__restrict__ c,
                                                      const int *__restrict__ india, int iopt, double *L, int8_t *bad,
                                                      int phase, int *unconv)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t k0l = (int64_t)p * ch + 1;
    if (k0l > n || (phase == 2 && p == 0)) return;
    const int k0 = (int)k0l, k1 = (int)min((int64_t)n, k0l + ch - 1);
    double ring[MW][MW];
#pragma unroll
    for (int q = 0; q < MW; q++)
#pragma unroll
        for (int o = 0; o < MW; o++) ring[q][o] = 0.0;
    if (phase == 2)
        for (int j = max(1, k0 - MW + 1); j < k0; j++) {
            const int sj = (j - 1) % MW;
#pragma unroll
            for (int o = 0; o < MW; o++) ring[sj][o] = L[(int64_t)(j - 1) * MW + o];
        }
    int match = 0;
    for (int k = k0; k <= k1; k++) {
        const int indk = india[k - 1], indk1 = k > 1 ? india[k - 2] : 0;
        const int mk = k - indk + indk1 + 1;
        double r[MW];
#pragma unroll
        for (int o = 0; o < MW; o++) {
            const int col = k - MW + 1 + o;
            r[o] = col >= mk ? c[indk - k + col - 1] : 0.0;
        }
        const int s = (k - 1) % MW;
        const int b = llt_row<MW>(k, mk, s, r, ring, iopt);
#pragma unroll
        for (int o = 0; o < MW; o++) ring[s][o] = r[o];
        double *Lk = L + (int64_t)(k - 1) * MW;
        if (phase == 2) {
            bool same = bad[k - 1] == b;
#pragma unroll
            for (int o = 0; o < MW; o++) same = same && __double_as_longlong(Lk[o]) == __double_as_longlong(r[o]);
            if (same) {
                if (++match >= MW - 1) return;
                continue;
            }
            match = 0;
        }
#pragma unroll
        for (int o = 0; o < MW; o++) Lk[o] = r[o];
        bad[k - 1] = (int8_t)b;
    }
    if (phase == 2 && k1 < n) atomicOr(unconv, 1);
}
// the factor rows back into pede's band storage; the dropped rows counted (aux[1]) and the first
// of them (aux[2])
template <int MW>
__global__ __launch_bounds__(256) void k_lltdec_store(int n, double *c, const int *india, const double *L,
                                                      const int8_t *bad, int *aux)
{
    int mc = 0, mf = INT_MAX;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x + 1; k <= n; k += gridDim.x * blockDim.x) {
        const int indk = india[k - 1], indk1 = k > 1 ? india[k - 2] : 0;
        const int mk = k - indk + indk1 + 1;
#pragma unroll
        for (int o = 0; o < MW; o++) {
            const int col = k - MW + 1 + o;
            if (col >= mk) c[indk - k + col - 1] = L[(int64_t)(k - 1) * MW + o];
        }
        if (bad[k - 1]) { mc++; mf = min(mf, k); }
    }
    if (mc) { atomicAdd(&aux[1], mc); atomicMin(&aux[2], mf); }
}
// the rank-defect code of lltdec: 0 none, -k for one dropped row k, else their number
__global__ void k_lltdec_nr(const int *aux, int *nrkd)
{
    if (threadIdx.x == 0) *nrkd = aux[1] == 0 ? 0 : (aux[1] == 1 ? -aux[2] : aux[1]);
}

template <int MW>
static void lltdec_chunked(int n, double *c, const int *india, int *nrkd, int iopt, double *L, int8_t *bad, int *flag,
                           hipStream_t s)
{
    const int ch = 256;
    const int nch = (n + ch - 1) / ch;
    const int init[3] = {0, 0, INT_MAX};   // not-caught-up flag, dropped rows, first dropped row
    int host_flag = 0;
    HIPCHK(hipMemcpyAsync(flag, init, sizeof(init), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_lltdec_chunks<MW>, dim3((nch + 63) / 64), dim3(64), 0, s, n, ch, c, india, iopt, L, bad, 1, flag);
    hipLaunchKernelGGL(k_lltdec_chunks<MW>, dim3((nch + 63) / 64), dim3(64), 0, s, n, ch, c, india, iopt, L, bad, 2, flag);
    HIPCHK(hipMemcpyAsync(&host_flag, flag, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (host_flag == 0) {
        hipLaunchKernelGGL(k_lltdec_store<MW>, dim3((unsigned)std::min(1024, (n + 255) / 256)), dim3(256), 0, s, n, c, india,
                           L, bad, flag);
        hipLaunchKernelGGL(k_lltdec_nr, dim3(1), dim3(64), 0, s, flag, nrkd);
    } else {
        hipLaunchKernelGGL(k_lltdec_d<MW>, dim3(1), dim3(64), 0, s, n, c, india, nrkd, iopt);
    }
}
bool lltdec_parallel(int n, int mbw) { return mbw >= 1 && mbw <= 8 && n >= 4096; }
void l_lltdec_par(int n, double *c, const int *india, int *nrkd, int iopt, int mbw, double *L, int8_t *bad, int *flag,
                  hipStream_t s)
{
    switch (mbw) {
    case 1: lltdec_chunked<1>(n, c, india, nrkd, iopt, L, bad, flag, s); break;
    case 2: lltdec_chunked<2>(n, c, india, nrkd, iopt, L, bad, flag, s); break;
    case 3: lltdec_chunked<3>(n, c, india, nrkd, iopt, L, bad, flag, s); break;
    case 4: lltdec_chunked<4>(n, c, india, nrkd, iopt, L, bad, flag, s); break;
    case 5: lltdec_chunked<5>(n, c, india, nrkd, iopt, L, bad, flag, s); break;
    case 6: lltdec_chunked<6>(n, c, india, nrkd, iopt, L, bad, flag, s); break;
    case 7: lltdec_chunked<7>(n, c, india, nrkd, iopt, L, bad, flag, s); break;
    default: lltdec_chunked<8>(n, c, india, nrkd, iopt, L, bad, flag, s); break;
    }
}
void l_lltdec(int n, double *c, const int *india, int *nrkd, int iopt, int mbw, hipStream_t s)
{
    switch (mbw) {
    case 1: hipLaunchKernelGGL(k_lltdec_d<1>, dim3(1), dim3(64), 0, s, n, c, india, nrkd, iopt); break;
    case 2: hipLaunchKernelGGL(k_lltdec_d<2>, dim3(1), dim3(64), 0, s, n, c, india, nrkd, iopt); break;
    case 3: hipLaunchKernelGGL(k_lltdec_d<3>, dim3(1), dim3(64), 0, s, n, c, india, nrkd, iopt); break;
    case 4: hipLaunchKernelGGL(k_lltdec_d<4>, dim3(1), dim3(64), 0, s, n, c, india, nrkd, iopt); break;
    case 5: hipLaunchKernelGGL(k_lltdec_d<5>, dim3(1), dim3(64), 0, s, n, c, india, nrkd, iopt); break;
    case 6: hipLaunchKernelGGL(k_lltdec_d<6>, dim3(1), dim3(64), 0, s, n, c, india, nrkd, iopt); break;
    case 7: hipLaunchKernelGGL(k_lltdec_d<7>, dim3(1), dim3(64), 0, s, n, c, india, nrkd, iopt); break;
    case 8: hipLaunchKernelGGL(k_lltdec_d<8>, dim3(1), dim3(64), 0, s, n, c, india, nrkd, iopt); break;
    default: hipLaunchKernelGGL(k_lltdec, dim3(1), dim3(256), 0, s, n, c, india, nrkd, iopt, mbw); break;
    }
}
void l_band_k(int32_t ncon, int n, const double *c, const int *india, const int *cb, const int64_t *koff, double *kmat,
              int mbw, hipStream_t s)
{
    if (ncon <= 0) return;
    const dim3 g((ncon + 63) / 64), b(64);
    switch (mbw) {
    case 1: hipLaunchKernelGGL(k_band_k_r<1>, g, b, 0, s, ncon, n, c, india, cb, koff, kmat); break;
    case 2: hipLaunchKernelGGL(k_band_k_r<2>, g, b, 0, s, ncon, n, c, india, cb, koff, kmat); break;
    case 3: hipLaunchKernelGGL(k_band_k_r<3>, g, b, 0, s, ncon, n, c, india, cb, koff, kmat); break;
    case 4: hipLaunchKernelGGL(k_band_k_r<4>, g, b, 0, s, ncon, n, c, india, cb, koff, kmat); break;
    case 5: hipLaunchKernelGGL(k_band_k_r<5>, g, b, 0, s, ncon, n, c, india, cb, koff, kmat); break;
    case 6: hipLaunchKernelGGL(k_band_k_r<6>, g, b, 0, s, ncon, n, c, india, cb, koff, kmat); break;
    case 7: hipLaunchKernelGGL(k_band_k_r<7>, g, b, 0, s, ncon, n, c, india, cb, koff, kmat); break;
    case 8: hipLaunchKernelGGL(k_band_k_r<8>, g, b, 0, s, ncon, n, c, india, cb, koff, kmat); break;
    default: hipLaunchKernelGGL(k_band_k, g, b, 0, s, ncon, n, c, india, cb, koff, kmat); break;
    }
}
void l_band_h(int32_t m, const int *cb, const int64_t *koff, const double *kmat, double *h, const int *hind, int *nrkd2,
              hipStream_t s)
{
    if (m <= 0) return;
    const int np = (m * m + m) / 2;
    hipLaunchKernelGGL(k_band_hprod, dim3((unsigned)std::min(1024, (np + 63) / 64)), dim3(64), 0, s, m, cb, koff, kmat, h);
    const size_t lds = np <= 6144 ? sizeof(double) * np : 0;
    hipLaunchKernelGGL(k_band_h, dim3(1), dim3(TPB), lds, s, m, cb, koff, kmat, h, hind, nrkd2, 0);
}
void l_equslvs(int n, int m, const double *c, const int *india, const int *cb, const int64_t *koff, const double *kmat,
               const double *h, const int *hind, double *x, hipStream_t s)
{ hipLaunchKernelGGL(k_equslvs, dim3(1), dim3(64), 0, s, n, m, c, india, cb, koff, kmat, h, hind, x); }


// ---------------------------------------------------------------------------------------------
// Band preconditioner apply, chunk-parallel ("spike" decomposition of the band recurrences of
// lltfwd/lltbwd, src/mpnum.f90:2237-2336).  Rows are split into chunks of CH rows; each chunk
// solves its rows with a zero boundary, the boundary dependence (w = mbandw-1 values of the
// neighbouring chunk) is linear and precomputed per chunk as an CH x w response matrix, and
// a short carry over the chunks plus a row-parallel correction completes the solve.  Valid
// for every coefficient set of the factor (rows reset by lltdec included), since the
// recurrences are linear in their inputs.
// Layout: row k (1-based) of the factor holds columns k-len+1..k at c[india(k)-k+j-1],
// len = india(k)-india(k-1) (india(0)=0).
// ---------------------------------------------------------------------------------------------
#define SPK_CH 256
#define SPK_WMAX 16

__device__ __forceinline__ int spk_len(const int *india, int k) { return india[k - 1] - (k > 1 ? india[k - 2] : 0); }

// forward: chunk-local solve of rows [k0, k1] (1-based) with boundary values bnd[0..w) holding
// x_{k0-w} .. x_{k0-1}; x holds the right-hand side on entry.
__device__ void spk_fwd_chunk(int k0, int k1, int w, const double *c, const int *india, double *x, const double *bnd)
{
    for (int k = k0; k <= k1; k++) {
        double xk = x[k - 1];
        int len = spk_len(india, k);
        for (int j = k - len + 1; j <= k - 1; j++) {
            double xj = j >= k0 ? x[j - 1] : (bnd ? bnd[j - (k0 - w)] : 0.0);
            xk = xk - c[india[k - 1] - k + j - 1] * xj;
        }
        x[k - 1] = xk * c[india[k - 1] - 1];
    }
}
// backward: chunk-local solve of rows [k1 down to k0] with boundary x_{k1+1}..x_{k1+w} in bnd
__device__ void spk_bwd_chunk(int k0, int k1, int w, int n, const double *c, const int *india, double *x, const double *bnd)
{
    // contributions of the boundary rows k in (k1, k1+w] to x_j, j <= k1
    if (bnd) {
        for (int k = k1 + 1; k <= min(n, k1 + w); k++) {
            double xk = bnd[k - k1 - 1];
            int len = spk_len(india, k);
            for (int j = k - len + 1; j <= k - 1; j++)
                if (j >= k0 && j <= k1) x[j - 1] = x[j - 1] - c[india[k - 1] - k + j - 1] * xk;
        }
    }
    for (int k = k1; k >= k0; k--) {
        double xk = x[k - 1] * (k > 1 ? c[india[k - 1] - 1] : c[india[0] - 1]);
        x[k - 1] = xk;
        if (k == 1) break;
        int len = spk_len(india, k);
        for (int j = k - len + 1; j <= k - 1; j++)
            if (j >= k0) x[j - 1] = x[j - 1] - c[india[k - 1] - k + j - 1] * xk;
    }
}

// response matrices: Sf[chunk][row][q] = forward response of the chunk's rows to boundary e_q,
// Sb likewise for the backward sweep.  One thread per (chunk, q); uses a scratch row buffer.
__global__ void k_spk_build(int n, int w, int nch, const double *c, const int *india, double *Sf, double *Sb, double *scr)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nch * w) return;
    int ch = t / w, q = t % w;
    int k0 = ch * SPK_CH + 1, k1 = min(n, k0 + SPK_CH - 1), L = k1 - k0 + 1;
    double *x = scr + (size_t)t * SPK_CH - (k0 - 1);   // x[k-1] for k in [k0,k1]
    double bnd[SPK_WMAX];
    for (int i = 0; i < w; i++) bnd[i] = (i == q) ? 1.0 : 0.0;
    for (int k = k0; k <= k1; k++) x[k - 1] = 0.0;
    if (ch > 0) spk_fwd_chunk(k0, k1, w, c, india, x, bnd);
    for (int i = 0; i < L; i++) Sf[((size_t)(k0 - 1 + i)) * w + q] = ch > 0 ? x[k0 - 1 + i] : 0.0;
    for (int k = k0; k <= k1; k++) x[k - 1] = 0.0;
    if (k1 < n) spk_bwd_chunk(k0, k1, w, n, c, india, x, bnd);
    for (int i = 0; i < L; i++) Sb[((size_t)(k0 - 1 + i)) * w + q] = k1 < n ? x[k0 - 1 + i] : 0.0;
}

// chunk-local solves with a zero boundary, one wavefront per chunk: the lanes stage the chunk's
// band rows (W+1 values, zero outside the band) and right-hand side in LDS, lane 0 runs the
// recurrence from LDS with the last W values of x in registers, the lanes write the result.
// Terms outside the band or before the chunk are exact zeros, so the recurrence needs no tests
// (the operations of spk_fwd_chunk / spk_bwd_chunk in the same order, up to the sign of zero).
template <int W>
__global__ __launch_bounds__(64) void k_spk_local(int n, int nch, const double *c, const int *india, double *x, int backward)
{
    __shared__ double sc[SPK_CH * (W + 1)];   // row r = k - k0: element (k, k-W+o) at [r*(W+1) + o]
    __shared__ double sx[SPK_CH];
    __shared__ int si[SPK_CH + 1];            // india of rows k0-1 .. k1 (0 for row 0)
    const int ch = blockIdx.x, lane = threadIdx.x;
    if (ch >= nch) return;
    const int k0 = ch * SPK_CH + 1, k1 = min(n, k0 + SPK_CH - 1), L = k1 - k0 + 1;
    // every lane issues all of its loads of a stage before it waits on any of them
    {
        constexpr int NI = (SPK_CH + 1 + 63) / 64, NX = (SPK_CH + 63) / 64;
        int iv[NI];
        double xv[NX];
#pragma unroll
        for (int u = 0; u < NI; u++) {
            const int r = lane + 64 * u, row = k0 - 1 + r;
            const bool ok = r <= L && row >= 1;
            const int ia = india[ok ? row - 1 : k0 - 1];
            iv[u] = ok ? ia : 0;
        }
#pragma unroll
        for (int u = 0; u < NX; u++) {
            const int r = lane + 64 * u;
            const double xa = x[k0 - 1 + (r < L ? r : 0)];
            xv[u] = r < L ? xa : 0.0;
        }
#pragma unroll
        for (int u = 0; u < NI; u++)
            if (lane + 64 * u <= L) si[lane + 64 * u] = iv[u];
#pragma unroll
        for (int u = 0; u < NX; u++)
            if (lane + 64 * u < L) sx[lane + 64 * u] = xv[u];
    }
    __syncthreads();
    constexpr int NB = 8;
    for (int t0 = 0; t0 < L * (W + 1); t0 += 64 * NB) {
        double v[NB];
#pragma unroll
        for (int u = 0; u < NB; u++) {
            // unconditional load at a clamped index (the row's diagonal when outside the band)
            const int t = t0 + lane + 64 * u;
            const bool tv = t < L * (W + 1);
            const int tt = tv ? t : 0;
            const int r = tt / (W + 1), o = tt - r * (W + 1);
            const int k = k0 + r, j = k - W + o;
            const int ik = si[r + 1], len = ik - si[r];
            const bool in = tv && j >= k - len + 1 && j >= k0;
            const double cv = c[in ? ik - k + j - 1 : ik - 1];
            v[u] = in ? cv : 0.0;
        }
#pragma unroll
        for (int u = 0; u < NB; u++) {
            const int t = t0 + lane + 64 * u;
            if (t < L * (W + 1)) sc[t] = v[u];
        }
    }
    __syncthreads();
    if (lane == 0) {
        double xw[W + 1];
#pragma unroll
        for (int q = 0; q <= W; q++) xw[q] = 0.0;
        // the next row's band values (and right-hand side) are read from LDS one row ahead
        double rw[W + 1], nw[W + 1];
        if (!backward) {
            // at row k: xw[q+1] = x_{k-W+q}
#pragma unroll
            for (int q = 0; q <= W; q++) rw[q] = sc[q];
            double xs = sx[0];
            for (int r = 0; r < L; r++) {
                const int rn = r + 1 < L ? r + 1 : r;
#pragma unroll
                for (int q = 0; q <= W; q++) nw[q] = sc[rn * (W + 1) + q];
                const double nxs = sx[rn];
                double xk = xs;
#pragma unroll
                for (int q = 0; q < W; q++) xk = xk - rw[q] * xw[q + 1];
                xk = xk * rw[W];
                sx[r] = xk;
#pragma unroll
                for (int q = 0; q < W; q++) xw[q] = xw[q + 1];
                xw[W] = xk;
#pragma unroll
                for (int q = 0; q <= W; q++) rw[q] = nw[q];
                xs = nxs;
            }
        } else {
            // at row k: xw[q] = pending x_{k-W+q} (q < W), xw[W] = pending x_k
#pragma unroll
            for (int q = 0; q <= W; q++) {
                const int r = L - 1 - W + q;
                xw[q] = r >= 0 ? sx[r] : 0.0;
            }
#pragma unroll
            for (int q = 0; q <= W; q++) rw[q] = sc[(L - 1) * (W + 1) + q];
            for (int r = L - 1; r >= 0; r--) {
                const int rp = r > 0 ? r - 1 : 0;
#pragma unroll
                for (int q = 0; q <= W; q++) nw[q] = sc[rp * (W + 1) + q];
                const int rn = r - 1 - W;
                const double pend = rn >= 0 ? sx[rn] : 0.0;
                const double xk = xw[W] * rw[W];
                sx[r] = xk;
#pragma unroll
                for (int q = 0; q < W; q++) xw[q] = xw[q] - rw[q] * xk;
#pragma unroll
                for (int q = W; q > 0; q--) xw[q] = xw[q - 1];
                xw[0] = pend;
#pragma unroll
                for (int q = 0; q <= W; q++) rw[q] = nw[q];
            }
        }
    }
    __syncthreads();
    for (int r = lane; r < L; r += 64) x[k0 + r - 1] = sx[r];
}

// carry: boundary values of every chunk, sequential over chunks (w x w products).  The data
// of a tile of chunks (their boundary rows of x and of the response matrix S) is staged in LDS
// by the whole workgroup; the first wavefront then runs the chunk-to-chunk recurrence with one
// boundary row per lane (lane i < w: its row of S and value of x in registers, the next chunk's
// row read ahead) and the carried values broadcast with readlane.
#define SPK_LDS_DOUBLES 6144
__device__ __forceinline__ double spk_rdlane(double v, int l)
{
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}
template <int W>
__global__ __launch_bounds__(256) void k_spk_carry(int n, int nch, const double *x, const double *S, double *t, int backward)
{
    constexpr int w = W;
    extern __shared__ double dyn[];
    const int tc = SPK_LDS_DOUBLES / (w + w * w);      // chunks per tile
    double *sx = dyn;                                  // [tc * w]
    double *sS = sx + tc * w;                          // [tc * w * w]
    const int tid = threadIdx.x, lane = tid & 63;
    double cur = 0.0;                                  // carried value of boundary row `lane` (wave 0)
    const int ntile = (nch + tc - 1) / tc;
    for (int tt = 0; tt < ntile; tt++) {
        // tiles in processing order (backward: from the last chunk)
        const int tile = backward ? ntile - 1 - tt : tt;
        const int c0 = tile * tc, c1 = min(nch, c0 + tc);
        __syncthreads();
        for (int e = tid; e < (c1 - c0) * w; e += 256) {
            const int ch = c0 + e / w, i = e % w;
            const int k0 = ch * SPK_CH + 1, k1 = min(n, k0 + SPK_CH - 1);
            const int k = backward ? k0 + i : k1 - w + 1 + i;
            const bool in = backward ? (k <= k1) : (k >= k0);
            sx[e] = in ? x[k - 1] : 0.0;
            for (int q = 0; q < w; q++) sS[e * w + q] = in ? S[(size_t)(k - 1) * w + q] : 0.0;
        }
        __syncthreads();
        if (tid < 64) {
            const int i = lane < w ? lane : 0;
            double xa = 0.0, sa[W];
            auto load = [&](int cc, double &xv, double *sv) {
                const int ch = backward ? c1 - 1 - cc : c0 + cc;
                const int e0 = (ch - c0) * w;
                xv = sx[e0 + i];
#pragma unroll
                for (int q = 0; q < W; q++) sv[q] = sS[(e0 + i) * w + q];
            };
            load(0, xa, sa);
            for (int cc = 0; cc < c1 - c0; cc++) {
                const int ch = backward ? c1 - 1 - cc : c0 + cc;
                const int k0 = ch * SPK_CH + 1, k1 = min(n, k0 + SPK_CH - 1), L = k1 - k0 + 1;
                double xb = 0.0, sb[W];
                if (cc + 1 < c1 - c0) load(cc + 1, xb, sb);   // read ahead
                if (lane < w) t[(size_t)ch * w + lane] = cur;
                const bool coupled = backward ? (k1 < n) : (ch > 0);
                double v = xa;
                if (coupled) {
#pragma unroll
                    for (int q = 0; q < W; q++) v += sa[q] * spk_rdlane(cur, q);
                }
                if (L < w) {
                    // chunk shorter than w (the last one): rows outside it carry through
                    const bool in = backward ? (lane < L) : (w - 1 - lane < L);
                    const int src = backward ? lane - L : lane + L;
                    const double through = __shfl(cur, src >= 0 && src < 64 ? src : 0, 64);
                    v = in ? v : (lane < w ? through : 0.0);
                }
                cur = lane < w ? v : 0.0;
                xa = xb;
#pragma unroll
                for (int q = 0; q < W; q++) sa[q] = sb[q];
            }
        }
    }
}

// Carry as a parallel scan.  The boundary values entering chunk k (in processing order) follow
// s_{k+1} = A_k s_k + b_k, with A_k the w x w response of chunk k's boundary rows to the
// incoming boundary (zero for the first chunk processed, which has none) and b_k its local
// solution there.  Composing the maps by doubling (Kogge-Stone): at level l every chunk k >= 2^l
// takes b_k += A_k^(l) b_{k-2^l}, A_k^(l) being the product of the maps (k-2^l, k].  The A^(l)
// depend only on the band factor and are formed once (k_spk_scan_a0, k_spk_scan_up); an
// application does the ~log2(nch) levels of w x w products on the vectors, in LDS.
__device__ __forceinline__ int spk_brow(int n, int w, int nch, int k, int i, int backward, bool &in)
{
    const int ch = backward ? nch - 1 - k : k;
    const int k0 = ch * SPK_CH + 1, k1 = min(n, k0 + SPK_CH - 1);
    const int row = backward ? k0 + i : k1 - w + 1 + i;
    in = backward ? (row <= k1) : (row >= k0);
    return row;
}
__global__ void k_spk_scan_a0(int n, int w, int nch, const double *S, double *A, int backward)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nch * w * w) return;
    const int k = e / (w * w), i = (e / w) % w, q = e % w;
    bool in;
    const int row = spk_brow(n, w, nch, k, i, backward, in);
    A[e] = (k > 0 && in) ? S[(size_t)(row - 1) * w + q] : 0.0;
}
__global__ void k_spk_scan_up(int w, int nch, const double *A, double *An, int step)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nch * w * w) return;
    const int k = e / (w * w), i = (e / w) % w, q = e % w;
    if (k < step) { An[e] = A[e]; return; }
    double v = 0.0;
    for (int r = 0; r < w; r++) v += A[((size_t)k * w + i) * w + r] * A[((size_t)(k - step) * w + r) * w + q];
    An[e] = v;
}
template <int W>
__global__ __launch_bounds__(1024) void k_spk_scan(int n, int nch, const double *x, const double *A, int nlev, double *t,
                                                   int backward)
{
    extern __shared__ double dyn[];
    double *b0 = dyn, *b1 = dyn + (size_t)nch * W;
    const int ne = nch * W;
    for (int e = threadIdx.x; e < ne; e += blockDim.x) {
        bool in;
        const int row = spk_brow(n, W, nch, e / W, e % W, backward, in);
        b0[e] = in ? x[row - 1] : 0.0;
    }
    for (int l = 0; l < nlev; l++) {
        __syncthreads();
        const int step = 1 << l;
        const double *Al = A + (size_t)l * ne * W;
        for (int e = threadIdx.x; e < ne; e += blockDim.x) {
            const int k = e / W, i = e - k * W;
            double v = b0[e];
            if (k >= step) {
                const double *a = Al + ((size_t)k * W + i) * W;
                const double *bp = b0 + (k - step) * W;
#pragma unroll
                for (int q = 0; q < W; q++) v += a[q] * bp[q];
            }
            b1[e] = v;
        }
        double *tmp = b0;
        b0 = b1;
        b1 = tmp;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < ne; e += blockDim.x) {
        const int k = e / W, i = e - k * W;
        const int ch = backward ? nch - 1 - k : k;
        t[(size_t)ch * W + i] = k > 0 ? b0[(k - 1) * W + i] : 0.0;
    }
}

__global__ void k_spk_fix(int n, int w, const double *S, const double *t, double *x, int backward)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (k > n) return;
    int ch = (k - 1) / SPK_CH;
    if (!backward && ch == 0) return;
    if (backward && min(n, ch * SPK_CH + SPK_CH) >= n) return;
    double v = x[k - 1];
    for (int q = 0; q < w; q++) v += S[(size_t)(k - 1) * w + q] * t[(size_t)ch * w + q];
    x[k - 1] = v;
}

// constraint part of equslvs, parallel: d_j = x[n+j] - K_j . x ; small solve ; x -= K^T lambda
__global__ void k_eqs_cdot(int n, const int *cb, const int64_t *koff, const double *kmat, const double *x, double *d)
{
    int j = blockIdx.x;
    __shared__ double red[TPB];
    int i0 = cb[2 * j], in = cb[2 * j + 1];
    const double *kj = kmat + koff[j];
    double acc = 0.0;
    // eight terms' loads issued together, the terms added in the same order as one at a time
    for (int i = threadIdx.x; i < in; i += TPB * 8) {
        double kv[8], xv[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int iu = i + u * TPB;
            const bool ok = iu < in && i0 + iu < n;
            kv[u] = kj[ok ? iu : 0];                 // clamped: loads unconditional
            xv[u] = x[ok ? i0 + iu : 0];
        }
#pragma unroll
        for (int u = 0; u < 8; u++)
            if (i + u * TPB < in && i0 + i + u * TPB < n) acc += kv[u] * xv[u];
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = TPB / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) d[j] = x[n + j] - red[0];
}
// the constraint block's factor and right-hand side are staged in LDS when they fit (the
// substitutions then run on LDS instead of global memory)
#define EQS_LDS 4096
__global__ __launch_bounds__(64) void k_eqs_small(int n, int m, const double *h, const int *hind, const double *d, double *x)
{
    __shared__ double sh_h[EQS_LDS];
    __shared__ int sh_i[EQS_LDS / 2];
    const int nh = m > 0 ? hind[m - 1] : 0;
    const bool fit = nh + m <= EQS_LDS && m <= EQS_LDS / 2;
    if (fit) {
        // loads in batches of 16 per lane, all issued before the first is used
        for (int t0 = threadIdx.x; t0 < nh; t0 += 64 * 16) {
            double v[16];
#pragma unroll
            for (int u = 0; u < 16; u++) v[u] = h[t0 + 64 * u < nh ? t0 + 64 * u : nh - 1];
#pragma unroll
            for (int u = 0; u < 16; u++)
                if (t0 + 64 * u < nh) sh_h[t0 + 64 * u] = v[u];
        }
        for (int t = threadIdx.x; t < m; t += 64) {
            sh_h[nh + t] = d[t];
            sh_i[t] = hind[t];
        }
        __syncthreads();
        if (m <= 64) {
            // one row per lane, both substitutions column by column: step j finalises x_j and
            // subtracts its column from the rows below (forward), or x_k's column from the rows
            // above (backward); every x_i receives its terms in the order of lltfwd / lltbwd
            const int lane = threadIdx.x, k = lane + 1;
            double xv = lane < m ? sh_h[nh + lane] : 0.0;
            const int indk = lane < m ? sh_i[lane] : 0, indk1 = (lane >= 1 && lane < m) ? sh_i[lane - 1] : 0;
            const int fk = k - indk + indk1 + 1;       // first column of row k
            for (int j = 1; j <= m; j++) {
                if (lane == j - 1) xv = xv * sh_h[indk - 1];
                const double xj = spk_rdlane(xv, j - 1);
                if (lane < m && k > j && j >= fk) xv = xv - sh_h[indk - k + j - 1] * xj;
            }
            for (int kk = m; kk >= 2; kk--) {
                const int ik = sh_i[kk - 1], ik1 = sh_i[kk - 2];
                if (lane == kk - 1) xv = xv * sh_h[ik - 1];
                const double xk = spk_rdlane(xv, kk - 1);
                const int fkk = kk - ik + ik1 + 1;
                if (k < kk && k >= fkk) xv = xv - sh_h[ik - kk + k - 1] * xk;
            }
            if (lane == 0) xv = xv * sh_h[indk - 1];
            if (lane < m) x[n + lane] = -xv;
            return;
        }
        if (threadIdx.x == 0) {
            double *xl = sh_h + nh;
            lltfwd_dev(m, sh_h, sh_i, xl);
            lltbwd_dev(m, sh_h, sh_i, xl);
        }
        __syncthreads();
        for (int t = threadIdx.x; t < m; t += 64) x[n + t] = -sh_h[nh + t];
        return;
    }
    if (threadIdx.x != 0) return;
    for (int j = 0; j < m; j++) x[n + j] = d[j];
    lltfwd_dev(m, h, hind, x + n);
    lltbwd_dev(m, h, hind, x + n);
    for (int i = 0; i < m; i++) x[n + i] = -x[n + i];
}
// x_i -= K_ji lambda_j over the constraints covering i in ascending j (the order of equslvs'
// sequential loop), one thread per parameter: deterministic, no atomics
__global__ __launch_bounds__(TPB) void k_eqs_upd(int n, int m, const int *cb, const int64_t *koff, const double *kmat,
                                                  double *x)
{
    // the constraints' ranges, offsets and multipliers are staged in LDS, TPB at a time; only the
    // constraints whose range meets the workgroup's parameters are kept
    __shared__ int s0[TPB], sn[TPB];
    __shared__ int64_t so[TPB];
    __shared__ double sl[TPB];
    __shared__ int cnt, wcnt[TPB / 64];
    const int ib = blockIdx.x * TPB, i = ib + threadIdx.x;
    double xi = i < n ? x[i] : 0.0;
    for (int j0 = 0; j0 < m; j0 += TPB) {
        __syncthreads();
        const int j = j0 + threadIdx.x;
        int a = 0, b = 0;
        if (j < m) { a = cb[2 * j]; b = cb[2 * j + 1]; }
        const bool meets = j < m && a < ib + TPB && a + b > ib;
        // keep ascending constraint order: position = number of meeting constraints before j
        const unsigned long long bal = __ballot(meets);
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
        if (lane == 0) wcnt[wv] = __popcll(bal);
        __syncthreads();
        int pos = __popcll(bal & ((1ull << lane) - 1ull));
        for (int w = 0; w < wv; w++) pos += wcnt[w];
        if (meets) {
            s0[pos] = a;
            sn[pos] = b;
            so[pos] = koff[j];
            sl[pos] = x[n + j];
        }
        if (threadIdx.x == TPB - 1) cnt = pos + (meets ? 1 : 0);
        __syncthreads();
        const int nc = cnt;
        if (i < n)
            for (int c = 0; c < nc; c++)
                if (i >= s0[c] && i < s0[c] + sn[c]) xi = xi - kmat[so[c] + (i - s0[c])] * sl[c];
    }
    if (i < n) x[i] = xi;
}

namespace {
inline int g1(int64_t n, int t = TPB) { return (int)((n + t - 1) / t); }
}

static void spk_local(int n, int w, int nch, const double *c, const int *india, double *x, int backward, hipStream_t s)
{
#define SPK_LOCAL(W) hipLaunchKernelGGL(k_spk_local<W>, dim3(nch), dim3(64), 0, s, n, nch, c, india, x, backward)
    switch (w) {
    case 1: SPK_LOCAL(1); break;
    case 2: SPK_LOCAL(2); break;
    case 3: SPK_LOCAL(3); break;
    case 4: SPK_LOCAL(4); break;
    case 5: SPK_LOCAL(5); break;
    case 6: SPK_LOCAL(6); break;
    case 7: SPK_LOCAL(7); break;
    case 8: SPK_LOCAL(8); break;
    case 9: SPK_LOCAL(9); break;
    case 10: SPK_LOCAL(10); break;
    case 11: SPK_LOCAL(11); break;
    case 12: SPK_LOCAL(12); break;
    case 13: SPK_LOCAL(13); break;
    case 14: SPK_LOCAL(14); break;
    case 15: SPK_LOCAL(15); break;
    default: SPK_LOCAL(16); break;
    }
#undef SPK_LOCAL
}

static void spk_carry(int n, int w, int nch, const double *x, const double *S, double *t, int backward, hipStream_t s)
{
    const size_t lds = SPK_LDS_DOUBLES * sizeof(double);
#define SPK_CARRY(W) hipLaunchKernelGGL(k_spk_carry<W>, dim3(1), dim3(256), lds, s, n, nch, x, S, t, backward)
    switch (w) {
    case 1: SPK_CARRY(1); break;
    case 2: SPK_CARRY(2); break;
    case 3: SPK_CARRY(3); break;
    case 4: SPK_CARRY(4); break;
    case 5: SPK_CARRY(5); break;
    case 6: SPK_CARRY(6); break;
    case 7: SPK_CARRY(7); break;
    case 8: SPK_CARRY(8); break;
    case 9: SPK_CARRY(9); break;
    case 10: SPK_CARRY(10); break;
    case 11: SPK_CARRY(11); break;
    case 12: SPK_CARRY(12); break;
    case 13: SPK_CARRY(13); break;
    case 14: SPK_CARRY(14); break;
    case 15: SPK_CARRY(15); break;
    default: SPK_CARRY(16); break;
    }
#undef SPK_CARRY
}

void l_spk_build(int n, int w, const double *c, const int *india, double *Sf, double *Sb, double *scr, hipStream_t s)
{
    int nch = (n + SPK_CH - 1) / SPK_CH;
    hipLaunchKernelGGL(k_spk_build, dim3(g1((int64_t)nch * w, 64)), dim3(64), 0, s, n, w, nch, c, india, Sf, Sb, scr);
}
int spk_chunk_rows() { return SPK_CH; }
// scan levels of the carry (nlev = ceil(log2 nch)); A holds nlev * nch * w * w values per direction
int spk_scan_levels(int n, int w)
{
    const int nch = (n + SPK_CH - 1) / SPK_CH;
    if (nch < 2 || w > 8 || (size_t)2 * nch * w * sizeof(double) > 144 * 1024) return 0;
    int l = 0;
    while ((1 << l) < nch) l++;
    return l;
}
void l_spk_scan_build(int n, int w, const double *Sf, const double *Sb, double *Af, double *Ab, int nlev, hipStream_t s)
{
    const int nch = (n + SPK_CH - 1) / SPK_CH;
    const size_t lv = (size_t)nch * w * w;
    const unsigned g = g1((int64_t)lv);
    hipLaunchKernelGGL(k_spk_scan_a0, dim3(g), dim3(TPB), 0, s, n, w, nch, Sf, Af, 0);
    hipLaunchKernelGGL(k_spk_scan_a0, dim3(g), dim3(TPB), 0, s, n, w, nch, Sb, Ab, 1);
    for (int l = 0; l + 1 < nlev; l++) {
        hipLaunchKernelGGL(k_spk_scan_up, dim3(g), dim3(TPB), 0, s, w, nch, Af + l * lv, Af + (l + 1) * lv, 1 << l);
        hipLaunchKernelGGL(k_spk_scan_up, dim3(g), dim3(TPB), 0, s, w, nch, Ab + l * lv, Ab + (l + 1) * lv, 1 << l);
    }
}
static void spk_scan(int n, int w, int nch, const double *x, const double *A, int nlev, double *t, int backward,
                     hipStream_t s)
{
    const size_t lds = (size_t)2 * nch * w * sizeof(double);
#define SPK_SCAN(W)                                                                                                   \
    do {                                                                                                              \
        if (lds > 65536) (void)hipFuncSetAttribute((const void *)k_spk_scan<W>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds); \
        hipLaunchKernelGGL(k_spk_scan<W>, dim3(1), dim3(1024), lds, s, n, nch, x, A, nlev, t, backward);                \
    } while (0)
    switch (w) {
    case 1: SPK_SCAN(1); break;
    case 2: SPK_SCAN(2); break;
    case 3: SPK_SCAN(3); break;
    case 4: SPK_SCAN(4); break;
    case 5: SPK_SCAN(5); break;
    case 6: SPK_SCAN(6); break;
    case 7: SPK_SCAN(7); break;
    default: SPK_SCAN(8); break;
    }
#undef SPK_SCAN
}

void l_equslvs_spike(int n, int m, int w, const double *c, const int *india, const double *Sf, const double *Sb,
                     double *tbuf, const int *cb, const int64_t *koff, const double *kmat, const double *h, const int *hind,
                     double *dtmp, double *x, const double *Af, const double *Ab, int nlev, hipStream_t s)
{
    int nch = (n + SPK_CH - 1) / SPK_CH;
    spk_local(n, w, nch, c, india, x, 0, s);
    if (Af && nlev > 0) spk_scan(n, w, nch, x, Af, nlev, tbuf, 0, s);
    else spk_carry(n, w, nch, x, Sf, tbuf, 0, s);
    hipLaunchKernelGGL(k_spk_fix, dim3(g1(n)), dim3(TPB), 0, s, n, w, Sf, tbuf, x, 0);
    if (m > 0) {
        hipLaunchKernelGGL(k_eqs_cdot, dim3(m), dim3(TPB), 0, s, n, cb, koff, kmat, x, dtmp);
        hipLaunchKernelGGL(k_eqs_small, dim3(1), dim3(64), 0, s, n, m, h, hind, dtmp, x);
        hipLaunchKernelGGL(k_eqs_upd, dim3(g1(n)), dim3(TPB), 0, s, n, m, cb, koff, kmat, x);
    }
    spk_local(n, w, nch, c, india, x, 1, s);
    if (Ab && nlev > 0) spk_scan(n, w, nch, x, Ab, nlev, tbuf, 1, s);
    else spk_carry(n, w, nch, x, Sb, tbuf, 1, s);
    hipLaunchKernelGGL(k_spk_fix, dim3(g1(n)), dim3(TPB), 0, s, n, w, Sb, tbuf, x, 1);
}

} // namespace mpgpu
