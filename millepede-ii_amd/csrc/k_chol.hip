// Dense blocked Cholesky (A = L L^T) for 'method decomposition' and the FP32 variant used
// with FP64 iterative refinement.  Column-major n x n, lower triangle, tile NB = 64.
// Right-looking: factor diagonal tile in LDS, triangular solve of the panel, then the
// trailing update with FP64 (or FP32) MFMA tiles on the matrix cores.
#include "mpgpu_device.h"
#include <float.h>

namespace mpgpu {

#define NB 64

// packed lower-row (pede) -> full column-major lower part (and mirror)
__global__ void k_packed_to_full(int n, const double *p, double *a, int lda)
{
    int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int64_t tot = (int64_t)n * (n + 1) / 2;
    if (t >= tot) return;
    int64_t i = (int64_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while (i * (i + 1) / 2 > t) i--;
    while ((i + 1) * (i + 2) / 2 <= t) i++;
    int64_t j = t - i * (i + 1) / 2; // 0-based, j <= i
    double v = p[t];
    a[j * lda + i] = v;
    a[i * lda + j] = v;
}

__global__ void k_full_to_f32(int64_t nn, const double *a, float *b)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nn; i += (int64_t)gridDim.x * blockDim.x) b[i] = (float)a[i];
}

// unblocked Cholesky of the diagonal tile (k0..k0+nb) in LDS; info = first failing column+1
template <typename T>
__global__ __launch_bounds__(256) void k_potrf_tile(int n, T *a, int lda, int k0, int *info)
{
    __shared__ T t[NB][NB + 1];
    const int nb = min(NB, n - k0);
    for (int e = threadIdx.x; e < nb * nb; e += blockDim.x) {
        int i = e % nb, j = e / nb;
        t[i][j] = a[(int64_t)(k0 + j) * lda + (k0 + i)];
    }
    __syncthreads();
    __shared__ int bad;
    if (threadIdx.x == 0) bad = 0;
    __syncthreads();
    for (int j = 0; j < nb; j++) {
        if (threadIdx.x == 0) {
            T d = t[j][j];
            if (!(d > (T)0)) { bad = 1; if (info[0] == 0) info[0] = k0 + j + 1; d = (T)1; }
            t[j][j] = sqrt(d);
        }
        __syncthreads();
        T djj = t[j][j];
        for (int i = j + 1 + threadIdx.x; i < nb; i += blockDim.x) t[i][j] = t[i][j] / djj;
        __syncthreads();
        // rank-1 update of the remaining lower part of the tile
        for (int e = threadIdx.x; e < (nb - j - 1) * (nb - j - 1); e += blockDim.x) {
            int i = j + 1 + e % (nb - j - 1), c = j + 1 + e / (nb - j - 1);
            if (i >= c) t[i][c] = t[i][c] - t[i][j] * t[c][j];
        }
        __syncthreads();
    }
    for (int e = threadIdx.x; e < nb * nb; e += blockDim.x) {
        int i = e % nb, j = e / nb;
        if (i >= j) a[(int64_t)(k0 + j) * lda + (k0 + i)] = t[i][j];
    }
}

// panel: rows below the diagonal tile, X L_kk^T = A_ik  -> one thread per row
template <typename T>
__global__ __launch_bounds__(NB) void k_trsm_panel(int n, T *a, int lda, int k0)
{
    __shared__ T l[NB][NB + 1];
    const int nb = min(NB, n - k0);
    for (int e = threadIdx.x; e < nb * nb; e += blockDim.x) {
        int i = e % nb, j = e / nb;
        l[i][j] = a[(int64_t)(k0 + j) * lda + (k0 + i)];
    }
    __syncthreads();
    int row = k0 + nb + blockIdx.x * NB + threadIdx.x;
    if (row >= n) return;
    T x[NB];
    for (int j = 0; j < nb; j++) x[j] = a[(int64_t)(k0 + j) * lda + row];
    for (int j = 0; j < nb; j++) {
        T s = x[j];
        for (int q = 0; q < j; q++) s -= x[q] * l[j][q];
        x[j] = s / l[j][j];
    }
    for (int j = 0; j < nb; j++) a[(int64_t)(k0 + j) * lda + row] = x[j];
}

// trailing update A_ij -= L_ik L_jk^T for tiles i >= j of the trailing matrix.
// 256 threads = 4 waves; each wave computes a 32x32 quarter of the 64x64 tile with
// 2x2 MFMA 16x16 accumulators (FP64: v_mfma_f64_16x16x4f64; FP32: v_mfma_f32_16x16x4f32).
typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

template <typename T> struct Mfma;
// Operands: lane l supplies A[row = l%16][k = l/16] and B[k = l/16][col = l%16].  Result
// element e of lane l is C[row][col = l%16] with row = l/16 + 4e for the FP64 instruction and
// row = 4*(l/16) + e for the FP32 one (measured on gfx950).
template <> struct Mfma<double> {
    typedef d4 acc_t;
    static __device__ __forceinline__ acc_t op(double a, double b, acc_t c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    static __device__ __forceinline__ int row(int lk, int e) { return lk + 4 * e; }
};
template <> struct Mfma<float> {
    typedef f4 acc_t;
    static __device__ __forceinline__ acc_t op(float a, float b, acc_t c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    static __device__ __forceinline__ int row(int lk, int e) { return 4 * lk + e; }
};

template <typename T>
__global__ __launch_bounds__(256) void k_syrk_update(int n, T *a, int lda, int k0)
{
    const int nb = min(NB, n - k0);
    const int t0 = (k0 + nb);              // first trailing row/col
    const int nt = (n - t0 + NB - 1) / NB; // trailing tiles per dimension
    int64_t tile = blockIdx.x;
    int ti = (int)((sqrt(8.0 * (double)tile + 1.0) - 1.0) * 0.5);
    while ((int64_t)ti * (ti + 1) / 2 > tile) ti--;
    while ((int64_t)(ti + 1) * (ti + 2) / 2 <= tile) ti++;
    int tj = (int)(tile - (int64_t)ti * (ti + 1) / 2);
    if (ti >= nt) return;
    const int r0 = t0 + ti * NB, c0 = t0 + tj * NB;
    __shared__ T la[NB][NB + 1]; // L_ik rows r0.., cols k0..
    __shared__ T lb[NB][NB + 1]; // L_jk rows c0..
    for (int e = threadIdx.x; e < NB * NB; e += blockDim.x) {
        int i = e % NB, k = e / NB;
        la[i][k] = (r0 + i < n && k < nb) ? a[(int64_t)(k0 + k) * lda + r0 + i] : (T)0;
        lb[i][k] = (c0 + i < n && k < nb) ? a[(int64_t)(k0 + k) * lda + c0 + i] : (T)0;
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    typedef typename Mfma<T>::acc_t acc_t;
    acc_t acc[2][2];
    for (int p = 0; p < 2; p++) for (int q = 0; q < 2; q++) for (int e = 0; e < 4; e++) acc[p][q][e] = (T)0;
    const int lr = lane & 15, lk = lane >> 4;
    for (int kk = 0; kk < NB; kk += 4) {
        for (int p = 0; p < 2; p++) {
            T av = la[wr + p * 16 + lr][kk + lk];
            for (int q = 0; q < 2; q++) {
                T bv = lb[wc + q * 16 + lr][kk + lk];
                acc[p][q] = Mfma<T>::op(av, bv, acc[p][q]);
            }
        }
    }
    for (int p = 0; p < 2; p++)
        for (int q = 0; q < 2; q++)
            for (int e = 0; e < 4; e++) {
                int i = r0 + wr + p * 16 + Mfma<T>::row(lk, e);
                int j = c0 + wc + q * 16 + lr;
                if (i < n && j < n && i >= j) {
                    int64_t o = (int64_t)j * lda + i;
                    a[o] = a[o] - acc[p][q][e];
                }
            }
}

// forward / backward substitution with the factor (column-major lower L)
// forward: L y = b ; tile-by-tile: solve diagonal tile (one block), then update the rest
template <typename T, typename V>
__global__ __launch_bounds__(256) void k_trsv_tile(int n, const T *a, int lda, int k0, V *x, int64_t ldx, int trans)
{
    x += blockIdx.y * ldx;   // right-hand side of this block
    __shared__ T l[NB][NB + 1];
    __shared__ V xs[NB];
    const int nb = min(NB, n - k0);
    for (int e = threadIdx.x; e < nb * nb; e += blockDim.x) {
        int i = e % nb, j = e / nb;
        l[i][j] = a[(int64_t)(k0 + j) * lda + (k0 + i)];
    }
    for (int i = threadIdx.x; i < nb; i += blockDim.x) xs[i] = x[k0 + i];
    __syncthreads();
    if (threadIdx.x == 0) {
        if (!trans) {
            for (int j = 0; j < nb; j++) {
                V s = xs[j];
                for (int q = 0; q < j; q++) s -= (V)l[j][q] * xs[q];
                xs[j] = s / (V)l[j][j];
            }
        } else {
            for (int j = nb - 1; j >= 0; j--) {
                V s = xs[j];
                for (int q = j + 1; q < nb; q++) s -= (V)l[q][j] * xs[q];
                xs[j] = s / (V)l[j][j];
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nb; i += blockDim.x) x[k0 + i] = xs[i];
}
// forward update: x[i] -= sum_{j in tile} L[i][j] x[j] for i >= k0+nb
template <typename T, typename V>
__global__ void k_gemv_fwd(int n, const T *a, int lda, int k0, V *x, int64_t ldx)
{
    x += blockIdx.y * ldx;
    const int nb = min(NB, n - k0);
    int i = k0 + nb + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    V s = x[i];
    for (int j = 0; j < nb; j++) s -= (V)a[(int64_t)(k0 + j) * lda + i] * x[k0 + j];
    x[i] = s;
}
// backward update: x[j] -= sum_{i in tile} L[i][j] x[i] for j < k0
template <typename T, typename V>
__global__ void k_gemv_bwd(int n, const T *a, int lda, int k0, V *x, int64_t ldx)
{
    x += blockIdx.y * ldx;
    const int nb = min(NB, n - k0);
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= k0) return;
    V s = x[j];
    for (int i = 0; i < nb; i++) s -= (V)a[(int64_t)j * lda + k0 + i] * x[k0 + i];
    x[j] = s;
}

template <typename T> static int chol_factor_t(int n, T *a, int lda, int *info_dev, hipStream_t s)
{
    hipMemsetAsync(info_dev, 0, sizeof(int), s);
    for (int k0 = 0; k0 < n; k0 += NB) {
        hipLaunchKernelGGL(k_potrf_tile<T>, dim3(1), dim3(256), 0, s, n, a, lda, k0, info_dev);
        int nb = min(NB, n - k0);
        int rest = n - k0 - nb;
        if (rest <= 0) break;
        hipLaunchKernelGGL(k_trsm_panel<T>, dim3((rest + NB - 1) / NB), dim3(NB), 0, s, n, a, lda, k0);
        int64_t nt = (rest + NB - 1) / NB;
        hipLaunchKernelGGL(k_syrk_update<T>, dim3((unsigned)(nt * (nt + 1) / 2)), dim3(256), 0, s, n, a, lda, k0);
    }
    int info = 0;
    hipMemcpyAsync(&info, info_dev, sizeof(int), hipMemcpyDeviceToHost, s);
    hipStreamSynchronize(s);
    return info;
}

int chol_factor(int n, double *a, int lda, double *work, int *info_dev, hipStream_t s, int fp32, float *a32)
{
    (void)work;
    if (!fp32) return chol_factor_t<double>(n, a, lda, info_dev, s);
    int64_t nn = (int64_t)lda * n;
    hipLaunchKernelGGL(k_full_to_f32, dim3(4096), dim3(256), 0, s, nn, a, a32);
    return chol_factor_t<float>(n, a32, lda, info_dev, s);
}

// nrhs right-hand sides, column q at x + q * ldx, solved side by side (grid dimension y)
template <typename T, typename V>
static void chol_solve_t(int n, const T *a, int lda, V *x, int nrhs, int64_t ldx, hipStream_t s, int first = 0)
{
    // right-hand sides that are zero above row 'first': the forward sweep starts at its tile
    for (int k0 = (first / NB) * NB; k0 < n; k0 += NB) {
        hipLaunchKernelGGL((k_trsv_tile<T, V>), dim3(1, nrhs), dim3(256), 0, s, n, a, lda, k0, x, ldx, 0);
        int rest = n - k0 - min(NB, n - k0);
        if (rest > 0)
            hipLaunchKernelGGL((k_gemv_fwd<T, V>), dim3((rest + 255) / 256, nrhs), dim3(256), 0, s, n, a, lda, k0, x, ldx);
    }
    int last = ((n - 1) / NB) * NB;
    for (int k0 = last; k0 >= 0; k0 -= NB) {
        hipLaunchKernelGGL((k_trsv_tile<T, V>), dim3(1, nrhs), dim3(256), 0, s, n, a, lda, k0, x, ldx, 1);
        if (k0 > 0) hipLaunchKernelGGL((k_gemv_bwd<T, V>), dim3((k0 + 255) / 256, nrhs), dim3(256), 0, s, n, a, lda, k0, x, ldx);
    }
}

void chol_solve(int n, const double *a, int lda, double *x, hipStream_t s) { chol_solve_t<double, double>(n, a, lda, x, 1, 0, s); }

void chol_solve_multi(int n, const double *a, int lda, double *x, int nrhs, int64_t ldx, hipStream_t s)
{
    if (nrhs > 0) chol_solve_t<double, double>(n, a, lda, x, nrhs, ldx, s);
}

// columns j0 .. j0+nc of the identity (column-major n x nc)
__global__ void k_unit_cols(int n, int j0, int nc, double *e)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)n * nc) return;
    const int c = (int)(t / n), i = (int)(t - (int64_t)c * n);
    e[t] = i == j0 + c ? 1.0 : 0.0;
}
__global__ void k_pick_diag(int n, int j0, int nc, const double *x, double *d)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < nc) d[c] = x[(int64_t)c * n + j0 + c];
}

// diagonal of A^-1 from the factor: columns of A^-1 in blocks of nbk, solved together
void chol_inv_diag(int n, const double *a, int lda, double *e, int nbk, double *d, hipStream_t s)
{
    for (int j0 = 0; j0 < n; j0 += nbk) {
        const int nc = min(nbk, n - j0);
        const int64_t tot = (int64_t)n * nc;
        hipLaunchKernelGGL(k_unit_cols, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, n, j0, nc, e);
        chol_solve_t<double, double>(n, a, lda, e, nc, n, s, j0);
        hipLaunchKernelGGL(k_pick_diag, dim3((nc + 255) / 256), dim3(256), 0, s, n, j0, nc, e, d + j0);
    }
}

// diagonal of a column-major FP32 factor
__global__ void k_diag_f32(int n, const float *a, int lda, float *d)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = a[(int64_t)i * lda + i];
}
void chol_diag32(int n, const float *a, int lda, float *d, hipStream_t s)
{
    hipLaunchKernelGGL(k_diag_f32, dim3((n + 255) / 256), dim3(256), 0, s, n, a, lda, d);
}

// FP32 factor, FP64 right-hand side accumulated in FP64 (the factor entries are widened)
void chol_solve32(int n, const float *a, int lda, double *x, float *xf, hipStream_t s)
{
    (void)xf;
    chol_solve_t<float, double>(n, a, lda, x, 1, 0, s);
}

// A += rho C^T C over the full column-major matrix (both triangles); one block per constraint
__global__ void k_add_cct(int32_t ncon, const int64_t *cptr, const int32_t *cidx, const double *cval, double rho,
                          double *a, int lda)
{
    const int j = blockIdx.x;
    const int64_t q0 = cptr[j];
    const int64_t len = cptr[j + 1] - q0;
    for (int64_t t = threadIdx.x; t < len * len; t += blockDim.x) {
        const int64_t p = t / len, q = t % len;
        const int64_t row = cidx[q0 + p] - 1, col = cidx[q0 + q] - 1;
        atomicAdd(&a[col * lda + row], rho * cval[q0 + p] * cval[q0 + q]);
    }
}
// y += rho C^T (C x) for the parameter block; one block per constraint
__global__ void k_cct_apply(int32_t ncon, const int64_t *cptr, const int32_t *cidx, const double *cval, double rho,
                            const double *x, double *y)
{
    __shared__ double red[256];
    const int j = blockIdx.x;
    double t = 0.0;
    for (int64_t q = cptr[j] + threadIdx.x; q < cptr[j + 1]; q += blockDim.x) t += cval[q] * x[cidx[q] - 1];
    red[threadIdx.x] = t;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    const double cx = red[0];
    for (int64_t q = cptr[j] + threadIdx.x; q < cptr[j + 1]; q += blockDim.x) atomicAdd(&y[cidx[q] - 1], rho * cval[q] * cx);
}
void l_add_cct(int32_t ncon, const int64_t *cptr, const int32_t *cidx, const double *cval, double rho, double *a,
               int lda, hipStream_t s)
{ if (ncon > 0) hipLaunchKernelGGL(k_add_cct, dim3(ncon), dim3(256), 0, s, ncon, cptr, cidx, cval, rho, a, lda); }
void l_cct_apply(int32_t ncon, const int64_t *cptr, const int32_t *cidx, const double *cval, double rho,
                 const double *x, double *y, hipStream_t s)
{ if (ncon > 0) hipLaunchKernelGGL(k_cct_apply, dim3(ncon), dim3(256), 0, s, ncon, cptr, cidx, cval, rho, x, y); }

void l_packed_to_full(int n, const double *packed, double *full, int lda, hipStream_t s)
{
    int64_t tot = (int64_t)n * (n + 1) / 2;
    hipLaunchKernelGGL(k_packed_to_full, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, n, packed, full, lda);
}
void l_full_to_f32(int64_t nn, const double *a, float *b, hipStream_t s)
{
    hipLaunchKernelGGL(k_full_to_f32, dim3(4096), dim3(256), 0, s, nn, a, b);
}

} // namespace mpgpu
