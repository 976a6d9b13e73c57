// Constraint elimination on the device: products with Q = H_m ... H_1 and Q^T in the compact WY
// form H_1 ... H_m = I - V T V^T (src/mpqldec.f90:404-454 qlmlq, applied in pede's avprod,
// src/pede.f90:5340-5358, and mminrs, 8270-8330), and pede's band of Q^T A Q for the MINRES
// preconditioner (src/mpqldec.f90:711-945 qlpssq).
//
// Reflector k is stored as a row over its parameter range [rf_lo, rf_lo + len) plus the element
// at rf_kn; tr_* holds the same values per parameter (the reflectors covering it).  All
// reductions run in a fixed order, so every product is repeatable bit for bit.
#include "mpgpu_internal.h"

namespace mpgpu {

#define QTPB 256

// w[col * ldw + k] = v_k . x_col for each reflector k (blockIdx.x) and column (blockIdx.y)
__global__ __launch_bounds__(QTPB) void k_ql_vtx(int32_t m, const int64_t *rf_ptr, const int32_t *rf_lo,
                                                  const int32_t *rf_kn, const double *rf_val, const double *rf_vk,
                                                  const double *x, int64_t ldx, double *w, int64_t ldw)
{
    __shared__ double red[QTPB];
    const int k = blockIdx.x;
    if (k >= m) return;
    const double *xc = x + (int64_t)blockIdx.y * ldx;
    const int64_t p0 = rf_ptr[k], p1 = rf_ptr[k + 1];
    const int64_t lo = rf_lo[k];
    double s = 0.0;
    for (int64_t q = p0 + threadIdx.x; q < p1; q += QTPB) s += rf_val[q] * xc[lo + (q - p0)];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = QTPB / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) w[(int64_t)blockIdx.y * ldw + k] = red[0] + rf_vk[k] * xc[rf_kn[k]];
}

// u = T^T w (trans 0) or u = T w (trans 1); T upper triangular, row-major m x m
__global__ void k_ql_tmul(int32_t m, const double *T, int trans, const double *w, double *u)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    double s = 0.0;
    if (trans)
        for (int j = i; j < m; j++) s += T[(int64_t)i * m + j] * w[j];
    else
        for (int j = 0; j <= i; j++) s += T[(int64_t)j * m + i] * w[j];
    u[i] = s;
}

// x[p] -= sum over the reflectors k covering p of v_k[p] u[k]
__global__ void k_ql_xmvu(int32_t n, const int64_t *tr_ptr, const int32_t *tr_k, const double *tr_v, const double *u,
                          double *x)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    double s = 0.0;
    for (int64_t q = tr_ptr[p]; q < tr_ptr[p + 1]; q++) s += tr_v[q] * u[tr_k[q]];
    x[p] -= s;
}

// e = [x(0 .. nfit); 0 (nfit .. n)]
__global__ void k_ql_pad(int32_t nfit, int32_t n, const double *x, double *e)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) e[p] = p < nfit ? x[p] : 0.0;
}

// V as dense columns (column-major n x m)
__global__ void k_ql_vdense(int32_t n, int32_t m, const int64_t *tr_ptr, const int32_t *tr_k, const double *tr_v,
                            double *Vd)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    for (int k = 0; k < m; k++) Vd[(int64_t)k * n + p] = 0.0;
    for (int64_t q = tr_ptr[p]; q < tr_ptr[p + 1]; q++) Vd[(int64_t)tr_k[q] * n + p] = tr_v[q];
}

// Y = 2 (P + V C^T) in place of P (column-major n x m): the products A v_k plus their corrections
// as combinations of the reflectors, C row-major m x m
__global__ void k_ql_y(int32_t n, int32_t m, const int64_t *tr_ptr, const int32_t *tr_k, const double *tr_v,
                       const double *Cf, double *P)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int64_t q0 = tr_ptr[p], q1 = tr_ptr[p + 1];
    for (int k = 0; k < m; k++) {
        double s = 0.0;
        for (int64_t q = q0; q < q1; q++) s += Cf[(int64_t)k * m + tr_k[q]] * tr_v[q];
        P[(int64_t)k * n + p] = 2.0 * (P[(int64_t)k * n + p] + s);
    }
}

// band update B_il += - sum_k V_ik P_lk - sum_k P_ik V_lk + sum_kk' V_ik S_kk' V_lk' for rows
// i < nfit and the columns of pede's variable-band layout (india: end of each row); P column-major
// n x m, S row-major m x m
__global__ void k_ql_band(int32_t nfit, int32_t mb, const int32_t *india, const int64_t *tr_ptr, const int32_t *tr_k,
                          const double *tr_v, const double *P, int64_t n, const double *S, int32_t m, double *band)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int i = (int)(t / mb), j = (int)(t - (int64_t)i * mb);
    if (i >= nfit) return;
    const int nb = min(i + 1, mb);
    if (j >= nb) return;
    const int l = i - nb + 1 + j;
    double corr = 0.0;
    const int64_t a0 = tr_ptr[i], a1 = tr_ptr[i + 1], b0 = tr_ptr[l], b1 = tr_ptr[l + 1];
    for (int64_t q = a0; q < a1; q++) corr -= tr_v[q] * P[(int64_t)tr_k[q] * n + l];
    for (int64_t q = b0; q < b1; q++) corr -= tr_v[q] * P[(int64_t)tr_k[q] * n + i];
    for (int64_t q = a0; q < a1; q++) {
        double s = 0.0;
        for (int64_t q2 = b0; q2 < b1; q2++) s += S[(int64_t)tr_k[q] * m + tr_k[q2]] * tr_v[q2];
        corr += tr_v[q] * s;
    }
    band[india[i] - nb + j] += corr;
}

static inline unsigned qblk(int64_t n) { return (unsigned)((n + QTPB - 1) / QTPB < 1 ? 1 : (n + QTPB - 1) / QTPB); }

void l_ql_apply(const QlDev &q, int trans, double *x, hipStream_t s)
{
    hipLaunchKernelGGL(k_ql_vtx, dim3(q.m), dim3(QTPB), 0, s, q.m, q.rf_ptr, q.rf_lo, q.rf_kn, q.rf_val, q.rf_vk, x,
                       (int64_t)0, q.w, (int64_t)0);
    hipLaunchKernelGGL(k_ql_tmul, dim3(qblk(q.m)), dim3(QTPB), 0, s, q.m, q.T, trans, q.w, q.u);
    hipLaunchKernelGGL(k_ql_xmvu, dim3(qblk(q.n)), dim3(QTPB), 0, s, q.n, q.tr_ptr, q.tr_k, q.tr_v, q.u, x);
}
void l_ql_pad(int32_t nfit, int32_t n, const double *x, double *e, hipStream_t s)
{
    hipLaunchKernelGGL(k_ql_pad, dim3(qblk(n)), dim3(QTPB), 0, s, nfit, n, x, e);
}
void l_ql_vtx_cols(const QlDev &q, const double *x, int64_t ldx, int32_t ncols, double *w, int64_t ldw, hipStream_t s)
{
    hipLaunchKernelGGL(k_ql_vtx, dim3(q.m, ncols), dim3(QTPB), 0, s, q.m, q.rf_ptr, q.rf_lo, q.rf_kn, q.rf_val, q.rf_vk,
                       x, ldx, w, ldw);
}
void l_ql_vdense(const QlDev &q, double *Vd, hipStream_t s)
{
    hipLaunchKernelGGL(k_ql_vdense, dim3(qblk(q.n)), dim3(QTPB), 0, s, q.n, q.m, q.tr_ptr, q.tr_k, q.tr_v, Vd);
}
void l_ql_y(const QlDev &q, const double *Cf, double *P, hipStream_t s)
{
    hipLaunchKernelGGL(k_ql_y, dim3(qblk(q.n)), dim3(QTPB), 0, s, q.n, q.m, q.tr_ptr, q.tr_k, q.tr_v, Cf, P);
}
void l_ql_band(const QlDev &q, int32_t nfit, int32_t mb, const int32_t *india, const double *P, const double *S,
               double *band, hipStream_t s)
{
    hipLaunchKernelGGL(k_ql_band, dim3(qblk((int64_t)nfit * mb)), dim3(QTPB), 0, s, nfit, mb, india, q.tr_ptr, q.tr_k,
                       q.tr_v, P, (int64_t)q.n, S, q.m, band);
}

} // namespace mpgpu
