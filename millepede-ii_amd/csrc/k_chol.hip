// Dense root-free LDL^T factorization and solves for 'method decomposition', the LAPACK
// methods and the FP32 factor that preconditions FP64 conjugate gradients.
//
// pede's chdec2 (src/mpnum.f90:891-944) eliminates from the last row upwards; pivot i with
// g_ii > 0 is replaced by w_i = 1/g_ii and counted in the rank, a non-positive pivot is kept
// as w_i = g_ii; the rows above are updated with w_i and row i is scaled by w_i.  chslv2
// (src/mpnum.f90:957-987) multiplies by w_i between the two substitutions.  The device works
// on the matrix in reversed index order (row i of pede's matrix is row n-1-i here), so the
// forward elimination below visits the pivots in chdec2's order and applies the same w_i.
//
// Blocked right-looking form, column-major full matrix (lower triangle), block width LB = 128:
//   k_ldl_diag   one workgroup factors the LB x LB diagonal block (each thread keeps 16
//                elements in registers, one barrier per column), writes the unit factor L11
//                (scaled by w), the raw pivots on the diagonal, w, and L11^-1 (kept per block:
//                it turns the panel solve and the diagonal steps of the substitutions into
//                products)
//   k_ldl_gemm   MODE_PANEL: W21 = A21 L11^-T on the matrix cores, L21 = W21 diag(w) into A,
//                W21 into a panel buffer;  MODE_SYRK: A22 -= L21 W21^T over the lower tiles.
//                128 x 128 tiles per workgroup, 4 waves of 64 x 64, v_mfma_f64_16x16x4f64 /
//                v_mfma_f32_16x16x4f32, operands straight from L2 with the next k-step's values
//                loaded while the current one multiplies.
#include "mpgpu_device.h"
#include <float.h>

namespace mpgpu {

#define LB 128

// ---------------------------------------------------------------------------------------------
// packed lower-row (pede) -> full column-major lower part, optionally index-reversed
// ---------------------------------------------------------------------------------------------
template <typename T>
__global__ void k_packed_to_full(int n, const double *p, T *a, int lda, int rev, const double *sc)
{
    int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int64_t tot = (int64_t)n * (n + 1) / 2;
    if (t >= tot) return;
    int64_t i = (int64_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while (i * (i + 1) / 2 > t) i--;
    while ((i + 1) * (i + 2) / 2 <= t) i++;
    int64_t j = t - i * (i + 1) / 2; // 0-based, j <= i
    const T v = sc ? (T)(p[t] * sc[i] * sc[j]) : (T)p[t];
    if (rev) {
        // (i, j) -> (n-1-i, n-1-j): row n-1-j >= column n-1-i
        const int64_t r = n - 1 - j, c = n - 1 - i;
        a[c * lda + r] = v;
    } else {
        a[j * lda + i] = v;
    }
}

// ---------------------------------------------------------------------------------------------
// diagonal block: factor, pivots, w, unit L11 and its inverse
// ---------------------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(1024) void k_ldl_diag(int n, T *a, int lda, int k0, T *w, T *linv)
{
    extern __shared__ double ldl_lds[];
    T *col0 = (T *)ldl_lds;                 // two column buffers of LB
    T *col1 = col0 + LB;
    T *wv = col1 + LB;                      // w of the block
    T *Lp = wv + LB;                        // unit L11, lower row-packed (diagonal = 1)
    T *Xp = Lp + LB * (LB + 1) / 2;         // L11^-1, lower row-packed
    const int nb = min(LB, n - k0);
    const int t = threadIdx.x;
    const int i = t & (LB - 1), cs = t >> 7;    // row, column class: columns cs + 8 m
    T v[16];
#pragma unroll
    for (int m = 0; m < 16; m++) {
        const int c = cs + 8 * m;
        v[m] = (i < nb && c <= i) ? a[(int64_t)(k0 + c) * lda + k0 + i] : (T)0;
    }
#pragma unroll
    for (int m = 0; m < 16; m++) {
        for (int q = 0; q < 8; q++) {
            const int j = 8 * m + q;
            if (j >= nb) break;
            T *buf = (j & 1) ? col1 : col0;
            // the owners of column j publish it (rows i >= j), after all updates of columns < j
            if (cs == q && i >= j && i < nb) buf[i] = v[m];
            __syncthreads();
            const T d = buf[j];
            const T wj = d > (T)0 ? (T)1 / d : d;
            if (t == 0) wv[j] = wj;
            const T f = (i > j && i < nb) ? buf[i] * wj : (T)0;
#pragma unroll
            for (int mm = 0; mm < 16; mm++) {
                const int c = cs + 8 * mm;
                if (c > j && c <= i) v[mm] -= f * buf[c];
            }
        }
    }
    __syncthreads();
    // L_ic = a_ic w_c below the diagonal, the raw pivot on it
#pragma unroll
    for (int m = 0; m < 16; m++) {
        const int c = cs + 8 * m;
        if (i < nb && c <= i) {
            const T l = c < i ? v[m] * wv[c] : v[m];
            a[(int64_t)(k0 + c) * lda + k0 + i] = l;
            Lp[i * (i + 1) / 2 + c] = c < i ? l : (T)1;
        }
    }
    if (t < nb) w[k0 + t] = wv[t];
    __syncthreads();
    // X = L11^-1 row by row: X_ic = -sum_{k=c}^{i-1} L_ik X_kc, 8 threads per column
    const int c = t >> 3, u = t & 7;
    if (c < nb) Xp[c * (c + 1) / 2 + c] = (T)1;
    __syncthreads();
    for (int r = 1; r < nb; r++) {
        T s = (T)0;
        if (c < r)
            for (int k = c + u; k < r; k += 8) s += Lp[r * (r + 1) / 2 + k] * Xp[k * (k + 1) / 2 + c];
        s += __shfl_xor(s, 1);
        s += __shfl_xor(s, 2);
        s += __shfl_xor(s, 4);
        if (c < r && u == 0) Xp[r * (r + 1) / 2 + c] = -s;
        __syncthreads();
    }
    // L11^-1 column-major LB x LB (zeros above the diagonal and outside the block)
    for (int e = t; e < LB * LB; e += blockDim.x) {
        const int r = e & (LB - 1), k = e >> 7;
        linv[e] = (r < nb && k <= r) ? Xp[r * (r + 1) / 2 + k] : (T)0;
    }
}

// ---------------------------------------------------------------------------------------------
// MFMA tile products
// ---------------------------------------------------------------------------------------------
typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

template <typename T> struct Mfma;
// Operands: lane l supplies A[row = l%16][k = l/16] and B[k = l/16][col = l%16].  Result
// element e of lane l is C[row][col = l%16] with row = l/16 + 4e for the FP64 instruction and
// row = 4*(l/16) + e for the FP32 one.
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

// One workgroup forms a 128 x 128 block C (+)= X Y^T, X rows = C rows, Y rows = C columns,
// both column-major (row index contiguous) with K columns (xrows, yrows >= 1 valid rows; K a
// multiple of KC).  K is consumed in chunks of KC columns staged in LDS, double-buffered.  The
// chunks travel from global memory straight into LDS (global_load_lds_dword: a wave copies one
// contiguous 256-byte run of a column per instruction, no registers involved), issued for chunk
// c+1 before the matrix-core products of chunk c.  Rows past the valid ones are read as copies
// of the last valid row: they only reach rows / columns of C that are not stored.  Four waves
// each own a 64 x 64 quarter (4 x 4 MFMA 16 x 16 accumulators); the products are formed as
// Y X^T on the matrix cores, so a lane's results lie in one column of C and 16 lanes store 16
// consecutive rows.
template <typename T> struct Chunk { static constexpr int KC = sizeof(T) == 8 ? 32 : 64; };

typedef __attribute__((address_space(3))) void *lds_ptr_t;

template <typename T>
__device__ __forceinline__ void dma_chunk(const T *X, int64_t ldx, int xrows, const T *Y, int64_t ldy, int yrows, int kb,
                                          T *b)
{
    constexpr int KC = Chunk<T>::KC;
    constexpr int RUN = 256 / (int)sizeof(T);   // elements of one 256-byte run
    constexpr int NRUN = 128 / RUN;             // runs per column of 128 rows
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int eo = (lane * 4) / (int)sizeof(T), bo = (lane * 4) % (int)sizeof(T);
#pragma unroll
    for (int j = 0; j < KC / 4; j++) {
        const int kl = wave + 4 * j;
        const int64_t k = kb + kl;
#pragma unroll
        for (int q = 0; q < NRUN; q++) {
            const int e = q * RUN + eo;
            const char *gx = (const char *)(X + k * ldx + min(e, xrows - 1)) + bo;
            const char *gy = (const char *)(Y + k * ldy + min(e, yrows - 1)) + bo;
            __builtin_amdgcn_global_load_lds((const void *)gx, (lds_ptr_t)(b + kl * 128 + q * RUN), 4, 0, 0);
            __builtin_amdgcn_global_load_lds((const void *)gy, (lds_ptr_t)(b + 128 * KC + kl * 128 + q * RUN), 4, 0, 0);
        }
    }
}

template <typename T>
__device__ __forceinline__ void chunk_mfma(const T *bx, typename Mfma<T>::acc_t (&acc)[4][4])
{
    constexpr int KC = Chunk<T>::KC;
    const T *by = bx + 128 * KC;
    const int t = threadIdx.x;
    const int wave = t >> 6, lane = t & 63, lr = lane & 15, lk = lane >> 4;
    const int wr = (wave >> 1) * 64, wc = (wave & 1) * 64;
#pragma unroll
    for (int kk = 0; kk < KC; kk += 4) {
        T xa[4], yb[4];
#pragma unroll
        for (int p = 0; p < 4; p++) {
            xa[p] = bx[(kk + lk) * 128 + wr + 16 * p + lr];
            yb[p] = by[(kk + lk) * 128 + wc + 16 * p + lr];
        }
#pragma unroll
        for (int q = 0; q < 4; q++)
#pragma unroll
            for (int p = 0; p < 4; p++) acc[q][p] = Mfma<T>::op(yb[q], xa[p], acc[q][p]);
    }
}

// the two LDS stages are distinct objects, so that the compiler can tell the products' LDS
// reads of one stage from the transfers into the other and keeps the transfers in flight
template <typename T>
__device__ __forceinline__ void block_nt(const T *X, int64_t ldx, int xrows, const T *Y, int64_t ldy, int yrows, int K,
                                         T *sa, T *sb, typename Mfma<T>::acc_t (&acc)[4][4])
{
    constexpr int KC = Chunk<T>::KC;
    const int nchunk = K / KC;     // even: K = 128
    dma_chunk<T>(X, ldx, xrows, Y, ldy, yrows, 0, sa);
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    for (int c = 0; c < nchunk; c += 2) {
        dma_chunk<T>(X, ldx, xrows, Y, ldy, yrows, (c + 1) * KC, sb);
        chunk_mfma<T>(sa, acc);
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
        if (c + 2 < nchunk) dma_chunk<T>(X, ldx, xrows, Y, ldy, yrows, (c + 2) * KC, sa);
        chunk_mfma<T>(sb, acc);
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
    }
}

enum { MODE_PANEL = 0, MODE_SYRK = 1 };

// MODE_PANEL: rows r >= k0+nb in tiles of 128 (blockIdx.x); W[r][j] = sum_k A[r][k0+k] Linv[j][k],
//             L[r][k0+j] = W[r][j] w[k0+j];  W stored column-major with leading dimension n.
// MODE_SYRK:  tiles (ti >= tj) of the trailing matrix, rows and columns counted in blocks from
//             t0 + LB * base; col_only: only the tiles of block column 0 (ti = 0..ntiles-1),
//             else the lower triangle of ntiles tiles; A[r][c] -= sum_k L[r][k0+k] W[c][k].
template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_ldl_gemm(int n, T *a, int lda, int k0, const T *linv, const T *w, T *wb,
                                                  int64_t ntiles, int base, int col_only)
{
    __shared__ T stage_a[2 * 128 * Chunk<T>::KC];
    __shared__ T stage_b[2 * 128 * Chunk<T>::KC];
    const int nb = min(LB, n - k0);
    const int t0 = k0 + nb + LB * base;
    int64_t tile = blockIdx.x;
    int64_t ti, tj;
    if (MODE == MODE_SYRK && !col_only) {
        // consecutive workgroups go to different XCDs: give each XCD a contiguous run of tiles
        // (tiles of one block row share the L21 rows in that XCD's L2)
        const int64_t per = (ntiles + 7) / 8;
        const int64_t x = tile & 7, k = tile >> 3;
        tile = x * per + k;
        if (tile >= ntiles) return;
        ti = (int64_t)((sqrt(8.0 * (double)tile + 1.0) - 1.0) * 0.5);
        while (ti * (ti + 1) / 2 > tile) ti--;
        while ((ti + 1) * (ti + 2) / 2 <= tile) ti++;
        tj = tile - ti * (ti + 1) / 2;
    } else {
        ti = tile;
        tj = 0;
    }
    const int r0 = t0 + (int)ti * LB;
    const int c0 = MODE == MODE_SYRK ? t0 + (int)tj * LB : 0;
    const int xrows = n - r0;
    const int yrows = MODE == MODE_SYRK ? n - c0 : nb;
    typedef typename Mfma<T>::acc_t acc_t;
    acc_t acc[4][4];
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
        for (int p = 0; p < 4; p++)
#pragma unroll
            for (int e = 0; e < 4; e++) acc[q][p][e] = (T)0;
    if (MODE == MODE_PANEL)
        block_nt<T>(a + (int64_t)k0 * lda + r0, lda, xrows, linv, LB, yrows, LB, stage_a, stage_b, acc);
    else
        block_nt<T>(a + (int64_t)k0 * lda + r0, lda, xrows, wb + c0, n, yrows, LB, stage_a, stage_b, acc);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4;
    const int rbase = r0 + (wave >> 1) * 64, cbase = c0 + (wave & 1) * 64;
    // The stores go out in four groups of 16 values per lane.  A group's loads (matrix elements,
    // pivots) come from clamped valid addresses and are issued together, one group ahead of the
    // group being stored, so the memory round trips overlap.
    T cur[2][4][4];
    auto fetch = [&](int q, T (&c)[4][4]) {
#pragma unroll
        for (int p = 0; p < 4; p++)
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int r = min(rbase + 16 * p + lr, n - 1);
                const int cc = cbase + 16 * q + Mfma<T>::row(lk, e);
                if (MODE == MODE_PANEL) c[p][e] = w[k0 + min(cc, nb - 1)];
                else c[p][e] = a[(int64_t)min(cc, n - 1) * lda + r];
            }
    };
    fetch(0, cur[0]);
#pragma unroll
    for (int q = 0; q < 4; q++) {
        if (q + 1 < 4) fetch(q + 1, cur[(q + 1) & 1]);
#pragma unroll
        for (int p = 0; p < 4; p++)
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int r = rbase + 16 * p + lr;
                const int cc = cbase + 16 * q + Mfma<T>::row(lk, e);
                const T cv = cur[q & 1][p][e];
                if (MODE == MODE_PANEL) {
                    if (r < n && cc < nb) {
                        wb[(int64_t)cc * n + r] = acc[q][p][e];
                        a[(int64_t)(k0 + cc) * lda + r] = acc[q][p][e] * cv;
                    }
                } else {
                    if (r < n && cc < n && cc <= r) a[(int64_t)cc * lda + r] = cv - acc[q][p][e];
                }
            }
    }
}

// ---------------------------------------------------------------------------------------------
// substitutions with the unit factor, the w scaling between them (chslv2), several right-hand
// sides side by side (grid dimension y, column q at x + q * ldx)
// ---------------------------------------------------------------------------------------------
// The substitution kernels use 512 threads: 128 rows (or columns) x 4 parts of 32 terms, each
// part's 32 loads unrolled so that they are all in flight together; the 4 partial sums meet in
// LDS.  Indices past the end of a partial block are clamped to a valid element and multiplied
// by a zero entry of the staged vector.
#define SUBT 512

// y_r = sum_k M[r][k] v_k over k < LB for a column-major LB x LB block M (ld LB when the block
// is an inverse diagonal block, lda for a matrix block), thread (r, part)
template <typename T, typename V>
__device__ __forceinline__ V part_rows(const T *M, int64_t ld, int r, int part, int kmax, const V *v)
{
    V s0 = 0, s1 = 0;
#pragma unroll
    for (int q = 0; q < 32; q += 2) {
        const int k = part * 32 + q;
        s0 += (V)M[(int64_t)min(k, kmax) * ld + r] * v[k];
        s1 += (V)M[(int64_t)min(k + 1, kmax) * ld + r] * v[k + 1];
    }
    return s0 + s1;
}
// y_c = sum_r M[r][c] v_r over r < LB (transposed product), thread (c, part): M column c is
// contiguous
template <typename T, typename V>
__device__ __forceinline__ V part_cols(const T *M, int64_t ld, int c, int part, int rmax, const V *v)
{
    V s0 = 0, s1 = 0;
#pragma unroll
    for (int q = 0; q < 32; q += 2) {
        const int r = part * 32 + q;
        s0 += (V)M[(int64_t)c * ld + min(r, rmax)] * v[r];
        s1 += (V)M[(int64_t)c * ld + min(r + 1, rmax)] * v[r + 1];
    }
    return s0 + s1;
}

// forward, first tile of the sweep: x_t = L_tt^-1 x_t
template <typename T, typename V>
__global__ __launch_bounds__(SUBT) void k_fwd_diag(int n, const T *linv_all, int t, V *x, int64_t ldx)
{
    x += blockIdx.y * ldx;
    __shared__ V xs[LB];
    __shared__ V part[4][LB];
    const int k0 = t * LB, nb = min(LB, n - k0);
    const T *li = linv_all + (int64_t)t * LB * LB;
    if (threadIdx.x < LB) xs[threadIdx.x] = threadIdx.x < nb ? x[k0 + threadIdx.x] : (V)0;
    __syncthreads();
    const int r = threadIdx.x & (LB - 1), h = threadIdx.x >> 7;
    part[h][r] = part_rows<T, V>(li, LB, r, h, LB - 1, xs);
    __syncthreads();
    if (threadIdx.x < nb) x[k0 + threadIdx.x] = part[0][threadIdx.x] + part[1][threadIdx.x] + part[2][threadIdx.x] + part[3][threadIdx.x];
}

// forward update from tile t: rows of the 128-row chunk blockIdx.x (chunks after tile t):
// x_r -= sum_j L[r][tile t col j] x_j; the chunk right after tile t then applies its own L^-1
template <typename T, typename V>
__global__ __launch_bounds__(SUBT) void k_fwd_update(int n, const T *a, int lda, const T *linv_all, int t, V *x,
                                                     int64_t ldx)
{
    x += blockIdx.y * ldx;
    __shared__ V xt[LB];
    __shared__ V part[4][LB];
    __shared__ V xs[LB];
    const int k0 = t * LB, nb = min(LB, n - k0);
    const int ck = t + 1 + blockIdx.x;      // chunk (tile index) updated by this workgroup
    const int c0 = ck * LB, cn = min(LB, n - c0);
    if (threadIdx.x < LB) xt[threadIdx.x] = threadIdx.x < nb ? x[k0 + threadIdx.x] : (V)0;
    __syncthreads();
    const int r = threadIdx.x & (LB - 1), h = threadIdx.x >> 7;
    // rows past the end of the matrix read row cn-1 and are not stored
    part[h][r] = part_rows<T, V>(a + (int64_t)k0 * lda + c0, lda, min(r, cn - 1), h, nb - 1, xt);
    __syncthreads();
    const V upd = threadIdx.x < LB ? part[0][threadIdx.x] + part[1][threadIdx.x] + part[2][threadIdx.x] + part[3][threadIdx.x] : (V)0;
    if (blockIdx.x != 0) {
        if (threadIdx.x < cn) x[c0 + threadIdx.x] -= upd;
        return;
    }
    // the next tile is complete: solve it with its inverse block
    if (threadIdx.x < LB) xs[threadIdx.x] = threadIdx.x < cn ? x[c0 + threadIdx.x] - upd : (V)0;
    __syncthreads();
    const T *li = linv_all + (int64_t)ck * LB * LB;
    part[h][r] = part_rows<T, V>(li, LB, r, h, LB - 1, xs);
    __syncthreads();
    if (threadIdx.x < cn) x[c0 + threadIdx.x] = part[0][threadIdx.x] + part[1][threadIdx.x] + part[2][threadIdx.x] + part[3][threadIdx.x];
}

// x <- w o x  (chslv2's multiplication by the stored pivot, src/mpnum.f90:979)
template <typename T, typename V>
__global__ void k_scale_w(int n, const T *w, V *x, int64_t ldx)
{
    x += blockIdx.y * ldx;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] *= (V)w[i];
}

// backward, last tile of the sweep: x_t = L_tt^-T x_t
template <typename T, typename V>
__global__ __launch_bounds__(SUBT) void k_bwd_diag(int n, const T *linv_all, int t, V *x, int64_t ldx)
{
    x += blockIdx.y * ldx;
    __shared__ V xs[LB];
    __shared__ V part[4][LB];
    const int k0 = t * LB, nb = min(LB, n - k0);
    const T *li = linv_all + (int64_t)t * LB * LB;
    if (threadIdx.x < LB) xs[threadIdx.x] = threadIdx.x < nb ? x[k0 + threadIdx.x] : (V)0;
    __syncthreads();
    const int c = threadIdx.x & (LB - 1), h = threadIdx.x >> 7;
    part[h][c] = part_cols<T, V>(li, LB, c, h, LB - 1, xs);
    __syncthreads();
    if (threadIdx.x < nb) x[k0 + threadIdx.x] = part[0][threadIdx.x] + part[1][threadIdx.x] + part[2][threadIdx.x] + part[3][threadIdx.x];
}

// backward update from tile t (solved): chunk blockIdx.x of the columns before tile t:
// x_j -= sum_{i in tile t} L[i][j] x_i; the chunk right before tile t then applies L^-T
template <typename T, typename V>
__global__ __launch_bounds__(SUBT) void k_bwd_update(int n, const T *a, int lda, const T *linv_all, int t, V *x,
                                                     int64_t ldx)
{
    x += blockIdx.y * ldx;
    __shared__ V xt[LB];
    __shared__ V part[4][LB];
    __shared__ V xs[LB];
    const int k0 = t * LB, nb = min(LB, n - k0);
    const int ck = t - 1 - blockIdx.x;
    const int c0 = ck * LB;               // full tiles only (ck < t)
    if (threadIdx.x < LB) xt[threadIdx.x] = threadIdx.x < nb ? x[k0 + threadIdx.x] : (V)0;
    __syncthreads();
    const int c = threadIdx.x & (LB - 1), h = threadIdx.x >> 7;
    // column c0 + c of L, rows k0 .. k0 + nb - 1
    part[h][c] = part_cols<T, V>(a + (int64_t)c0 * lda + k0, lda, c, h, nb - 1, xt);
    __syncthreads();
    const V upd = threadIdx.x < LB ? part[0][threadIdx.x] + part[1][threadIdx.x] + part[2][threadIdx.x] + part[3][threadIdx.x] : (V)0;
    if (blockIdx.x != 0) {
        if (threadIdx.x < LB) x[c0 + threadIdx.x] -= upd;
        return;
    }
    if (threadIdx.x < LB) xs[threadIdx.x] = x[c0 + threadIdx.x] - upd;
    __syncthreads();
    const T *li = linv_all + (int64_t)ck * LB * LB;
    part[h][c] = part_cols<T, V>(li, LB, c, h, LB - 1, xs);
    __syncthreads();
    if (threadIdx.x < LB) x[c0 + threadIdx.x] = part[0][threadIdx.x] + part[1][threadIdx.x] + part[2][threadIdx.x] + part[3][threadIdx.x];
}

// x <- reversed x (index i <-> n-1-i), several right-hand sides
template <typename V>
__global__ void k_reverse(int n, V *x, int64_t ldx)
{
    x += blockIdx.y * ldx;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n / 2) { V tmp = x[i]; x[i] = x[n - 1 - i]; x[n - 1 - i] = tmp; }
}

template <typename T>
__global__ void k_pivots(int n, const T *a, int lda, double *d)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = (double)a[(int64_t)i * lda + i];
}

// ---------------------------------------------------------------------------------------------
// host wrappers
// ---------------------------------------------------------------------------------------------
static int nblocks_of(int n) { return (n + LB - 1) / LB; }

int64_t ldl_work_elems(int n) { return 2 * (int64_t)n * LB + (int64_t)nblocks_of(n) * LB * LB; }

template <typename T> static int diag_lds() { return (int)sizeof(T) * (3 * LB + LB * (LB + 1)); }


// Right-looking factorization with one block of look-ahead: at step t the trailing update first
// brings block column t+1 up to date (on the caller's stream), then a second stream factors
// diagonal block t+1 and forms its panel while the caller's stream updates the rest of the
// trailing matrix.  Two panel buffers alternate, so the panel of t+1 never overwrites the W of
// step t that the concurrent update still reads.
template <typename T> static hipError_t ldl_factor_t(int n, T *a, int lda, T *w, T *work, hipStream_t s)
{
    T *wbuf[2] = {work, work + (int64_t)n * LB};
    T *linv_all = work + 2 * (int64_t)n * LB;
    const int ldsd = diag_lds<T>(), ldsg = 0;   // the products' LDS stages are static
    hipError_t e = hipFuncSetAttribute((const void *)k_ldl_diag<T>, hipFuncAttributeMaxDynamicSharedMemorySize, ldsd);
    if (e != hipSuccess) return e;
    const int nblk = nblocks_of(n);
    auto diag_panel = [&](int t, hipStream_t st) {
        const int k0 = t * LB;
        T *linv = linv_all + (int64_t)t * LB * LB;
        hipLaunchKernelGGL(k_ldl_diag<T>, dim3(1), dim3(1024), ldsd, st, n, a, lda, k0, w, linv);
        const int rest = n - k0 - min(LB, n - k0);
        if (rest > 0) {
            const int64_t nt = (rest + LB - 1) / LB;
            hipLaunchKernelGGL((k_ldl_gemm<T, MODE_PANEL>), dim3((unsigned)nt), dim3(256), ldsg, st, n, a, lda, k0, linv, w,
                               wbuf[t & 1], nt, 0, 0);
        }
    };
    diag_panel(0, s);
    if (nblk == 1) return hipGetLastError();
    hipStream_t s2 = nullptr;
    hipEvent_t e1 = nullptr, e2 = nullptr;
    if ((e = hipStreamCreateWithFlags(&s2, hipStreamNonBlocking)) != hipSuccess) return e;
    if ((e = hipEventCreateWithFlags(&e1, hipEventDisableTiming)) != hipSuccess) { (void)hipStreamDestroy(s2); return e; }
    if ((e = hipEventCreateWithFlags(&e2, hipEventDisableTiming)) != hipSuccess) {
        (void)hipEventDestroy(e1); (void)hipStreamDestroy(s2); return e;
    }
    for (int t = 0; t + 1 < nblk; t++) {
        const int k0 = t * LB;
        const int rest = n - k0 - LB;           // rows after block t (block t is full here)
        const int64_t nt = (rest + LB - 1) / LB;
        T *wb = wbuf[t & 1];
        const T *linv = linv_all + (int64_t)t * LB * LB;
        // block column t+1 of the trailing matrix
        hipLaunchKernelGGL((k_ldl_gemm<T, MODE_SYRK>), dim3((unsigned)nt), dim3(256), ldsg, s, n, a, lda, k0, linv, w, wb,
                           nt, 0, 1);
        (void)hipEventRecord(e1, s);
        (void)hipStreamWaitEvent(s2, e1, 0);
        diag_panel(t + 1, s2);
        (void)hipEventRecord(e2, s2);
        // the rest of the trailing matrix, concurrently with block t+1
        if (nt > 1) {
            const int64_t ntiles = (nt - 1) * nt / 2;
            const int64_t grid = ((ntiles + 7) / 8) * 8;
            hipLaunchKernelGGL((k_ldl_gemm<T, MODE_SYRK>), dim3((unsigned)grid), dim3(256), ldsg, s, n, a, lda, k0, linv, w,
                               wb, ntiles, 1, 0);
        }
        (void)hipStreamWaitEvent(s, e2, 0);
    }
    e = hipGetLastError();
    (void)hipEventDestroy(e1);
    (void)hipEventDestroy(e2);
    (void)hipStreamDestroy(s2);
    return e;
}

hipError_t ldl_factor(int n, double *a, int lda, double *w, double *work, hipStream_t s)
{ return ldl_factor_t<double>(n, a, lda, w, work, s); }
hipError_t ldl_factor32(int n, float *a, int lda, float *w, float *work, hipStream_t s)
{ return ldl_factor_t<float>(n, a, lda, w, work, s); }

template <typename T, typename V>
static void ldl_solve_t(int n, const T *a, int lda, const T *w, const T *work, V *x, int nrhs, int64_t ldx, int first,
                        int fwd_only, hipStream_t s)
{
    const T *linv_all = work + 2 * (int64_t)n * LB;
    const int nt = nblocks_of(n);
    const int tf = first / LB;       // right-hand sides are zero above row 'first'
    hipLaunchKernelGGL((k_fwd_diag<T, V>), dim3(1, nrhs), dim3(SUBT), 0, s, n, linv_all, tf, x, ldx);
    for (int t = tf; t < nt - 1; t++)
        hipLaunchKernelGGL((k_fwd_update<T, V>), dim3(nt - 1 - t, nrhs), dim3(SUBT), 0, s, n, a, lda, linv_all, t, x, ldx);
    if (fwd_only) return;
    hipLaunchKernelGGL((k_scale_w<T, V>), dim3((n + 255) / 256, nrhs), dim3(256), 0, s, n, w, x, ldx);
    hipLaunchKernelGGL((k_bwd_diag<T, V>), dim3(1, nrhs), dim3(SUBT), 0, s, n, linv_all, nt - 1, x, ldx);
    for (int t = nt - 1; t > 0; t--)
        hipLaunchKernelGGL((k_bwd_update<T, V>), dim3(t, nrhs), dim3(SUBT), 0, s, n, a, lda, linv_all, t, x, ldx);
}

void ldl_solve(int n, const double *a, int lda, const double *w, const double *work, double *x, int nrhs, int64_t ldx,
               hipStream_t s)
{ if (nrhs > 0) ldl_solve_t<double, double>(n, a, lda, w, work, x, nrhs, ldx, 0, 0, s); }
void ldl_solve32(int n, const float *a, int lda, const float *w, const float *work, double *x, int nrhs, int64_t ldx,
                 hipStream_t s)
{ if (nrhs > 0) ldl_solve_t<float, double>(n, a, lda, w, work, x, nrhs, ldx, 0, 0, s); }

void l_reverse(int n, double *x, int nrhs, int64_t ldx, hipStream_t s)
{
    if (n > 1 && nrhs > 0)
        hipLaunchKernelGGL(k_reverse<double>, dim3((n / 2 + 255) / 256, nrhs), dim3(256), 0, s, n, x, ldx);
}

template <typename T> static void pivots_t(int n, const T *a, int lda, double *d, hipStream_t s)
{ hipLaunchKernelGGL(k_pivots<T>, dim3((n + 255) / 256), dim3(256), 0, s, n, a, lda, d); }
void ldl_pivots(int n, const double *a, int lda, double *d, hipStream_t s) { pivots_t<double>(n, a, lda, d, s); }
void ldl_pivots32(int n, const float *a, int lda, double *d, hipStream_t s) { pivots_t<float>(n, a, lda, d, s); }

// columns j0 .. j0+nc of the identity (column-major n x nc)
__global__ void k_unit_cols(int n, int j0, int nc, double *e)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)n * nc) return;
    const int c = (int)(t / n), i = (int)(t - (int64_t)c * n);
    e[t] = i == j0 + c ? 1.0 : 0.0;
}
// d_c = sum_k w_k y_kc^2 for y = L^-1 e_c (one workgroup per column)
__global__ __launch_bounds__(256) void k_wsq(int n, int nc, const double *w, const double *y, double *d)
{
    __shared__ double red[256];
    const double *yc = y + (int64_t)blockIdx.x * n;
    double s = 0.0;
    for (int k = threadIdx.x; k < n; k += blockDim.x) s += w[k] * yc[k] * yc[k];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) d[blockIdx.x] = red[0];
}

// diagonal of A^-1 = L^-T W L^-1 (reversed order, as the factor): (A^-1)_cc = sum_k w_k (L^-1 e_c)_k^2,
// columns in blocks of nbk solved together by the forward sweep only
void ldl_inv_diag(int n, const double *a, int lda, const double *w, const double *work, double *e, int nbk, double *d,
                  hipStream_t s)
{
    for (int j0 = 0; j0 < n; j0 += nbk) {
        const int nc = min(nbk, n - j0);
        const int64_t tot = (int64_t)n * nc;
        hipLaunchKernelGGL(k_unit_cols, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, n, j0, nc, e);
        ldl_solve_t<double, double>(n, a, lda, w, work, e, nc, n, j0, 1, s);
        hipLaunchKernelGGL(k_wsq, dim3(nc), dim3(256), 0, s, n, nc, w, e, d + j0);
    }
}

// A += rho C^T C over the lower triangle of the (optionally reversed) column-major matrix;
// one block per constraint
template <typename T>
__global__ void k_add_cct(int32_t ncon, const int64_t *cptr, const int32_t *cidx, const double *cval, double rho,
                          T *a, int lda, int n, int rev, const double *sc)
{
    const int j = blockIdx.x;
    const int64_t q0 = cptr[j];
    const int64_t len = cptr[j + 1] - q0;
    for (int64_t t = threadIdx.x; t < len * len; t += blockDim.x) {
        const int64_t p = t / len, q = t % len;
        int64_t row = cidx[q0 + p] - 1, col = cidx[q0 + q] - 1;
        const double f = sc ? sc[row] * sc[col] : 1.0;
        if (rev) { row = n - 1 - row; col = n - 1 - col; }
        if (row < col) continue;
        atomicAdd(&a[col * lda + row], (T)(rho * cval[q0 + p] * cval[q0 + q] * f));
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
               int lda, int n, int rev, hipStream_t s)
{ if (ncon > 0) hipLaunchKernelGGL(k_add_cct<double>, dim3(ncon), dim3(256), 0, s, ncon, cptr, cidx, cval, rho, a, lda, n, rev, (const double *)nullptr); }
void l_add_cct32(int32_t ncon, const int64_t *cptr, const int32_t *cidx, const double *cval, double rho, float *a,
                 int lda, int n, int rev, const double *sc, hipStream_t s)
{ if (ncon > 0) hipLaunchKernelGGL(k_add_cct<float>, dim3(ncon), dim3(256), 0, s, ncon, cptr, cidx, cval, rho, a, lda, n, rev, sc); }
void l_cct_apply(int32_t ncon, const int64_t *cptr, const int32_t *cidx, const double *cval, double rho,
                 const double *x, double *y, hipStream_t s)
{ if (ncon > 0) hipLaunchKernelGGL(k_cct_apply, dim3(ncon), dim3(256), 0, s, ncon, cptr, cidx, cval, rho, x, y); }

void l_packed_to_full(int n, const double *packed, double *full, int lda, int rev, hipStream_t s)
{
    int64_t tot = (int64_t)n * (n + 1) / 2;
    hipLaunchKernelGGL(k_packed_to_full<double>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, n, packed, full, lda, rev,
                       (const double *)nullptr);
}
void l_packed_to_full32(int n, const double *packed, float *full, int lda, int rev, const double *sc, hipStream_t s)
{
    int64_t tot = (int64_t)n * (n + 1) / 2;
    hipLaunchKernelGGL(k_packed_to_full<float>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, n, packed, full, lda, rev, sc);
}
// x[:, q] <- d o x[:, q]
__global__ void k_scale_cols(int n, const double *d, double *x, int64_t ldx)
{
    x += blockIdx.y * ldx;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] *= d[i];
}
void l_scale_cols(int n, const double *d, double *x, int nrhs, int64_t ldx, hipStream_t s)
{
    if (n > 0 && nrhs > 0) hipLaunchKernelGGL(k_scale_cols, dim3((n + 255) / 256, nrhs), dim3(256), 0, s, n, d, x, ldx);
}

// ---------------------------------------------------------------------------------------------
// column-wise vector operations of the block conjugate gradients (nrhs columns of length n)
// ---------------------------------------------------------------------------------------------
// out[q] = sum_i a[i,q] b[i,q]; one workgroup per column, fixed-order tree
__global__ __launch_bounds__(256) void k_col_dots(int n, const double *a, const double *b, double *out)
{
    __shared__ double red[256];
    const int64_t o = (int64_t)blockIdx.x * n;
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += a[o + i] * b[o + i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}
// y[:,q] += sg * alpha[q] * x[:,q]
__global__ void k_col_axpy(int n, const double *alpha, double sg, const double *x, double *y)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t o = (int64_t)blockIdx.y * n;
    if (i < n) y[o + i] += sg * alpha[blockIdx.y] * x[o + i];
}
// p[:,q] = z[:,q] + beta[q] * p[:,q]
__global__ void k_col_xpby(int n, const double *z, const double *beta, double *p)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t o = (int64_t)blockIdx.y * n;
    if (i < n) p[o + i] = z[o + i] + beta[blockIdx.y] * p[o + i];
}
// squared Frobenius norm of the packed symmetric matrix (off-diagonal elements twice):
// per-workgroup partial sums, then one workgroup adds them in order
__global__ __launch_bounds__(256) void k_packed_fro_part(int n, const double *a, double *part)
{
    __shared__ double red[256];
    const int64_t tot = (int64_t)n * (n + 1) / 2;
    double s = 0.0;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < tot; t += (int64_t)gridDim.x * blockDim.x) {
        int64_t i = (int64_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
        while (i * (i + 1) / 2 > t) i--;
        while ((i + 1) * (i + 2) / 2 <= t) i++;
        const bool diag = t - i * (i + 1) / 2 == i;
        s += (diag ? 1.0 : 2.0) * a[t] * a[t];
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}
__global__ __launch_bounds__(256) void k_sum_parts(int np, const double *part, double *out)
{
    __shared__ double red[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < np; i += blockDim.x) s += part[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

void l_col_dots(int n, int nrhs, const double *a, const double *b, double *out, hipStream_t s)
{ hipLaunchKernelGGL(k_col_dots, dim3(nrhs), dim3(256), 0, s, n, a, b, out); }
void l_col_axpy(int n, int nrhs, const double *alpha, double sg, const double *x, double *y, hipStream_t s)
{ hipLaunchKernelGGL(k_col_axpy, dim3((n + 255) / 256, nrhs), dim3(256), 0, s, n, alpha, sg, x, y); }
void l_col_xpby(int n, int nrhs, const double *z, const double *beta, double *p, hipStream_t s)
{ hipLaunchKernelGGL(k_col_xpby, dim3((n + 255) / 256, nrhs), dim3(256), 0, s, n, z, beta, p); }
// part: at least 1024 doubles of scratch; returns ||A||_F^2 (synchronises the stream)
double l_packed_fro2(int n, const double *a, double *part, double *scal, hipStream_t s)
{
    const int g = 1024;
    hipLaunchKernelGGL(k_packed_fro_part, dim3(g), dim3(256), 0, s, n, a, part);
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(256), 0, s, g, part, scal);
    double r = 0.0;
    (void)hipMemcpyAsync(&r, scal, sizeof(double), hipMemcpyDeviceToHost, s);
    (void)hipStreamSynchronize(s);
    return r;
}
__global__ void k_zero64(int64_t n, double *y)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) y[i] = 0.0;
}
void l_zero(int64_t n, double *y, hipStream_t s)
{
    int64_t g = (n + 255) / 256;
    if (g > 65536) g = 65536;
    if (n > 0) hipLaunchKernelGGL(k_zero64, dim3((unsigned)g), dim3(256), 0, s, n, y);
}

} // namespace mpgpu
