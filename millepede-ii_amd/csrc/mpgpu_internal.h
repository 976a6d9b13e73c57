// Host-side internals of libmpgpu (not part of the C-ABI).
#pragma once
#include "mpgpu_device.h"
#include "../../include/mpgpu.h"
#include <string>
#include <vector>
#include <cstring>
#include <utility>

namespace mpgpu {

#define HIPCHK(x)                                                                                   \
    do {                                                                                            \
        hipError_t e_ = (x);                                                                        \
        if (e_ != hipSuccess) throw MpError(30, std::string("HIP error: ") + hipGetErrorString(e_) + \
                                                    " at " + __FILE__ + ":" + std::to_string(__LINE__)); \
    } while (0)

struct MpError {
    int code;
    std::string msg;
    MpError(int c, std::string m) : code(c), msg(std::move(m)) {}
};

// Stream of the library call in progress on this thread.  Device buffers made during a call are
// stream-ordered on its context's stream (hipMallocAsync / hipFreeAsync): hipFree synchronizes
// the whole device, which would make concurrent contexts (serving workers) wait for each other's
// kernels at every freed buffer.
inline hipStream_t &alloc_stream()
{
    static thread_local hipStream_t s = nullptr;
    return s;
}
struct AllocScope {
    hipStream_t prev;
    explicit AllocScope(hipStream_t s) : prev(alloc_stream()) { alloc_stream() = s; }
    ~AllocScope() { alloc_stream() = prev; }
    AllocScope(const AllocScope &) = delete;
    AllocScope &operator=(const AllocScope &) = delete;
};

// owning device buffer; freed in the order of the stream it was made on (see alloc_stream)
template <typename T> struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    hipStream_t as = nullptr;
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    ~DBuf() { release(); }
    void release()
    {
        if (p) {
            if (as) (void)hipFreeAsync(p, as);
            else (void)hipFree(p);
        }
        p = nullptr;
        n = 0;
        as = nullptr;
    }
    void alloc(size_t count)
    {
        if (count == n && p) return;
        release();
        if (count == 0) return;
        as = alloc_stream();
        if (as) HIPCHK(hipMallocAsync((void **)&p, sizeof(T) * count, as));
        else HIPCHK(hipMalloc((void **)&p, sizeof(T) * count));
        n = count;
    }
    void upload(const T *h, size_t count, hipStream_t s)
    {
        alloc(count);
        if (count) HIPCHK(hipMemcpyAsync(p, h, sizeof(T) * count, hipMemcpyHostToDevice, s));
    }
    void upload(const std::vector<T> &v, hipStream_t s) { upload(v.data(), v.size(), s); }
    void download(T *h, size_t count, hipStream_t s) const
    {
        if (count) HIPCHK(hipMemcpyAsync(h, p, sizeof(T) * count, hipMemcpyDeviceToHost, s));
    }
    void zero(hipStream_t s)
    {
        if (n) HIPCHK(hipMemsetAsync(p, 0, sizeof(T) * n, s));
    }
    void swap(DBuf &o)
    {
        std::swap(p, o.p);
        std::swap(n, o.n);
        std::swap(as, o.as);
    }
};

// HIP events that live until the end of the enclosing scope
struct EventPool {
    std::vector<hipEvent_t> ev;
    EventPool() = default;
    EventPool(const EventPool &) = delete;
    EventPool &operator=(const EventPool &) = delete;
    ~EventPool()
    {
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    }
    hipEvent_t get()
    {
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        ev.push_back(e);
        return e;
    }
};

struct EqBlock {
    int kind; // 0 constraint, 1 measurement
    double rhs, sigma;
    std::vector<int32_t> labels;
    std::vector<double> factors;
};

// kernels (k_records.hip, k_solve.hip, k_chol.hip)
hipError_t launch_records(const DevRecords &R, const DevModel &M, const PassOpts &o, const RecOut &out, const LdsPlan &P,
                          int64_t rec_begin, int64_t rec_end, int grid, hipStream_t s, const MonCtx &mon);
int lds_bytes(LdsPlan &P);
hipError_t launch_block_gather(int64_t nblock, const int64_t *boff, const int32_t *brg, const int32_t *bcol,
                               const int32_t *gfirst, const int64_t *ptr, const int64_t *src, const double *cbuf,
                               double *mat, double *vec, hipStream_t s);
hipError_t launch_records_band(const DevRecords &R, const DevModel &M, const PassOpts &o, const RecOut &out,
                               const BandPlan &P, const int64_t *ids, int64_t nb, const int32_t *binfo,
                               double *scratch, int grid, hipStream_t s);
int band_lds_bytes(BandPlan &P);
void l_mupdat(const DevModel &M, int32_t n, const int32_t *i, const int32_t *j, const double *v, hipStream_t s);
void l_get_diag(const DevModel &M, int32_t n, double *d, hipStream_t s);
void l_get_band(const DevModel &M, int32_t n, int32_t mbandw, const int32_t *india, double *c, hipStream_t s);
void l_dot(int64_t n, const double *a, const double *b, double *part, double *out, hipStream_t s);
void l_pass_stats(int64_t n, const double *chi2, const int32_t *ndf, const int8_t *st, double *part, long long *ndfsum,
                  unsigned long long *hist, double *chi_out, hipStream_t s);
void l_gj_init(int n, const double *v, double *diag, int *used, int *info, hipStream_t s);
void l_gj_step(int n, double *v, const double *diag, int *used, int *info, double *vk, double *b, hipStream_t s);
void l_gj_finish(int n, double *v, double *b, const int *used, const int *info, hipStream_t s);
// y = A x for the packed symmetric matrix; scratch: packed_symv_scratch(n) doubles
void l_packed_symv(int n, const double *a, const double *x, double *y, double *scratch, hipStream_t s);
int64_t packed_symv_scratch(int n);
// the same product with the partial sums in a fixed file-scope device array (no allocation);
// n <= noalloc_max_n()
hipError_t noalloc_packed_symv(int device, int n, const double *a, const double *x, double *y, hipStream_t s);
int noalloc_max_n();
void l_mr_scale(int64_t n, const MrState *st, const double *y, double *v, hipStream_t s);
void l_mr_axpy(int64_t n, const MrState *st, int which, const double *r, double *y, hipStream_t s);
void l_mr_step(MrState *st, hipStream_t s);
void l_mr_update(int64_t n, const MrState *st, const double *v, double *w, double *w1, double *w2, double *x, hipStream_t s);
void l_block_symv(int32_t ngroup, const int64_t *brow, const int32_t *bcf, const int64_t *boff, const int32_t *gfirst,
                  const int64_t *ccol_ptr, const int64_t *cc_off, const int32_t *cc_fn, const double *a,
                  const double *x, double *y, int32_t nvar, int64_t nall, int32_t maxgrp, const int32_t *cc_blk,
                  double *tb, int32_t ugrp, const int32_t *bcol, int64_t nblock, int variant, hipStream_t s);
// variants of the product with identical arithmetic (for the autotuner); returns their number
int block_symv_variants(int32_t ugrp, int32_t maxgrp, int *out);
void l_cons_symv(int32_t ncon, int32_t nvar, const int64_t *cptr, const int32_t *cidx, const double *cval, const double *x,
                 double *y, hipStream_t s);
void l_scale_copy(int64_t n, double sc, const double *a, double *o, hipStream_t s);
void l_axpy_neg(int64_t n, double c, const double *r, double *y, hipStream_t s);
void l_minres_update(int64_t n, const double *v, double *w, double *w1, double *w2, double *x, double oldeps, double delta,
                     double denom, double phi, hipStream_t s);
void l_qlp_minres(int64_t n, const double *v, double *w, double *wl, double *wl2, double *x, double epln, double dlta,
                  double sc, double tau, int upd_w, int upd_x, hipStream_t s);
void l_qlp_step(int64_t n, double *v, double *w, double *wl, double *wl2, double *x, double *xl2, const QlpStep &c,
                hipStream_t s);
void l_copy(int64_t n, const double *a, double *b, hipStream_t s);
void l_xpby(int64_t n, const double *z, double beta, double *p, hipStream_t s);
void l_sub(int64_t n, const double *b, const double *ax, double *r, hipStream_t s);
void l_precons(int32_t n, int32_t p, const double *diag, double *cu, int *nrkd, const int64_t *cptr, const int32_t *cidx,
               const double *cval, double *sm, double *scratch, hipStream_t s);
void l_presols(int32_t n, int32_t p, const double *cu, const int64_t *cptr, const int32_t *cidx, const double *cval,
               const double *sm, const double *y, double *x, hipStream_t s);
void l_lltdec(int n, double *c, const int *india, int *nrkd, int iopt, int mbw, hipStream_t s);
// the band factorization by chunks of rows in parallel, with the sequential result bit for bit
// (falls back to the sequential kernel when a chunk does not catch up); L: n * mbw doubles,
// bad: n bytes, flag: three ints of device scratch
bool lltdec_parallel(int n, int mbw);
void l_lltdec_par(int n, double *c, const int *india, int *nrkd, int iopt, int mbw, double *L, int8_t *bad, int *flag,
                  hipStream_t s);
void l_band_k(int32_t ncon, int n, const double *c, const int *india, const int *cb, const int64_t *koff, double *kmat,
              int mbw, hipStream_t s);
void l_band_h(int32_t m, const int *cb, const int64_t *koff, const double *kmat, double *h, const int *hind, int *nrkd2,
              hipStream_t s);
void l_equslvs(int n, int m, const double *c, const int *india, const int *cb, const int64_t *koff, const double *kmat,
               const double *h, const int *hind, double *x, hipStream_t s);

void l_spk_build(int n, int w, const double *c, const int *india, double *Sf, double *Sb, double *scr, hipStream_t s);
void l_equslvs_spike(int n, int m, int w, const double *c, const int *india, const double *Sf, const double *Sb,
                     double *tbuf, const int *cb, const int64_t *koff, const double *kmat, const double *h, const int *hind,
                     double *dtmp, double *x, const double *Af, const double *Ab, int nlev, hipStream_t s);
int spk_chunk_rows();   // rows per chunk of the chunk-parallel band solves
int spk_scan_levels(int n, int w);
void l_spk_scan_build(int n, int w, const double *Sf, const double *Sb, double *Af, double *Ab, int nlev, hipStream_t s);

// dense LDL^T (k_chol.hip): full column-major n x n matrix (lda = n), lower triangle.  The
// factor keeps the raw pivots on the diagonal, w_i = 1/d_i (d_i > 0) or d_i in w, and needs
// ldl_work_elems(n) elements of work (panel buffer and the inverse diagonal blocks, used again
// by the solves).
int64_t ldl_work_elems(int n);
hipError_t ldl_factor(int n, double *a, int lda, double *w, double *work, hipStream_t s);
hipError_t ldl_factor32(int n, float *a, int lda, float *w, float *work, hipStream_t s);
void ldl_solve(int n, const double *a, int lda, const double *w, const double *work, double *x, int nrhs, int64_t ldx,
               hipStream_t s);
void ldl_solve32(int n, const float *a, int lda, const float *w, const float *work, double *x, int nrhs, int64_t ldx,
                 hipStream_t s);
void ldl_pivots(int n, const double *a, int lda, double *d, hipStream_t s);
void ldl_pivots32(int n, const float *a, int lda, double *d, hipStream_t s);
// diagonal of A^-1 from the FP64 factor (in the factor's index order); e: scratch of n * nbk
void ldl_inv_diag(int n, const double *a, int lda, const double *w, const double *work, double *e, int nbk, double *d,
                  hipStream_t s);
void l_reverse(int n, double *x, int nrhs, int64_t ldx, hipStream_t s);
void l_packed_to_full(int n, const double *packed, double *full, int lda, int rev, hipStream_t s);
// sc (may be NULL): symmetric scaling, element (i, j) multiplied by sc[i] sc[j]
void l_packed_to_full32(int n, const double *packed, float *full, int lda, int rev, const double *sc, hipStream_t s);
void l_scale_cols(int n, const double *d, double *x, int nrhs, int64_t ldx, hipStream_t s);
void l_add_cct(int32_t ncon, const int64_t *cptr, const int32_t *cidx, const double *cval, double rho, double *a,
               int lda, int n, int rev, hipStream_t s);
void l_add_cct32(int32_t ncon, const int64_t *cptr, const int32_t *cidx, const double *cval, double rho, float *a,
                 int lda, int n, int rev, const double *sc, hipStream_t s);
void l_cct_apply(int32_t ncon, const int64_t *cptr, const int32_t *cidx, const double *cval, double rho,
                 const double *x, double *y, hipStream_t s);
void l_col_dots(int n, int nrhs, const double *a, const double *b, double *out, hipStream_t s);
void l_col_axpy(int n, int nrhs, const double *alpha, double sg, const double *x, double *y, hipStream_t s);
void l_col_xpby(int n, int nrhs, const double *z, const double *beta, double *p, hipStream_t s);
double l_packed_fro2(int n, const double *a, double *part, double *scal, hipStream_t s);
void l_zero(int64_t n, double *y, hipStream_t s);

// factor of the dense solve methods, kept until the matrix changes (as pede's mchdec decomposes
// only after a matrix pass and otherwise reuses the decomposition, src/pede.f90:8024-8047)
struct DenseFactor {
    int n = 0;
    int metsol = 0;      // method the factor was made for
    bool rev = false;    // factor of the index-reversed matrix (chdec2's pivot order)
    bool fp32 = false;   // FP32 factor (refinement in FP64); false: FP64 factor
    double rho = 0.0, anorm = 0.0;   // weight of the augmenting C^T C term, ||A + rho C^T C||_F bound
    int nrank = 0;
    double evmax = 0.0, evmin = 0.0;
    DBuf<double> a, w, work;
    DBuf<float> a32, w32, work32;
    DBuf<double> sc;     // FP32 factor: it factors D A D with D = diag(A)^-1/2 (unit diagonal)
    std::vector<double> piv;
    // v <- A^-1 v for nrhs columns of length n (leading dimension n) with the FP64 factor
    void solve64(double *v, int nrhs, hipStream_t s)
    {
        if (rev) l_reverse(n, v, nrhs, n, s);
        ldl_solve(n, a.p, n, w.p, work.p, v, nrhs, n, s);
        if (rev) l_reverse(n, v, nrhs, n, s);
    }
    // A^-1 = D (D A D)^-1 D
    void solve32(double *v, int nrhs, hipStream_t s)
    {
        l_scale_cols(n, sc.p, v, nrhs, n, s);
        if (rev) l_reverse(n, v, nrhs, n, s);
        ldl_solve32(n, a32.p, n, w32.p, work32.p, v, nrhs, n, s);
        if (rev) l_reverse(n, v, nrhs, n, s);
        l_scale_cols(n, sc.p, v, nrhs, n, s);
    }
};

// ---- index maps and matrix structure on the device (k_struct.hip) ----
// the compiled records on the device, as the structure kernels read them
struct StructIn {
    int64_t nrec = 0, neq = 0;
    const int64_t *rec_eq = nullptr, *rec_w = nullptr, *eq_sl = nullptr;   // rec_w: word range of each record
    const uint32_t *eq_n = nullptr;
    const int32_t *lab = nullptr;       // the record words as read (global words hold labels)
    const float *vf = nullptr;
    const double *vd = nullptr;
};
// labels in use: bitmap over [lmin, lmin + 32 nw) and the rank of every 32-bit word
struct LabelMap {
    int32_t lmin = 0;
    int64_t nw = 0;
    int64_t neq_glob = 0, nglob_words = 0;   // equations with global derivatives, global words
    DBuf<uint32_t> bits;
    DBuf<int64_t> wrank;
};
// group pairs in use: bitmap over the lower triangle (a, b <= a) at bit a (a + 1) / 2 + b
struct PairMap {
    int64_t nw = 0;
    DBuf<uint32_t> bits;
    DBuf<int64_t> wrank;
};
// group-block storage: blocks in row-major order of the lower block triangle
struct BlockDev {
    int64_t nblock = 0, nstored = 0, ncc = 0;
    DBuf<int64_t> brow, boff, ccol_ptr, cc_off;
    DBuf<int32_t> brg, bcol, bcf, cc_fn, cc_blk;
};
// equation tables of the records (decodeNextMeasurement): fills in.neq, in.rec_eq, in.eq_sl,
// in.eq_n; error 20 names the first record that breaks the format
int64_t st_decode(StructIn &in, DBuf<int64_t> &rec_eq, DBuf<int64_t> &eq_sl, DBuf<uint32_t> &eq_n, hipStream_t s);
// labels -> total indices; c.labels (sorted) on the host; extra: labels of steering parameters,
// constraints and external measurements
void st_label_table(const StructIn &in, const std::vector<int32_t> &extra, LabelMap &map, std::vector<int32_t> &labels,
                    hipStream_t s);
// per record: sorted distinct total indices (gl lists), the position of every global word in its
// record's list (widx), the entry counters and the group-boundary flags (per record or per
// equation counting)
void st_record_lists(const StructIn &in, const LabelMap &map, int32_t ntot, bool per_record, DBuf<int32_t> &widx,
                     DBuf<int64_t> &rec_gl, DBuf<int32_t> &gl_tix, std::vector<int32_t> &ngl_rec,
                     std::vector<int64_t> &counters, std::vector<uint8_t> &bnd, hipStream_t s);
// per record: st[k * nrec + r], k = variables, segments, largest local index, local x global
// words, equations per variable, first group; cnt_r: values of its matrix update; sparse: the
// group pairs it couples into pm
// a context on a stream the caller keeps (serving workers reuse one stream per worker), or on
// a stream of its own when stream is null
int create_context(const mpgpu_options *opt, int32_t device, hipStream_t stream, mpgpu_ctx **out);
// FP32 values widened to FP64 (device to device)
void st_widen(const float *src, double *dst, int64_t n, hipStream_t s);
// residual monitoring: label of the first global word (0: none) and error of every equation
void st_eq_first(const StructIn &in, std::vector<int32_t> &lab, std::vector<double> &err, hipStream_t s);
// iterated entries cut: recount the global words of the equations without a low parameter
void st_entries_recount(const StructIn &in, const LabelMap &map, const std::vector<uint8_t> &low,
                        std::vector<int64_t> &counts, hipStream_t s);
void st_record_stats(const StructIn &in, const DBuf<int32_t> &widx, const DBuf<int64_t> &rec_gl,
                     const DBuf<int32_t> &gl_tix, const int32_t *var, const int32_t *group_of, int32_t ngroup,
                     int32_t max_ngl, bool sparse, PairMap &pm, std::vector<int32_t> &st, std::vector<int64_t> &cnt_r,
                     DBuf<int32_t> &d_st, hipStream_t s);
void st_blocks(PairMap &pm, int32_t ngroup, const int32_t *gfirst, const std::vector<int64_t> &extra_pairs, BlockDev &B,
               hipStream_t s);
void st_pair_blocks(const StructIn &in, const DBuf<int64_t> &rec_gl, const DBuf<int32_t> &gl_tix, const int32_t *var,
                    const int32_t *group_of, int32_t max_nseg, const DBuf<int32_t> &d_st, const PairMap &pm,
                    DBuf<int64_t> &rec_pb, DBuf<int32_t> &pair_blk, hipStream_t s);
// processing order: records stably sorted by their first parameter group
void st_order(int64_t nrec, int32_t ngroup, const DBuf<int32_t> &d_st, DBuf<int64_t> &order, std::vector<int64_t> &h_order,
              hipStream_t s);
// incidence lists of the batch order[q0, q1): ptr[nblock + 1], src[nsrc]
void st_incidence(const StructIn &in, const DBuf<int64_t> &rec_gl, const DBuf<int32_t> &gl_tix, const int32_t *var,
                  const int32_t *group_of, int32_t max_nseg, const DBuf<int64_t> &order, const DBuf<int64_t> &rec_pb,
                  const DBuf<int32_t> &pair_blk, const DBuf<int64_t> &rec_cb, int64_t q0, int64_t q1, int64_t nblock,
                  int64_t *ptr, int64_t *src, int64_t nsrc, hipStream_t s);

// ---- constraint elimination (mpgpu_qldec.cpp, k_qlelim.hip) ----
// device view of the reflectors: rows over parameter ranges, the transpose per parameter, the WY
// factor T and two m-vectors of scratch
struct QlDev {
    int32_t n = 0, m = 0;
    const int64_t *rf_ptr = nullptr, *tr_ptr = nullptr;
    const int32_t *rf_lo = nullptr, *rf_kn = nullptr, *tr_k = nullptr;
    const double *rf_val = nullptr, *rf_vk = nullptr, *tr_v = nullptr, *T = nullptr;
    double *w = nullptr, *u = nullptr;
};
struct QlElim {
    int32_t nvar = 0, ncon = 0, nfit = 0, nblk = 0;
    double evmin = 0.0, evmax = 0.0;
    std::vector<int> order;            // input constraint of each QL column (pede's sorted order)
    std::vector<int32_t> rcon, bcon;   // matConsRanges (4 per constraint), matConsBlocks (3 per block + end)
    std::vector<int64_t> ioffRow;
    std::vector<int32_t> ioffPar, irng;
    std::vector<double> matV, vecVk, matL;
    std::vector<int64_t> rf_ptr, tr_ptr;
    std::vector<int32_t> rf_lo, rf_kn, tr_k;
    std::vector<double> rf_val, rf_vk, tr_v, T;
    std::vector<double> vtvp;          // V^T V with entries of magnitude <= 16 eps cleared, unit diagonal
    DBuf<int64_t> d_rf_ptr, d_tr_ptr;
    DBuf<int32_t> d_rf_lo, d_rf_kn, d_tr_k;
    DBuf<double> d_rf_val, d_rf_vk, d_tr_v, d_T, d_w, d_u;
    QlDev dev;

    void build(int32_t nvar, const std::vector<std::vector<std::pair<int32_t, double>>> &con_tot,
               const std::vector<int32_t> &var_index);
    void upload(hipStream_t s);
    std::vector<double> bsub(const std::vector<double> &d) const;
    // coefficients of pede's band transform from X = V^T A V (column-major): Cf (row-major m x m)
    // and the diagonal weights D
    void band_coef(const std::vector<double> &X, std::vector<double> &Cf, std::vector<double> &D) const;

  private:
    void decompose();
    void wy();
};
void l_ql_apply(const QlDev &q, int trans, double *x, hipStream_t s);   // x <- Q x (0) or Q^T x (1)
void l_ql_pad(int32_t nfit, int32_t n, const double *x, double *e, hipStream_t s);
void l_ql_vtx_cols(const QlDev &q, const double *x, int64_t ldx, int32_t ncols, double *w, int64_t ldw, hipStream_t s);
void l_ql_vdense(const QlDev &q, double *Vd, hipStream_t s);
void l_ql_y(const QlDev &q, const double *Cf, double *P, hipStream_t s);
void l_ql_band(const QlDev &q, int32_t nfit, int32_t mb, const int32_t *india, const double *P, const double *S,
               double *band, hipStream_t s);

} // namespace mpgpu
