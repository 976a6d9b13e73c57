// libmpgpu host side: context, record compilation, structure build (loop1/loop2
// equivalents), data passes (loopn) and the global solve dispatch (xloopn's metsol switch).
#include "mpgpu_internal.h"
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <functional>
#include <limits>
#include <numeric>
#include <omp.h>
#include <zlib.h>
#include <sys/stat.h>

using namespace mpgpu;

namespace {
struct Precond;

// Two pinned host buffers through which records travel to the device: the reader fills one
// while the copy of the other runs.  Each holds CHUNK words and their values (FP64-sized).
struct Staging {
    static constexpr int64_t CHUNK = (int64_t)4 << 20;
    char *pin[2] = {nullptr, nullptr};
    hipEvent_t done[2] = {nullptr, nullptr};
    bool inflight[2] = {false, false};
    int k = 0;
    int64_t n = 0;     // words in buffer k
    void init()
    {
        for (int i = 0; i < 2; i++) {
            HIPCHK(hipHostMalloc((void **)&pin[i], (size_t)CHUNK * 12, hipHostMallocDefault));
            HIPCHK(hipEventCreateWithFlags(&done[i], hipEventDisableTiming));
        }
    }
    ~Staging()
    {
        for (int i = 0; i < 2; i++) {
            if (done[i]) {
                if (inflight[i]) (void)hipEventSynchronize(done[i]);
                (void)hipEventDestroy(done[i]);
            }
            if (pin[i]) (void)hipHostFree(pin[i]);
        }
    }
    int32_t *words() { return (int32_t *)pin[k]; }
    char *vals() { return pin[k] + CHUNK * 4; }
};
}

struct mpgpu_ctx {
    mpgpu_options opt{};
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = true;       // false: a stream the caller keeps (serving workers)
    std::string err;

    // ---- input records: record pointers on the host; words and values go to the device as
    // they are added (d_in, and d_vf or d_vd below), through pinned staging buffers for files
    // and large inputs ----
    std::vector<int64_t> rec_ptr{0};
    DBuf<int32_t> d_in;        // words as read (global words hold labels); capacity d_in.n
    int64_t nwords = 0;        // words on the device or in flight (staged ones not included)
    bool dbl = false;          // values held as FP64
    std::unique_ptr<Staging> stage;

    // ---- steering ----
    std::vector<int32_t> sp_label;
    std::vector<double> sp_value, sp_presig;
    std::vector<EqBlock> cons, meas;

    // ---- structure ----
    bool built = false;
    int32_t ntot = 0, nvar = 0, ncon = 0, nall = 0, ngroup = 0;
    std::vector<int32_t> labels, var_index, group_start_label, var_to_tot;
    std::vector<int64_t> counters;
    std::vector<double> par, presig_tot, pre_weight;
    std::vector<int32_t> gfirst, group_of; // variable groups
    std::vector<int64_t> brow, boff;
    std::vector<int32_t> bcol, brg;
    int32_t uniform_grp = 0;              // size of every parameter group when they are all equal, else 0
    int spmv_variant = -1;                // block product variant chosen by the autotuner (-1: default)
    std::string tuning;                   // the autotuner's measurements
    mpgpu_record_summary summary{};       // loop1 / loop2 counts for millepede.log
    bool blocks_on_host = true;           // brow/bcol/brg/boff copied from the device (on first use)
    int64_t nblock = 0, nstored = 0;
    // constraints (CSR over variable indices) and residual bookkeeping
    std::vector<int64_t> cptr;
    std::vector<int32_t> cidx;
    std::vector<double> cval;
    std::vector<double> crhs;
    std::vector<std::vector<std::pair<int32_t, double>>> con_tot; // (total index, factor) per constraint
    // constraint elimination (the MINRES methods without 'withmultipliers'): QL decomposition of
    // the constraint matrix; the solve works on the first nfit components of Q^T x
    bool elim = false;
    int32_t nfit = 0;
    std::unique_ptr<QlElim> ql;
    // residual monitoring: measurement groups (total index of the first global parameter), their
    // mean errors and the histograms of the last pass
    std::vector<int32_t> mon_tot;
    std::vector<double> mon_res;
    int32_t mon_bins = 0;
    DBuf<int32_t> d_mon_grp;
    DBuf<double> d_mon_res;
    DBuf<unsigned int> d_mon_hist;

    // ---- device ----
    DBuf<int64_t> d_rec_eq, d_rec_gl, d_rec_pb, d_eq_sl, d_order, d_rec_w;
    DBuf<uint32_t> d_eq_n;
    DBuf<float> d_vf;
    DBuf<double> d_vd;
    DBuf<int32_t> d_widx, d_gl_tix, d_pair_blk;
    DBuf<double> d_par, d_mat, d_vec, d_pre_weight;
    DBuf<int32_t> d_var, d_group_of, d_gfirst, d_bcol, d_brg;
    DBuf<int32_t> d_grp_tot;        // total index -> group of the variable (-1: fixed)
    DBuf<int64_t> d_brow, d_boff, d_ccol_ptr, d_cc_off;
    DBuf<int32_t> d_bcf, d_cc_fn;   // per block / per column-list entry: first index and size of the other group
    DBuf<int32_t> d_cc_blk;         // per column-list entry: the block
    DBuf<double> d_tb;              // per block: its transposed product (max_grp values, groups <= 8)
    DBuf<double> d_psymv;           // full storage: partial sums of the packed product (deterministic)
    DBuf<unsigned long long> d_cnt;
    DBuf<double> d_chi2;
    DBuf<int32_t> d_ndf;
    DBuf<int8_t> d_status;
    DBuf<double> d_part, d_scal;
    DBuf<long long> d_ndfsum;
    DBuf<unsigned long long> d_hist;
    DBuf<int64_t> d_cptr;
    DBuf<int32_t> d_cidx;
    DBuf<double> d_cval;
    LdsPlan plan{};
    // bordered-band local fits
    DBuf<int32_t> d_band_info;
    DBuf<int64_t> d_band_ids;
    DBuf<double> d_band_scratch;
    BandPlan bplan{};
    int64_t nband = 0;
    // two-phase matrix build (group-block storage): record updates -> buffer -> block gather
    struct Batch { int64_t q0, q1, base, len, inc0; };
    std::vector<Batch> batches;
    DBuf<int64_t> d_rec_cb;      // per record: offset of its update in the batch buffer (-1: none)
    DBuf<int32_t> d_rec_cn;      // per record: number of update values
    DBuf<int64_t> d_inc_ptr;     // per batch: (nblock+1) pointers into d_inc_src
    DBuf<int64_t> d_inc_src;     // buffer offsets of the records' values for each block
    DBuf<double> d_cbuf;         // update buffer (largest batch)
    int band_grid = 0;
    int band_lfitbb = -1;
    int64_t nrec = 0;
    int64_t max_pair = 0;

    // ---- solve state ----
    bool have_inverse = false;
    std::vector<double> diag_before; // workspaceDiag (matrix diagonal before inversion)
    bool have_cov = false;           // cov_diag holds the last LAPACK-method solve's covariance diagonal
    // MINRES preconditioner of the current matrix (pede builds it only after a matrix pass,
    // src/pede.f90:8292-8306): kept until the matrix changes; pc_bw/pc_lp: its mbandw/lprecm
    std::shared_ptr<Precond> pc;
    int pc_bw = 0, pc_lp = 0;
    // factor of the dense methods for the current matrix (built on first use after it changed)
    std::unique_ptr<DenseFactor> dfac;
    std::vector<double> cov_diag;
    std::vector<double> last_vec;
    std::vector<unsigned long long> last_cnt;
};

static int fail(mpgpu_ctx *c, int code, const std::string &m)
{
    if (c) c->err = m;
    return code;
}

#define API_TRY try { AllocScope alloc_scope_(c ? c->stream : nullptr);
#define API_CATCH(ctx)                                                                   \
    }                                                                                    \
    catch (const MpError &e) { return fail(ctx, e.code, e.msg); }                        \
    catch (const std::bad_alloc &) { return fail(ctx, 30, "host allocation failed"); }   \
    catch (const std::exception &e) { return fail(ctx, 30, e.what()); }

static double now_ms()
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

extern "C" void mpgpu_default_options(mpgpu_options *o)
{
    std::memset(o, 0, sizeof(*o));
    o->metsol = MPGPU_INVERSION;
    o->matsto = MPGPU_FULL;
    o->mbandw = 0;
    o->lprecm = 0;
    o->lhuber = 0;
    o->lfitbb = 1;
    o->mreqenf = 25;
    o->mreqena = 10;
    o->mcount = 0;
    o->itnlim = 2000;
    o->withmultipliers = 0;   // elimination (icelim = 1, src/mpmod.f90:104)
    o->monitor = 0;
    o->monitor_bins = 100;
    o->mrmode = 0;
    o->chicut = 0.0f;
    o->chirem = 0.0f;
    o->chhuge = 50.0f;
    o->dwcut = 0.0f;
    o->cndlmx = 100.0f;
    o->regula = 1.0f;
    o->default_presigma = 0.0;
    o->mrestl = 1.0e-6;
    o->mrtcnd = 1.0e7;
    o->nregul = 0;
    o->refine_steps = 20;
    o->pair_entries = 1;
    o->skip_empty_cons = 0;
    o->resolve_redundancy_cons = 0;
    o->scale_errors[0] = 1.0;
    o->scale_errors[1] = 1.0;
    o->refine_tol = 0.0;
    o->fp32_min_n = 16384;
}

extern "C" int mpgpu_create(const mpgpu_options *opt, int32_t device, mpgpu_ctx **out)
{
    return mpgpu::create_context(opt, device, nullptr, out);
}

// A context on a stream of the caller's (kept when the context is destroyed), or on a stream of
// its own (stream null)
int mpgpu::create_context(const mpgpu_options *opt, int32_t device, hipStream_t stream, mpgpu_ctx **out)
{
    if (!out) return 30;
    mpgpu_ctx *c = nullptr;
    API_TRY
    c = new mpgpu_ctx();
    if (opt) c->opt = *opt; else mpgpu_default_options(&c->opt);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) throw MpError(30, "no HIP device available");
    if (device < 0 || device >= ndev) throw MpError(30, "bad device ordinal");
    c->device = device;
    HIPCHK(hipSetDevice(device));
    if (stream) {
        c->stream = stream;
        c->own_stream = false;
    } else {
        HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    }
    {
        // freed stream-ordered buffers stay in the device's pool up to 32 GB, so that the
        // buffers of the next calls are made without returning to the driver
        hipMemPool_t pool = nullptr;
        if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess && pool) {
            uint64_t keep = (uint64_t)32 << 30;
            (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
        }
        (void)hipGetLastError();
    }
    *out = c;
    return 0;
    }
    catch (const MpError &e) { std::fprintf(stderr, "mpgpu_create: %s\n", e.msg.c_str()); delete c; *out = nullptr; return e.code; }
    catch (...) { delete c; *out = nullptr; return 30; }
}

extern "C" int mpgpu_destroy(mpgpu_ctx *c)
{
    if (!c) return 0;
    (void)hipSetDevice(c->device);
    // the context's buffers are freed in the order of its stream: the stream goes last
    hipStream_t s = c->stream;
    const bool own = c->own_stream;
    if (s) (void)hipStreamSynchronize(s);
    delete c;
    if (s) {
        (void)hipStreamSynchronize(s);
        if (own) (void)hipStreamDestroy(s);
    }
    return 0;
}

extern "C" const char *mpgpu_last_error(const mpgpu_ctx *c) { return c ? c->err.c_str() : "null context"; }

extern "C" int mpgpu_set_options(mpgpu_ctx *c, const mpgpu_options *o)
{
    if (!c || !o) return 30;
    if (c->built) {
        // options the structure build consumed: the device matrix, the update buffers, the
        // pre-weights and the monitoring histograms are sized and filled for them, so a change
        // would make later passes index memory laid out for another storage
        const mpgpu_options &b = c->opt;
        const char *which = nullptr;
        if (o->matsto != b.matsto) which = "matsto";
        else if (o->withmultipliers != b.withmultipliers) which = "withmultipliers";
        else if (o->mcount != b.mcount) which = "mcount";
        else if (o->mreqenf != b.mreqenf) which = "mreqenf";
        else if (o->monitor != b.monitor || o->monitor_bins != b.monitor_bins) which = "monitor";
        else if (o->lfitbb != b.lfitbb) which = "lfitbb";
        else if (o->default_presigma != b.default_presigma || o->regula != b.regula) which = "presigma/regula";
        else if (o->scale_errors[0] != b.scale_errors[0] || o->scale_errors[1] != b.scale_errors[1]) which = "scaleerrors";
        else if (o->pair_entries != b.pair_entries) which = "pairentries";
        else if (o->iteren != b.iteren) which = "iteren";
        if (which)
            return fail(c, 30, std::string("option ") + which + " is fixed once the structure is built (rebuild the context to change it)");
    }
    c->opt = *o;
    return 0;
}

// ---- record ingest ----------------------------------------------------------------------------
// Device arrays that grow keep their contents (stream-ordered copy; the old buffer is freed after
// it on the same stream).
template <typename T> static void grow_dev(DBuf<T> &b, int64_t used, int64_t cap, hipStream_t s)
{
    DBuf<T> nb;
    nb.alloc((size_t)cap);
    if (used > 0) HIPCHK(hipMemcpyAsync(nb.p, b.p, sizeof(T) * used, hipMemcpyDeviceToDevice, s));
    b.swap(nb);
}
// room on the device for `need` words and their values
static void dev_reserve(mpgpu_ctx *c, int64_t need)
{
    if ((int64_t)c->d_in.n >= need) return;
    const int64_t cap = std::max<int64_t>({need, (int64_t)c->d_in.n + (int64_t)c->d_in.n / 2, (int64_t)1 << 20});
    grow_dev(c->d_in, c->nwords, cap, c->stream);
    if (c->dbl) grow_dev(c->d_vd, c->nwords, cap, c->stream);
    else grow_dev(c->d_vf, c->nwords, cap, c->stream);
}
// the values held so far become FP64 (the first FP64 record or payload after FP32 ones)
static void widen_values(mpgpu_ctx *c)
{
    if (c->dbl) return;
    DBuf<double> vd;
    vd.alloc(std::max<size_t>(c->d_in.n, 1));
    st_widen(c->d_vf.p, vd.p, c->nwords, c->stream);
    c->d_vd.swap(vd);
    c->d_vf.release();
    c->dbl = true;
}
// send the staged words of the current buffer; the next buffer is free to fill on return
static void stage_flush(mpgpu_ctx *c)
{
    Staging &S = *c->stage;
    if (S.n == 0) return;
    hipStream_t s = c->stream;
    dev_reserve(c, c->nwords + S.n);
    HIPCHK(hipMemcpyAsync(c->d_in.p + c->nwords, S.words(), sizeof(int32_t) * S.n, hipMemcpyHostToDevice, s));
    if (c->dbl) HIPCHK(hipMemcpyAsync(c->d_vd.p + c->nwords, S.vals(), sizeof(double) * S.n, hipMemcpyHostToDevice, s));
    else HIPCHK(hipMemcpyAsync(c->d_vf.p + c->nwords, S.vals(), sizeof(float) * S.n, hipMemcpyHostToDevice, s));
    HIPCHK(hipEventRecord(S.done[S.k], s));
    S.inflight[S.k] = true;
    c->nwords += S.n;
    S.n = 0;
    S.k ^= 1;
    if (S.inflight[S.k]) {
        HIPCHK(hipEventSynchronize(S.done[S.k]));
        S.inflight[S.k] = false;
    }
}
// room for a record of nw words in the staging buffer (FP64 values: widen what came before)
static void stage_room(mpgpu_ctx *c, int64_t nw, bool fp64)
{
    if (!c->stage) {
        c->stage = std::make_unique<Staging>();
        c->stage->init();
    }
    if (nw > Staging::CHUNK) throw MpError(18, "record of " + std::to_string(nw) + " words exceeds the reader's buffer");
    if (fp64 && !c->dbl) {
        stage_flush(c);
        widen_values(c);
    }
    if (c->stage->n + nw > Staging::CHUNK) stage_flush(c);
}
// all records on the device: the last staged words are sent and the staging buffers freed
static void stage_finish(mpgpu_ctx *c)
{
    if (!c->stage) return;
    stage_flush(c);
    HIPCHK(hipStreamSynchronize(c->stream));
    c->stage.reset();
}

extern "C" int mpgpu_add_records(mpgpu_ctx *c, int64_t nrec, const int64_t *rec_ptr, const int32_t *inder,
                                 const float *gf, const double *gd)
{
    if (!c) return 30;
    API_TRY
    HIPCHK(hipSetDevice(c->device));
    if (c->built) throw MpError(24, "records cannot be added after the structure was built");
    if (nrec < 0 || (nrec > 0 && (!rec_ptr || !inder || (!gf && !gd)))) throw MpError(24, "bad record arrays");
    if (gf && gd) throw MpError(24, "give either FP32 or FP64 payload");
    if (nrec == 0) return 0;
    int64_t w0 = rec_ptr[0], w1 = rec_ptr[nrec];
    if (w1 < w0) throw MpError(24, "record pointers not ascending");
    for (int64_t r = 0; r < nrec; r++)
        if (rec_ptr[r + 1] < rec_ptr[r]) throw MpError(24, "record pointers not ascending");
    hipStream_t s = c->stream;
    const int64_t total = w1 - w0;
    // records a file read has staged go first; then the arrays are copied straight from the
    // caller's memory (the runtime stages pageable memory itself; the copies end before return,
    // so the caller may reuse its arrays)
    if (c->stage) stage_flush(c);
    const int64_t base = c->nwords;
    if (gd) widen_values(c);
    dev_reserve(c, c->nwords + total);
    HIPCHK(hipMemcpyAsync(c->d_in.p + c->nwords, inder + w0, sizeof(int32_t) * total, hipMemcpyHostToDevice, s));
    std::vector<double> wide;
    if (c->dbl && gd) {
        HIPCHK(hipMemcpyAsync(c->d_vd.p + c->nwords, gd + w0, sizeof(double) * total, hipMemcpyHostToDevice, s));
    } else if (c->dbl) {
        wide.assign(gf + w0, gf + w1);
        HIPCHK(hipMemcpyAsync(c->d_vd.p + c->nwords, wide.data(), sizeof(double) * total, hipMemcpyHostToDevice, s));
    } else {
        HIPCHK(hipMemcpyAsync(c->d_vf.p + c->nwords, gf + w0, sizeof(float) * total, hipMemcpyHostToDevice, s));
    }
    HIPCHK(hipStreamSynchronize(s));
    c->nwords += total;
    for (int64_t r = 0; r < nrec; r++) c->rec_ptr.push_back(base + rec_ptr[r + 1] - w0);
    return 0;
    API_CATCH(c)
}

// Mille binary files (src/mille.f90:208-219; C files as written by the Mille library, which
// pede reads plain or gzip-compressed, src/pede.f90:1689): zlib reads both, so one reader covers
// plain and .gz files of either flavour.  Records are read straight into the pinned staging
// buffers; each full buffer is sent to the device while the next one fills.
extern "C" int mpgpu_read_mille_file(mpgpu_ctx *c, const char *path, int32_t fortran, int64_t *nrec_out)
{
    if (!c) return 30;
    const size_t rp_saved = c->rec_ptr.size();
    const int64_t total_saved = c->nwords + (c->stage ? c->stage->n : 0);
    try {
        AllocScope alloc_scope_(c->stream);
        HIPCHK(hipSetDevice(c->device));
        if (c->built) throw MpError(24, "records cannot be added after the structure was built");
        gzFile f = gzopen(path, "rb");
        if (!f) throw MpError(18, std::string("cannot open binary file ") + path);
        struct Closer { gzFile f; ~Closer() { gzclose(f); } } closer{f};
        gzbuffer(f, 1 << 20);
        {
            // a plain file holds at most size / 8 words (FP32 value and index per word): the
            // device arrays are sized once for it instead of growing
            struct stat sb;
            if (gzdirect(f) && stat(path, &sb) == 0 && sb.st_size > 0)
                dev_reserve(c, c->nwords + (c->stage ? c->stage->n : 0) + (int64_t)sb.st_size / 8 + 1);
        }
        auto get = [&](void *dst, size_t size, size_t count) -> size_t {
            const size_t want = size * count;
            size_t got = 0;
            while (got < want) {
                const unsigned chunk = (unsigned)std::min<size_t>(want - got, 1u << 30);
                const int r = gzread(f, (char *)dst + got, chunk);
                if (r <= 0) break;
                got += (size_t)r;
            }
            return got / size;
        };
        int64_t nrec = 0;
        std::vector<float> tmpf;
        for (;;) {
            int32_t marker = 0, nr = 0;
            if (fortran) {
                if (get(&marker, 4, 1) != 1) break;
                if (get(&nr, 4, 1) != 1) throw MpError(18, "truncated Fortran record");
            } else {
                if (get(&nr, 4, 1) != 1) break;
            }
            const bool fp64 = nr < 0;
            const int64_t nw = (fp64 ? -(int64_t)nr : (int64_t)nr) / 2;
            stage_room(c, nw, fp64);
            Staging &S = *c->stage;
            if (fp64) {
                if (get((double *)S.vals() + S.n, 8, nw) != (size_t)nw) throw MpError(18, "truncated record");
            } else if (c->dbl) {
                tmpf.resize(nw);
                if (get(tmpf.data(), 4, nw) != (size_t)nw) throw MpError(18, "truncated record");
                double *v = (double *)S.vals() + S.n;
                for (int64_t i = 0; i < nw; i++) v[i] = tmpf[i];
            } else {
                if (get((float *)S.vals() + S.n, 4, nw) != (size_t)nw) throw MpError(18, "truncated record");
            }
            if (get(S.words() + S.n, 4, nw) != (size_t)nw) throw MpError(18, "truncated record");
            if (fortran) {
                int32_t m2 = 0;
                if (get(&m2, 4, 1) != 1 || m2 != marker) throw MpError(18, "bad Fortran record marker");
            }
            S.n += nw;
            c->rec_ptr.push_back(c->nwords + S.n);
            nrec++;
        }
        int err = 0;
        gzerror(f, &err);
        if (err != Z_OK && err != Z_STREAM_END) throw MpError(18, std::string("corrupt binary file ") + path);
        if (nrec_out) *nrec_out = nrec;
        return 0;
    } catch (const MpError &e) {
        // the records of this file are dropped (those of earlier calls stay)
        c->rec_ptr.resize(rp_saved);
        if (c->nwords >= total_saved) {
            c->nwords = total_saved;
            if (c->stage) c->stage->n = 0;
        } else if (c->stage) {
            c->stage->n = total_saved - c->nwords;
        }
        return fail(c, e.code, e.msg);
    } catch (const std::bad_alloc &) {
        return fail(c, 30, "host allocation failed");
    } catch (const std::exception &e) {
        return fail(c, 30, e.what());
    }
}

extern "C" int mpgpu_set_steering_parameters(mpgpu_ctx *c, int32_t n, const int32_t *labels, const double *values,
                                             const double *presigma)
{
    if (!c) return 30;
    if (c->built) return fail(c, 24, "parameters must be given before the structure is built");
    for (int32_t k = 0; k < n; k++) {
        c->sp_label.push_back(labels[k]);
        c->sp_value.push_back(values ? values[k] : 0.0);
        c->sp_presig.push_back(presigma ? presigma[k] : 0.0);
    }
    return 0;
}

extern "C" int mpgpu_add_equations(mpgpu_ctx *c, int32_t kind, int32_t nblock, const int64_t *ptr, const int32_t *labels,
                                   const double *factors, const double *rhs, const double *sigma)
{
    if (!c) return 30;
    if (c->built) return fail(c, 24, "equations must be given before the structure is built");
    for (int32_t b = 0; b < nblock; b++) {
        EqBlock e;
        e.kind = kind;
        e.rhs = rhs ? rhs[b] : 0.0;
        e.sigma = sigma ? sigma[b] : 0.0;
        for (int64_t p = ptr[b]; p < ptr[b + 1]; p++) {
            e.labels.push_back(labels[p]);
            e.factors.push_back(factors[p]);
        }
        (kind == 0 ? c->cons : c->meas).push_back(std::move(e));
    }
    return 0;
}

// ---------------------------------------------------------------------------------------------
// structure build
// ---------------------------------------------------------------------------------------------
// Bordered-band structure of one record's local normal matrix: the band-width histogram and
// cost model of loopbf (src/pede.f90:3815-3874).  Returns 0 for the full (sqminv) path, else
// (mbnd+1) | mbdr << 10 | mside << 20.
static int32_t band_info(const int32_t *inder, const int64_t *eq_sl, const uint32_t *eq_n, int64_t e0, int64_t e1,
                         int nalc, std::vector<int> &ibh)
{
    ibh.assign(2 * (size_t)nalc + 2, 0);
    for (int64_t e = e0; e < e1; e++) {
        const int nl = eq_n[e] & 0xffff;
        const int64_t sl = eq_sl[e];
        for (int j = 1; j <= nl; j++) {
            const int ij = inder[sl + j];
            for (int k = 1; k <= j; k++) {
                const int ik = inder[sl + k];
                const int id = std::abs(ij - ik) + 1;
                ibh[id] = std::max(ibh[id], std::min(ij, ik));
                ibh[nalc + id] = std::max(ibh[nalc + id], std::min(nalc + 1 - ij, nalc + 1 - ik));
            }
        }
    }
    int kx = -1, kbdrx = 0, mside = 0;
    int64_t icmn = (int64_t)nalc * nalc * nalc;
    for (int side = 1; side <= 2 && kx < 0; side++) {
        int kbdr = 0;
        for (int k = nalc; k >= 2; k--) {
            const int kbnd = k - 2;
            kbdr = std::max(kbdr, ibh[side == 1 ? k : k + nalc]);
            const int64_t t = kbnd + kbdr + 1;
            const int64_t icost = 6 * (int64_t)(nalc - kbdr) * t * t + 2 * (int64_t)kbdr * kbdr * kbdr;
            if (icost < icmn) { icmn = icost; kx = k; kbdrx = kbdr; mside = side; }
        }
    }
    if (kx <= 0) return 0;
    return (kx - 1) | (kbdrx << 10) | (mside << 20);
}

// Autotuner of the group-block product: the variants with identical arithmetic (rows-pass kernel
// and wavefronts per workgroup; bit-identical results, so the choice never changes a solution)
// are timed on this structure, one warm-up and three timed products each (HIP events on the
// context's stream; the time does not depend on the values), and the fastest median is kept.
static const char *kind_name(int kind) { return kind == 0 ? "rows_u" : kind == 2 ? "rows_u16" : "rows_st"; }
static void tune_block_symv(mpgpu_ctx *c)
{
    c->spmv_variant = -1;
    c->tuning.clear();
    int cand[16];
    const int nc = block_symv_variants(c->uniform_grp, c->plan.max_grp, cand);
    if (nc <= 1 || c->nblock == 0 || !c->d_tb.p || !c->d_cc_blk.p) return;
    // small matrices (below 16 M stored words, 128 MB) keep the default: the timing would cost
    // more than it can save (serving jobs)
    if (c->nstored < ((int64_t)16 << 20) && !std::getenv("MPGPU_SPMV_VARIANT")) {
        c->tuning = "default variant (matrix below 16 M words)";
        return;
    }
    // MPGPU_SPMV_VARIANT=kind*16+w: a fixed variant instead of the timing (tests)
    if (const char *fix = std::getenv("MPGPU_SPMV_VARIANT")) {
        const int v = std::atoi(fix);
        for (int q = 0; q < nc; q++)
            if (cand[q] == v) {
                c->spmv_variant = v;
                c->tuning = "fixed variant " + std::to_string(v);
            }
        return;
    }
    hipStream_t s = c->stream;
    DBuf<double> x, y;
    x.alloc(c->nall);
    y.alloc(c->nall);
    l_zero(c->nall, x.p, s);
    EventPool ev;
    double best = 1e30;
    char buf[96];
    for (int q = 0; q < nc; q++) {
        auto run = [&]() {
            l_block_symv(c->ngroup, c->d_brow.p, c->d_bcf.p, c->d_boff.p, c->d_gfirst.p, c->d_ccol_ptr.p, c->d_cc_off.p,
                         c->d_cc_fn.p, c->d_mat.p, x.p, y.p, c->nvar, c->nall, c->plan.max_grp, c->d_cc_blk.p,
                         c->d_tb.p, c->uniform_grp, c->d_bcol.p, c->nblock, cand[q], s);
        };
        run();
        float ms[3];
        for (int k = 0; k < 3; k++) {
            hipEvent_t e0 = ev.get(), e1 = ev.get();
            HIPCHK(hipEventRecord(e0, s));
            run();
            HIPCHK(hipEventRecord(e1, s));
            HIPCHK(hipEventSynchronize(e1));
            HIPCHK(hipEventElapsedTime(&ms[k], e0, e1));
        }
        std::sort(ms, ms + 3);
        std::snprintf(buf, sizeof buf, "%s%s/w%d %.3f ms", q ? ", " : "", kind_name(cand[q] / 16),
                      cand[q] % 16, ms[1]);
        c->tuning += buf;
        if (ms[1] < best) { best = ms[1]; c->spmv_variant = cand[q]; }
    }
    std::snprintf(buf, sizeof buf, "; chosen %s/w%d", kind_name(c->spmv_variant / 16),
                  c->spmv_variant % 16);
    c->tuning += buf;
    if (std::getenv("MPGPU_VERBOSE")) std::fprintf(stderr, "block product autotuner: %s\n", c->tuning.c_str());
}

extern "C" int mpgpu_build_structure(mpgpu_ctx *c, int32_t *ntot_out, int32_t *nvar_out, int32_t *ncon_out)
{
    if (!c) return 30;
    API_TRY
    HIPCHK(hipSetDevice(c->device));
    // 'pairentries' above 1 drops the matrix updates of rarely filled parameter-group pairs
    // (src/pede.f90:6564, 7036, src/mpbits.f90): not built, refused as the steering parser does
    if (c->opt.pair_entries > 1) throw MpError(40, "pairentries > 1 (dropping rare parameter pairs) is not supported");
    if (c->built) throw MpError(24, "the structure is already built");
    stage_finish(c);
    const int64_t nrec = (int64_t)c->rec_ptr.size() - 1;
    const int64_t nwords = c->nwords;
    c->nrec = nrec;
    // MPGPU_VERBOSE: host time of each phase on stderr
    const bool verbose = std::getenv("MPGPU_VERBOSE") != nullptr;
    double t_phase = now_ms();
    auto phase = [&](const char *name) {
        if (!verbose) return;
        const double t = now_ms();
        std::fprintf(stderr, "build_structure: %-28s %9.1f ms\n", name, t - t_phase);
        t_phase = t;
    };

    hipStream_t s = c->stream;
    // ---- the records are on the device since they were added; the words as read (labels) are
    // kept apart from the words the record kernels use (global words replaced by their position
    // in the record's list, see below) ----
    c->d_rec_w.upload(c->rec_ptr, s);
    DBuf<int32_t> d_lab;
    d_lab.swap(c->d_in);
    if ((int64_t)d_lab.n > nwords + nwords / 16 + 1024) {
        // growth slack of a file read: the value array is kept for the passes, at its size
        if (c->dbl) grow_dev(c->d_vd, nwords, std::max<int64_t>(nwords, 1), s);
        else grow_dev(c->d_vf, nwords, std::max<int64_t>(nwords, 1), s);
    }
    c->d_widx.alloc(std::max<int64_t>(nwords, 1));
    if (nwords > 0)
        HIPCHK(hipMemcpyAsync(c->d_widx.p, d_lab.p, sizeof(int32_t) * nwords, hipMemcpyDeviceToDevice, s));
    StructIn sin;
    sin.nrec = nrec;
    sin.rec_w = c->d_rec_w.p;
    sin.lab = d_lab.p;
    sin.vf = c->dbl ? nullptr : c->d_vf.p;
    sin.vd = c->dbl ? c->d_vd.p : nullptr;
    phase("records to the device");
    // ---- decode all records (equation tables) on the device ----
    const int64_t neq = st_decode(sin, c->d_rec_eq, c->d_eq_sl, c->d_eq_n, s);
    std::vector<int64_t> rec_eq(nrec + 1);
    c->d_rec_eq.download(rec_eq.data(), nrec + 1, s);
    HIPCHK(hipStreamSynchronize(s));
    // host copies of the equation tables, only for the host passes that need them
    std::vector<int64_t> eq_sl;
    std::vector<uint32_t> eq_n;
    auto host_eq = [&]() {
        if ((int64_t)eq_sl.size() == neq) return;
        eq_sl.resize(neq);
        eq_n.resize(neq);
        c->d_eq_sl.download(eq_sl.data(), neq, s);
        c->d_eq_n.download(eq_n.data(), neq, s);
        HIPCHK(hipStreamSynchronize(s));
    };
    phase("decode");

    // ---- label table: all labels of global words, steering parameters and equations; the
    // rank of a label among them is its total index ----
    std::vector<int32_t> ngl_rec;
    std::vector<uint8_t> bnd;
    std::vector<int64_t> recount;   // iterated entries cut: entries in equations without low parameters
    std::vector<int32_t> mon_lab;   // residual monitoring: first global label and error of each equation
    std::vector<double> mon_err;
    {
        std::vector<int32_t> extra(c->sp_label.begin(), c->sp_label.end());
        for (auto &b : c->cons) extra.insert(extra.end(), b.labels.begin(), b.labels.end());
        for (auto &b : c->meas) extra.insert(extra.end(), b.labels.begin(), b.labels.end());
        LabelMap lmap;
        st_label_table(sin, extra, lmap, c->labels, s);
        c->ntot = (int32_t)c->labels.size();
        c->summary.neq_glob = lmap.neq_glob;
        c->summary.nglob_deriv = lmap.nglob_words;
        phase("label table");
        // ---- per record: sorted distinct total indices of its global words (gl lists), the
        // position of each global word in them (widx), entry counters and the group boundaries
        // of pargrp (src/pede.f90:2311-2404): at the first element of every sorted list, after
        // its last, and at both sides of every step that is not +1 -- a label repeated within one
        // equation (pede sorts equation lists without removing duplicates, 2259-2270) is such a
        // step, so it closes the group before it and starts its own.  Counting and lists are per
        // equation, or per record with 'countrecords'; derivatives of zero do not count.
        st_record_lists(sin, lmap, c->ntot, c->opt.mcount != 0, c->d_widx, c->d_rec_gl, c->d_gl_tix, ngl_rec, c->counters,
                        bnd, s);
        if (c->opt.monitor != 0) st_eq_first(sin, mon_lab, mon_err, s);
        // iterated entries cut (loop1i, src/pede.f90:5975-5981, 6164-6265): only with counting
        // per equation; the parameters fixed by the entries cut are those without a negative
        // pre-sigma and with fewer than mreqenf entries
        recount.clear();
        if (c->opt.mcount == 0 && c->opt.iteren > c->opt.mreqenf) {
            std::vector<uint8_t> low(c->ntot, 0);
            std::vector<char> neg(c->ntot, 0);
            for (size_t k = 0; k < c->sp_label.size(); k++) {
                const int32_t t = (int32_t)(std::lower_bound(c->labels.begin(), c->labels.end(), c->sp_label[k]) -
                                            c->labels.begin());
                neg[t] = c->sp_presig[k] < 0.0;
            }
            for (int32_t t = 0; t < c->ntot; t++) low[t] = !neg[t] && c->counters[t] < c->opt.mreqenf;
            st_entries_recount(sin, lmap, low, recount, s);
        }
        d_lab.release();
        sin.lab = nullptr;
    }
    const int32_t ntot = c->ntot;
    bnd[0] = 1;
    auto tix_of = [&](int32_t l) -> int32_t {   // 0-based total index of a label in use
        return (int32_t)(std::lower_bound(c->labels.begin(), c->labels.end(), l) - c->labels.begin());
    };

    // residual monitoring: every measurement with local and global derivatives belongs to the
    // group of its first global parameter; the group's mean error is summed in record order
    // (loop2, src/pede.f90:6700-6710 and 7075-7082)
    c->mon_tot.clear();
    c->mon_res.clear();
    std::vector<int32_t> mon_grp;
    if (c->opt.monitor != 0) {
        std::vector<int64_t> mcnt(ntot, 0);
        std::vector<double> msum(ntot, 0.0);
        for (int64_t e = 0; e < neq; e++) {
            if (mon_lab[e] == 0) continue;
            const int32_t t = tix_of(mon_lab[e]);
            mcnt[t]++;
            // 'scaleerrors' applies to the errors pede reads here as well (prepareRecords
            // before loop2, src/pede.f90:6638)
            msum[t] = msum[t] + mon_err[e] * c->opt.scale_errors[0];
        }
        mon_grp.assign(ntot, -1);
        for (int32_t t = 0; t < ntot; t++)
            if (mcnt[t] > 0) {
                mon_grp[t] = (int32_t)c->mon_tot.size();
                c->mon_tot.push_back(t);
                c->mon_res.push_back(msum[t] / (double)mcnt[t]);
            }
        c->mon_bins = std::max(1, c->opt.monitor_bins);
    }
    phase("record label lists");
    // ---- parameters: start values, pre-sigmas, variable / fixed (loop1, src/pede.f90:5920-6062) ----
    c->par.assign(ntot, 0.0);
    c->presig_tot.assign(ntot, 0.0);
    for (size_t k = 0; k < c->sp_label.size(); k++) {
        int32_t t = tix_of(c->sp_label[k]);
        c->par[t] = c->sp_value[k];
        c->presig_tot[t] = c->sp_presig[k];
    }
    // constraint groups (grpcon, src/pede.f90:854-1060).  The constraints are taken in the order
    // of their (first, last) parameter; one that no earlier constraint has marked joins the group
    // of the first related constraint (one sharing a parameter) that has a group, or opens a new
    // group; it then sets the group of its parameters and marks every related constraint with it
    // (constraints marked before keep their group unless re-marked, parameters keep the group of
    // the last constraint that set them).  A group counts the parameters in its range whose group
    // is that group; with as many constraints as such parameters and 'resolveredundancycons' the
    // parameters become fixed (-4) and the group's constraints are dropped.  Constraints with the
    // same (first, last) parameter share a parameter and end in one group whatever their order.
    std::vector<char> redundant(ntot, 0), con_dropped(c->cons.size(), 0);
    if (c->opt.resolve_redundancy_cons && !c->cons.empty()) {
        const int nc = (int)c->cons.size();
        std::vector<std::vector<int32_t>> plist(nc);
        std::vector<std::vector<int32_t>> parcons(ntot);
        for (int j = 0; j < nc; j++) {
            for (int32_t lab : c->cons[j].labels) {
                const int32_t t = tix_of(lab);
                plist[j].push_back(t);
                parcons[t].push_back(j);            // every occurrence, in input order
            }
            std::sort(plist[j].begin(), plist[j].end());
            plist[j].erase(std::unique(plist[j].begin(), plist[j].end()), plist[j].end());
        }
        std::vector<int> ord;                       // constraints with parameters
        for (int j = 0; j < nc; j++) if (!plist[j].empty()) ord.push_back(j);
        std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) {
            if (plist[x].front() != plist[y].front()) return plist[x].front() < plist[y].front();
            return plist[x].back() < plist[y].back();
        });
        std::vector<int> grp(nc, 0), gpc(ntot, 0);
        int ngrp = 0;
        for (int jc : ord) {
            int g = grp[jc];
            if (g == 0) {
                for (int32_t t : plist[jc]) {
                    for (int j2 : parcons[t]) if ((g = grp[j2]) > 0) break;
                    if (g > 0) break;
                }
                if (g == 0) g = ++ngrp;
            }
            for (int32_t t : plist[jc]) {
                gpc[t] = g;
                for (int j2 : parcons[t]) grp[j2] = g;
            }
        }
        std::vector<int64_t> ncon_g(ngrp + 1, 0);
        std::vector<int32_t> lo(ngrp + 1, INT32_MAX), hi(ngrp + 1, -1);
        for (int j = 0; j < nc; j++) {
            const int g = grp[j];
            if (g == 0) continue;
            ncon_g[g]++;
            lo[g] = std::min(lo[g], plist[j].front());
            hi[g] = std::max(hi[g], plist[j].back());
        }
        std::vector<char> red_g(ngrp + 1, 0);
        for (int g = 1; g <= ngrp; g++) {
            if (ncon_g[g] == 0) continue;
            int64_t npar = 0;
            for (int32_t t = lo[g]; t <= hi[g]; t++) npar += gpc[t] == g;
            if (npar == ncon_g[g]) {
                red_g[g] = 1;
                for (int32_t t = lo[g]; t <= hi[g]; t++) if (gpc[t] == g) redundant[t] = 1;
            }
        }
        for (int j = 0; j < nc; j++) if (grp[j] > 0 && red_g[grp[j]]) con_dropped[j] = 1;
    }
    c->var_index.assign(ntot, 0);
    c->var_to_tot.clear();
    int32_t nv = 0;
    for (int32_t t = 0; t < ntot; t++) {
        if (c->presig_tot[t] < 0.0) c->var_index[t] = -1;
        else if (c->counters[t] < c->opt.mreqenf) c->var_index[t] = -2;
        else if (redundant[t]) c->var_index[t] = -4;
        else if (!recount.empty() && recount[t] < c->opt.mreqenf && c->counters[t] < c->opt.iteren) c->var_index[t] = -3;
        else { c->var_index[t] = ++nv; c->var_to_tot.push_back(t); }
    }
    c->nvar = nv;
    if (nv <= 0) throw MpError(22, "Aborted, no variable global parameters");
    // group start labels over total indices
    c->group_start_label.assign(ntot, 0);
    {
        int32_t start = 0;
        for (int32_t t = 0; t < ntot; t++) {
            if (bnd[t]) start = t;
            c->group_start_label[t] = c->labels[start];
        }
    }
    // variable groups: consecutive variables of the same total-index group
    c->group_of.assign(nv + 1, 0);
    c->gfirst.clear();
    {
        int32_t lastgrp_start = -1;
        for (int32_t v = 1; v <= nv; v++) {
            int32_t t = c->var_to_tot[v - 1];
            int32_t gs = c->group_start_label[t];
            if (gs != lastgrp_start || c->gfirst.empty()) {
                c->gfirst.push_back(v);
                lastgrp_start = gs;
            }
            c->group_of[v] = (int32_t)c->gfirst.size() - 1;
        }
        c->ngroup = (int32_t)c->gfirst.size();
        c->gfirst.push_back(nv + 1);
    }
    // pre-weights (src/pede.f90:6050-6062)
    c->pre_weight.assign(nv, 0.0);
    for (int32_t v = 1; v <= nv; v++) {
        double ps = c->presig_tot[c->var_to_tot[v - 1]];
        double w = 0.0;
        if (ps > 0.0) w = 1.0 / (ps * ps);
        else if (ps == 0.0 && c->opt.default_presigma > 0.0) w = 1.0 / (c->opt.default_presigma * c->opt.default_presigma);
        c->pre_weight[v - 1] = (double)c->opt.regula * w;
    }

    phase("parameters and groups");
    // ---- constraints: Lagrange rows over variable indices (addcst, src/pede.f90:704-771) ----
    c->cptr.assign(1, 0);
    c->cidx.clear();
    c->cval.clear();
    c->con_tot.clear();
    c->crhs.clear();
    int64_t nempty = 0;
    for (size_t jc = 0; jc < c->cons.size(); jc++) {
        const auto &b = c->cons[jc];
        if (con_dropped[jc]) continue;
        std::vector<std::pair<int32_t, double>> tl;
        for (size_t q = 0; q < b.labels.size(); q++) tl.push_back({tix_of(b.labels[q]), b.factors[q]});
        // a constraint without any variable parameter is empty: skipped with 'skipemptycons';
        // without it pede checks the input only and ends without a solution (exit 5,
        // src/pede.f90:1160-1190)
        bool any = false;
        for (auto &p : tl) if (c->var_index[p.first] > 0) any = true;
        if (!any) { nempty++; continue; }
        c->con_tot.push_back(tl);
        c->crhs.push_back(b.rhs);
        // accumulate factors per variable (duplicates add, as repeated mupdat calls do)
        std::vector<std::pair<int32_t, double>> vl;
        for (auto &p : tl) if (c->var_index[p.first] > 0) vl.push_back({c->var_index[p.first], p.second});
        std::sort(vl.begin(), vl.end(), [](auto &a, auto &b2) { return a.first < b2.first; });
        for (size_t q = 0; q < vl.size();) {
            size_t q2 = q;
            double s = 0.0;
            while (q2 < vl.size() && vl[q2].first == vl[q].first) s += vl[q2++].second;
            c->cidx.push_back(vl[q].first);
            c->cval.push_back(s);
            q = q2;
        }
        c->cptr.push_back((int64_t)c->cidx.size());
    }
    if (nempty > 0 && !c->opt.skip_empty_cons)
        throw MpError(5, std::to_string(nempty) + " empty constraints detected (use option 'skipemptycons' to skip "
                         "those): ended without solution");
    // Without 'withmultipliers' the MINRES methods eliminate the constraints (pede's default,
    // src/pede.f90:6492-6498); inversion and decomposition keep Lagrange multipliers (the same
    // constrained solution; decomposition augments the parameter block, see mpgpu_solve)
    const bool minres_method = c->opt.metsol == MPGPU_MINRES || c->opt.metsol == MPGPU_MINRES_QLP;
    c->elim = !c->opt.withmultipliers && !c->con_tot.empty() && minres_method;
    c->ncon = c->elim ? 0 : (int32_t)c->con_tot.size();
    c->nall = c->nvar + c->ncon;
    c->ql.reset();
    c->nfit = c->nvar;
    if (c->elim) {
        c->ql = std::make_unique<QlElim>();
        c->ql->build(c->nvar, c->con_tot, c->var_index);
        c->nfit = c->ql->nfit;
    }

    phase("constraints");
    // ---- matrix structure ----
    const bool sparse = c->opt.matsto == MPGPU_SPARSE;
    c->d_group_of.upload(c->group_of, s);
    c->d_gfirst.upload(c->gfirst, s);
    c->d_var.upload(c->var_index, s);
    {
        std::vector<int32_t> gt(ntot, -1);
        for (int32_t t = 0; t < ntot; t++)
            if (c->var_index[t] > 0) gt[t] = c->group_of[c->var_index[t]];
        c->d_grp_tot.upload(gt, s);
    }
    // per record on the device: segments (parameter groups of its variables), local-fit sizes,
    // update size, first group; sparse storage: the group pairs it couples
    std::vector<int32_t> rst;
    std::vector<int64_t> cnt_r;
    DBuf<int32_t> d_rst;
    PairMap pm;
    int32_t max_ngl_all = 1;
    for (int32_t x : ngl_rec) max_ngl_all = std::max(max_ngl_all, x);
    st_record_stats(sin, c->d_widx, c->d_rec_gl, c->d_gl_tix, c->d_var.p, c->d_group_of.p, c->ngroup, max_ngl_all, sparse,
                    pm, rst, cnt_r, d_rst, s);
    const int32_t *st_nv = rst.data(), *st_nseg = rst.data() + nrec, *st_nalc = rst.data() + 2 * nrec,
                  *st_sup = rst.data() + 3 * nrec, *st_qe = rst.data() + 4 * nrec;
    phase("record stats and group pairs");
    {
        // loop1 / loop2 counts for millepede.log.  A label belongs to a parameter group when it
        // occurs with a non-zero derivative (zero derivatives are skipped for grouping,
        // src/pede.f90:2219-2222), so the groups over all parameters are the group starts among
        // labels with entries; labels only named in text files are in none (src/pede.f90:5986-5991).
        mpgpu_record_summary &sm = c->summary;
        sm.nrec = nrec;
        sm.neq = neq;
        sm.fp64_payload = c->dbl ? 1 : 0;
        sm.ntot = ntot;
        sm.nvar = c->nvar;
        sm.ngroup_var = c->ngroup;
        sm.ncon = (int32_t)c->con_tot.size();
        sm.ngroup_tot = 0;
        for (int32_t t = 0; t < ntot; t++) sm.ngroup_tot += bnd[t] && c->counters[t] > 0;
        sm.max_glob_rec = sm.max_loc_rec = sm.max_eq_rec = 0;
        sm.nempty_rec = 0;
        for (int64_t r = 0; r < nrec; r++) {
            sm.max_glob_rec = std::max(sm.max_glob_rec, st_nv[r]);
            sm.max_loc_rec = std::max(sm.max_loc_rec, st_nalc[r]);
            sm.max_eq_rec = std::max(sm.max_eq_rec, (int32_t)(rec_eq[r + 1] - rec_eq[r]));
            sm.nempty_rec += st_nv[r] == 0;
        }
    }
    int32_t max_gl = 1, max_eq = 1, max_nalc = 1, max_var = 1, max_seg = 1;
    int32_t bmax_gl = 1, bmax_eq = 1, bmax_nalc = 1, bmax_var = 1, bmax_seg = 1, bmax_mp1 = 1, bmax_bdr = 0, bmax_sup = 1;
    int32_t max_qe = 1;   // most equations any record variable occurs in (full local fits)
    std::vector<int32_t> binfo(nrec, 0);
    const int lfitbb = c->opt.lfitbb;
    // the record words with local indices (band detection; global words hold their list position)
    std::vector<int32_t> words;
    if (lfitbb > 0)
        for (int64_t r = 0; r < nrec; r++)
            if (st_nalc[r] > 5) {
                host_eq();
                words.resize(std::max<int64_t>(nwords, 1));
                c->d_widx.download(words.data(), nwords, s);
                HIPCHK(hipStreamSynchronize(s));
                break;
            }
#pragma omp parallel for schedule(dynamic, 4096) reduction(max : max_gl, max_eq, max_nalc, max_var, max_seg, bmax_gl, \
    bmax_eq, bmax_nalc, bmax_var, bmax_seg, bmax_mp1, bmax_bdr, bmax_sup, max_qe)
    for (int64_t r = 0; r < nrec; r++) {
        thread_local std::vector<int> ibh;
        const int32_t ngl = ngl_rec[r], neq_r = (int32_t)(rec_eq[r + 1] - rec_eq[r]);
        const int32_t nalc = st_nalc[r], nv_r = st_nv[r], nseg = st_nseg[r];
        // bordered-band local fit (decided once, as pede does in the first iteration)
        int32_t bi = 0;
        if (nalc > 5 && lfitbb > 0) bi = band_info(words.data(), eq_sl.data(), eq_n.data(), rec_eq[r], rec_eq[r + 1], nalc, ibh);
        binfo[r] = bi;
        if (bi) {
            bmax_gl = std::max(bmax_gl, ngl);
            bmax_eq = std::max(bmax_eq, neq_r);
            bmax_nalc = std::max(bmax_nalc, nalc);
            bmax_var = std::max(bmax_var, nv_r);
            bmax_seg = std::max(bmax_seg, nseg);
            bmax_mp1 = std::max(bmax_mp1, bi & 0x3ff);
            bmax_bdr = std::max(bmax_bdr, (bi >> 10) & 0x3ff);
            bmax_sup = std::max(bmax_sup, st_sup[r]);
        } else {
            max_gl = std::max(max_gl, ngl);
            max_eq = std::max(max_eq, neq_r);
            max_nalc = std::max(max_nalc, nalc);
            max_var = std::max(max_var, nv_r);
            max_qe = std::max(max_qe, st_qe[r]);
        }
        max_seg = std::max(max_seg, nseg);
    }
    BlockDev blk;
    if (sparse) {
        // external measurements couple their parameters too; every diagonal block exists
        // (pre-sigma, preconditioner)
        std::vector<int64_t> xp;
        for (auto &b : c->meas) {
            std::vector<int32_t> gs;
            for (int32_t l : b.labels) { int v = c->var_index[tix_of(l)]; if (v > 0) gs.push_back(c->group_of[v]); }
            for (int32_t a : gs) for (int32_t b2 : gs) if (a >= b2) xp.push_back((int64_t)a * (a + 1) / 2 + b2);
        }
        st_blocks(pm, c->ngroup, c->d_gfirst.p, xp, blk, s);
        c->nblock = blk.nblock;
        c->nstored = blk.nstored;
        c->blocks_on_host = false;
        c->d_brow.swap(blk.brow);
        c->d_bcol.swap(blk.bcol);
        c->d_brg.swap(blk.brg);
        c->d_boff.swap(blk.boff);
        c->d_bcf.swap(blk.bcf);
        c->d_ccol_ptr.swap(blk.ccol_ptr);
        c->d_cc_off.swap(blk.cc_off);
        c->d_cc_fn.swap(blk.cc_fn);
        // per record: block id of each segment pair
        st_pair_blocks(sin, c->d_rec_gl, c->d_gl_tix, c->d_var.p, c->d_group_of.p, max_seg, d_rst, pm, c->d_rec_pb,
                       c->d_pair_blk, s);
    } else {
        int64_t n = c->nall;
        c->nstored = n * (n + 1) / 2;
        c->nblock = 0;
        c->blocks_on_host = true;
        c->brow.clear();
        c->bcol.clear();
        c->brg.clear();
        c->boff.clear();
    }
    c->max_pair = (int64_t)max_seg * (max_seg + 1) / 2;
    c->plan.max_gl = max_gl;
    c->plan.max_eq = max_eq;
    c->plan.nalc_max = std::max(1, std::min(max_nalc, 64));
    c->plan.max_var = max_var;
    c->plan.max_pair = (int32_t)std::max<int64_t>(1, c->max_pair);
    c->plan.qe = std::min(8, std::max(1, max_qe));
    {
        int64_t mw = 1;
        for (int64_t r = 0; r < nrec; r++)
            if (!binfo[r]) mw = std::max<int64_t>(mw, c->rec_ptr[r + 1] - c->rec_ptr[r]);
        c->plan.max_w = (int32_t)mw;
    }
    c->plan.gh = sparse ? 0 : 1;
    {
        int32_t maxgrp = 1;
        for (int32_t g = 0; g < c->ngroup; g++) maxgrp = std::max(maxgrp, c->gfirst[g + 1] - c->gfirst[g]);
        // group-block storage: row descriptors of the update buffer
        c->plan.stage = sparse ? 1 : 0;
        c->plan.max_grp = maxgrp;
        // every group of one size: the block SpMV knows the block layout
        bool uni = true;
        for (int32_t g = 0; g < c->ngroup; g++) uni = uni && c->gfirst[g + 1] - c->gfirst[g] == maxgrp;
        c->uniform_grp = uni ? maxgrp : 0;
    }
    lds_bytes(c->plan);
    if (c->plan.bytes > 160 * 1024)
        throw MpError(30, "record too large for the record kernel (LDS " + std::to_string(c->plan.bytes) + " bytes)");
    // bordered-band records: their own kernel, LDS plan and per-workgroup inverse scratch
    c->nband = 0;
    for (int64_t r = 0; r < nrec; r++) if (binfo[r]) c->nband++;
    c->band_lfitbb = lfitbb;
    if (c->nband > 0) {
        if (bmax_nalc > 256) throw MpError(30, "bordered-band local fits support at most 256 local parameters");
        if (bmax_bdr > 16) throw MpError(30, "bordered-band local fits support at most 16 border rows");
        BandPlan &B = c->bplan;
        B.max_gl = bmax_gl;
        B.max_eq = bmax_eq;
        B.nalc_max = bmax_nalc;
        B.max_var = bmax_var;
        B.max_pair = sparse ? (int32_t)((int64_t)bmax_seg * (bmax_seg + 1) / 2) : 1;
        B.mp1_max = bmax_mp1;
        B.nbdr_max = std::max(1, bmax_bdr);
        B.max_sup = bmax_sup;
        B.vinv_stride = (int64_t)bmax_nalc * (bmax_nalc + 1) / 2;
        band_lds_bytes(B);
        if (B.bytes > 160 * 1024)
            throw MpError(30, "record too large for the bordered-band kernel (LDS " + std::to_string(B.bytes) + " bytes)");
        const int per_cu = std::max(1, std::min(8, (160 * 1024) / B.bytes));
        c->band_grid = (int)std::min<int64_t>(c->nband, 256 * per_cu);
    }

    phase("matrix structure");
    // ---- processing order: records sorted by the first parameter group they touch, so that the
    // records in flight at the same time update a compact set of matrix blocks (cache reuse of
    // the FP64 atomics); results do not depend on the order beyond FP64 summation order
    std::vector<int64_t> order;
    st_order(nrec, c->ngroup, d_rst, c->d_order, order, s);
    phase("processing order");
    // ---- two-phase build index (group-block storage) ----
    // Each record writes its update of every segment pair (r-major block, lower triangle for the
    // diagonal pairs) contiguously into a buffer; the block gather sums, for every stored block,
    // the values of the records that touch it, in processing order.  Records are taken in
    // batches whose updates fit the buffer budget.
    c->batches.clear();
    if (sparse) {
        std::vector<int64_t> rec_cb(nrec, -1);
        std::vector<int32_t> rec_cn(nrec, 0);
        for (int64_t r = 0; r < nrec; r++)
            if (binfo[r]) cnt_r[r] = 0;             // bordered-band records update the matrix directly
        // update-buffer cap: MPGPU_UPDATE_BUFFER_GB, else 45 % of the free device memory up to
        // 128 GB (a whole pass of the large workloads in one batch: the matrix is read and
        // written once per pass, one record-kernel launch)
        const char *env = std::getenv("MPGPU_UPDATE_BUFFER_GB");
        double gb = 64.0;
        if (env) gb = std::atof(env);
        else {
            size_t fr = 0, tot = 0;
            if (hipMemGetInfo(&fr, &tot) == hipSuccess && fr > 0) gb = std::min(128.0, 0.45 * (double)fr / (1ull << 30));
            else (void)hipGetLastError();
        }
        const int64_t budget = std::max<int64_t>(1 << 20, (int64_t)(gb * (1ull << 30) / 8.0));
        std::vector<int64_t> bat_inc;                   // incidences of every batch
        int64_t q0 = 0, acc = 0, ninc = 0;
        for (int64_t q = 0; q < nrec; q++) {
            const int64_t r = order[q];
            if (acc + cnt_r[r] > budget && q > q0) {
                c->batches.push_back({q0, q, 0, acc, 0});
                bat_inc.push_back(ninc);
                q0 = q;
                acc = 0;
                ninc = 0;
            }
            rec_cb[r] = cnt_r[r] > 0 ? acc : -1;
            rec_cn[r] = (int32_t)cnt_r[r];
            acc += cnt_r[r];
            if (cnt_r[r] > 0) ninc += (int64_t)st_nseg[r] * (st_nseg[r] + 1) / 2;
        }
        if (nrec > q0 || c->batches.empty()) {
            c->batches.push_back({q0, nrec, 0, acc, 0});
            bat_inc.push_back(ninc);
        }
        c->d_rec_cb.upload(rec_cb, s);
        c->d_rec_cn.upload(rec_cn, s);
        // incidences: per batch, per block, buffer offsets of the (record, pair) values
        const int64_t nb = c->nblock, nbat = (int64_t)c->batches.size();
        int64_t inc_total = 0;
        for (int64_t bt = 0; bt < nbat; bt++) {
            c->batches[bt].inc0 = inc_total;
            inc_total += bat_inc[bt];
        }
        c->d_inc_ptr.alloc((size_t)nbat * (nb + 1));
        c->d_inc_src.alloc(std::max<int64_t>(inc_total, 1));
        for (int64_t bt = 0; bt < nbat; bt++) {
            const auto &B = c->batches[bt];
            st_incidence(sin, c->d_rec_gl, c->d_gl_tix, c->d_var.p, c->d_group_of.p, max_seg, c->d_order, c->d_rec_pb,
                         c->d_pair_blk, c->d_rec_cb, B.q0, B.q1, nb, c->d_inc_ptr.p + (size_t)bt * (nb + 1),
                         c->d_inc_src.p + B.inc0, bat_inc[bt], s);
        }
        int64_t mx = 1;
        for (auto &B : c->batches) mx = std::max(mx, B.len);
        c->d_cbuf.alloc((size_t)mx);
        // the transposed products of the off-diagonal blocks (small groups): column-list blocks
        if (c->plan.max_grp <= 8) {
            c->d_cc_blk.swap(blk.cc_blk);
            // transposed products of the blocks, and a spare slot per lane for the uniform kernel
            c->d_tb.alloc((size_t)c->plan.max_grp * std::max<int64_t>(1, c->nblock) + 64);
        } else {
            c->d_cc_blk.release();
            c->d_tb.release();
        }
    } else {
        c->d_rec_cb.release();
        c->d_rec_cn.release();
        c->d_inc_ptr.release();
        c->d_inc_src.release();
        c->d_cbuf.release();
        c->d_rec_pb.release();
        c->d_pair_blk.release();
    }
    phase("two-phase index");
    // ---- remaining device state ----
    if (c->nband > 0) {
        std::vector<int64_t> ids;
        ids.reserve(c->nband);
        for (int64_t q = 0; q < nrec; q++) if (binfo[order[q]]) ids.push_back(order[q]);
        c->d_band_ids.upload(ids, s);
        c->d_band_info.upload(binfo, s);
        c->d_band_scratch.alloc((size_t)c->band_grid * c->bplan.vinv_stride);
    } else {
        c->d_band_ids.release();
        c->d_band_info.release();
        c->d_band_scratch.release();
    }
    c->d_par.upload(c->par, s);
    c->d_pre_weight.upload(c->pre_weight, s);
    c->d_mat.alloc(std::max<int64_t>(1, c->nstored));
    c->d_vec.alloc(c->nall);
    c->d_cnt.alloc(c->nvar);
    c->d_chi2.alloc(std::max<int64_t>(1, nrec));
    c->d_ndf.alloc(std::max<int64_t>(1, nrec));
    c->d_status.alloc(std::max<int64_t>(1, nrec));
    c->d_part.alloc(2048);
    l_zero(1, c->d_part.p + 2047, s);       // the completion counter of the one-launch dot product
    c->d_scal.alloc(16);
    if (!sparse) c->d_psymv.alloc(packed_symv_scratch(c->nall));
    c->d_ndfsum.alloc(1);
    c->d_hist.alloc(24);
    if (c->ncon > 0) {
        c->d_cptr.upload(c->cptr, s);
        c->d_cidx.upload(c->cidx, s);
        c->d_cval.upload(c->cval, s);
    }
    if (c->elim) c->ql->upload(s);
    if (c->opt.monitor != 0 && !c->mon_tot.empty()) {
        c->d_mon_grp.upload(mon_grp, s);
        c->d_mon_res.upload(c->mon_res, s);
        c->d_mon_hist.alloc((size_t)c->mon_tot.size() * c->mon_bins);
        c->d_mon_hist.zero(s);
    } else {
        c->d_mon_grp.release();
        c->d_mon_res.release();
        c->d_mon_hist.release();
    }
    HIPCHK(hipStreamSynchronize(s));
    phase("device upload");
    if (sparse) tune_block_symv(c);
    phase("product autotuner");
    c->built = true;
    c->pc.reset();
    c->dfac.reset();
    if (ntot_out) *ntot_out = c->ntot;
    if (nvar_out) *nvar_out = c->nvar;
    if (ncon_out) *ncon_out = c->ncon;
    return 0;
    API_CATCH(c)
}

extern "C" int mpgpu_get_labels(const mpgpu_ctx *c, int32_t *labels, int32_t *var_index, int64_t *counters,
                                int32_t *group_first_label)
{
    if (!c || !c->built) return 30;
    for (int32_t t = 0; t < c->ntot; t++) {
        if (labels) labels[t] = c->labels[t];
        if (var_index) var_index[t] = c->var_index[t];
        if (counters) counters[t] = c->counters[t];
        if (group_first_label) group_first_label[t] = c->group_start_label[t];
    }
    return 0;
}

extern "C" int mpgpu_get_record_summary(const mpgpu_ctx *c, mpgpu_record_summary *out)
{
    if (!c || !out) return 30;
    if (!c->built) return 24;
    *out = c->summary;
    return 0;
}

extern "C" int mpgpu_get_tuning(const mpgpu_ctx *c, char *text, int32_t len)
{
    if (!c || !c->built || !text || len <= 0) return 30;
    std::snprintf(text, (size_t)len, "%s", c->tuning.c_str());
    return 0;
}

extern "C" int mpgpu_matrix_info(const mpgpu_ctx *c, int64_t *stored, int64_t *nblocks, int32_t *ngroups)
{
    if (!c || !c->built) return 30;
    if (stored) *stored = c->nstored;
    if (nblocks) *nblocks = c->nblock;
    if (ngroups) *ngroups = c->ngroup;
    return 0;
}

// host copy of the block structure (built on the device)
static void host_blocks(mpgpu_ctx *c)
{
    if (c->blocks_on_host) return;
    const int64_t nb = c->nblock;
    c->brow.resize((size_t)c->ngroup + 1);
    c->bcol.resize(nb);
    c->brg.resize(nb);
    c->boff.resize(nb);
    c->d_brow.download(c->brow.data(), c->brow.size(), c->stream);
    c->d_bcol.download(c->bcol.data(), nb, c->stream);
    c->d_brg.download(c->brg.data(), nb, c->stream);
    c->d_boff.download(c->boff.data(), nb, c->stream);
    HIPCHK(hipStreamSynchronize(c->stream));
    c->blocks_on_host = true;
}

extern "C" int mpgpu_get_block_structure(const mpgpu_ctx *c0, int32_t *gfirst, int64_t *brow, int32_t *bcol,
                                         int64_t *boff)
{
    mpgpu_ctx *c = const_cast<mpgpu_ctx *>(c0);
    if (!c || !c->built) return 30;
    if (c->opt.matsto != MPGPU_SPARSE) return fail(c, 24, "block structure exists for sparse storage only");
    API_TRY
    host_blocks(c);
    if (gfirst) std::copy(c->gfirst.begin(), c->gfirst.end(), gfirst);
    if (brow) std::copy(c->brow.begin(), c->brow.end(), brow);
    if (bcol) std::copy(c->bcol.begin(), c->bcol.end(), bcol);
    if (boff) std::copy(c->boff.begin(), c->boff.begin() + c->nblock, boff);
    return 0;
    API_CATCH(c)
}

extern "C" int mpgpu_set_parameters(mpgpu_ctx *c, const double *p)
{
    if (!c || !c->built) return 30;
    API_TRY
    std::copy(p, p + c->ntot, c->par.begin());
    c->d_par.upload(c->par, c->stream);
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
    API_CATCH(c)
}

extern "C" int mpgpu_get_parameters(const mpgpu_ctx *c, double *p)
{
    if (!c || !c->built) return 30;
    std::copy(c->par.begin(), c->par.end(), p);
    return 0;
}

static DevModel model_of(mpgpu_ctx *c)
{
    DevModel M{};
    M.par = c->d_par.p;
    M.var = c->d_var.p;
    M.nvar = c->nvar;
    M.sparse = c->opt.matsto == MPGPU_SPARSE;
    M.mat = c->d_mat.p;
    M.group_of = c->d_group_of.p;
    M.grp_tot = c->d_grp_tot.p;
    M.gfirst = c->d_gfirst.p;
    M.brow = c->d_brow.p;
    M.bcol = c->d_bcol.p;
    M.boff = c->d_boff.p;
    M.vec = c->d_vec.p;
    M.cnt = c->d_cnt.p;
    return M;
}

static DevRecords records_of(mpgpu_ctx *c)
{
    DevRecords R{};
    R.nrec = c->nrec;
    R.rec_eq = c->d_rec_eq.p;
    R.rec_word = c->d_rec_w.p;
    R.rec_gl = c->d_rec_gl.p;
    R.rec_pb = c->d_rec_pb.p;
    R.eq_sl = c->d_eq_sl.p;
    R.eq_n = c->d_eq_n.p;
    R.vf = c->dbl ? nullptr : c->d_vf.p;
    R.vd = c->dbl ? c->d_vd.p : nullptr;
    R.widx = c->d_widx.p;
    R.gl_tix = c->d_gl_tix.p;
    R.pair_blk = c->d_pair_blk.p;
    R.rec_w = nullptr;
    R.order = c->d_order.p;
    R.band = c->nband > 0 ? c->d_band_info.p : nullptr;
    const bool two = c->opt.matsto == MPGPU_SPARSE;
    R.rec_cb = two ? c->d_rec_cb.p : nullptr;
    R.rec_cn = two ? c->d_rec_cn.p : nullptr;
    R.cbuf = two ? c->d_cbuf.p : nullptr;
    R.max_gl = c->plan.max_gl;
    R.max_eq = c->plan.max_eq;
    R.max_nalc = c->plan.nalc_max;
    R.max_var = c->plan.max_var;
    R.max_seg = 0;
    return R;
}

// host-side element updates routed to the device matrix (mupdat semantics)
static void device_mupdat(mpgpu_ctx *c, const std::vector<int32_t> &ii, const std::vector<int32_t> &jj,
                          const std::vector<double> &vv)
{
    if (ii.empty()) return;
    DBuf<int32_t> di, dj;
    DBuf<double> dv;
    di.upload(ii, c->stream);
    dj.upload(jj, c->stream);
    dv.upload(vv, c->stream);
    l_mupdat(model_of(c), (int32_t)ii.size(), di.p, dj.p, dv.p, c->stream);
    HIPCHK(hipStreamSynchronize(c->stream));
    c->pc.reset();
    c->dfac.reset();
}

extern "C" int mpgpu_loopn(mpgpu_ctx *c, int32_t icalcm, int32_t nloopn, float chicut, mpgpu_pass_stats *st)
{
    if (!c || !c->built) return fail(c, 30, "structure not built");
    if (c->opt.lfitbb != c->band_lfitbb)
        return fail(c, 30, "the local-fit band option changed after the structure was built");
    API_TRY
    HIPCHK(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    EventPool ev;    // destroys its events on every exit
    hipEvent_t e0 = ev.get(), e1 = ev.get(), e2 = ev.get();
    std::vector<std::pair<hipEvent_t, hipEvent_t>> t_rec, t_gat;   // per launch
    auto timed = [&](std::vector<std::pair<hipEvent_t, hipEvent_t>> &v) {
        v.push_back({ev.get(), ev.get()});
        return v.back();
    };
    HIPCHK(hipEventRecord(e0, s));
    c->d_vec.zero(s);
    c->d_cnt.zero(s);
    if (icalcm == 1) { c->d_mat.zero(s); c->have_inverse = false; c->have_cov = false; c->pc.reset(); c->dfac.reset(); }
    PassOpts o{};
    o.icalcm = icalcm;
    o.nloopn = nloopn;
    o.lhuber = c->opt.lhuber;
    o.lfitbb = c->opt.lfitbb;
    o.chicut = chicut;
    o.chhuge = c->opt.chhuge;
    o.dwcut = c->opt.dwcut;
    o.cndlmx = c->opt.cndlmx;
    o.esc[0] = c->opt.scale_errors[0];
    o.esc[1] = c->opt.scale_errors[1];
    RecOut out{c->d_chi2.p, c->d_ndf.p, c->d_status.p};
    DevModel M = model_of(c);
    DevRecords R = records_of(c);
    MonCtx mon{};
    if (c->opt.monitor != 0 && c->d_mon_hist.p) {
        mon = MonCtx{c->d_mon_grp.p, c->d_mon_res.p, c->d_mon_hist.p, c->mon_bins, c->opt.monitor};
        c->d_mon_hist.zero(s);   // histograms of this pass only (src/pede.f90:2607)
    }
    int grid = (int)std::min<int64_t>(c->nrec, 256 * 40);
    HIPCHK(hipEventRecord(e1, s));
    if (c->nrec > 0) {
        if (icalcm == 1 && c->opt.matsto == MPGPU_SPARSE) {
            // two-phase build: each batch of records writes its updates, then the blocks gather them
            for (size_t bt = 0; bt < c->batches.size(); bt++) {
                const auto &B = c->batches[bt];
                if (B.q1 <= B.q0) continue;
                const int g = (int)std::min<int64_t>(B.q1 - B.q0, 256 * 40);
                auto tr = timed(t_rec);
                HIPCHK(hipEventRecord(tr.first, s));
                HIPCHK(launch_records(R, M, o, out, c->plan, B.q0, B.q1, std::max(g, 1), s, mon));
                HIPCHK(hipEventRecord(tr.second, s));
                auto tg = timed(t_gat);
                HIPCHK(hipEventRecord(tg.first, s));
                HIPCHK(launch_block_gather(c->nblock, c->d_boff.p, c->d_brg.p, c->d_bcol.p, c->d_gfirst.p,
                                           c->d_inc_ptr.p + bt * (c->nblock + 1), c->d_inc_src.p + B.inc0, c->d_cbuf.p,
                                           c->d_mat.p, c->d_vec.p, s));
                HIPCHK(hipEventRecord(tg.second, s));
            }
        } else {
            auto tr = timed(t_rec);
            HIPCHK(hipEventRecord(tr.first, s));
            HIPCHK(launch_records(R, M, o, out, c->plan, 0, c->nrec, std::max(grid, 1), s, mon));
            HIPCHK(hipEventRecord(tr.second, s));
        }
    }
    if (c->nband > 0)
        HIPCHK(launch_records_band(R, M, o, out, c->bplan, c->d_band_ids.p, c->nband, c->d_band_info.p,
                                   c->d_band_scratch.p, c->band_grid, s));
    HIPCHK(hipEventRecord(e2, s));
    // pass statistics
    c->d_ndfsum.zero(s);
    c->d_hist.zero(s);
    l_pass_stats(c->nrec, c->d_chi2.p, c->d_ndf.p, c->d_status.p, c->d_part.p, c->d_ndfsum.p, c->d_hist.p, c->d_scal.p, s);
    double chi2 = 0.0;
    long long ndfsum = 0;
    unsigned long long hist[24];
    c->d_scal.download(&chi2, 1, s);
    c->d_ndfsum.download(&ndfsum, 1, s);
    c->d_hist.download(hist, 24, s);
    HIPCHK(hipStreamSynchronize(s));
    if (hist[9] > 0) throw MpError(30, "record exceeds record-kernel limits (" + std::to_string(hist[9]) + " records)");
    if (hist[8] > 0) throw MpError(30, "records left without a local fit (" + std::to_string(hist[8]) + ")");
    // pre-sigma diagonal (icalcm=1), regularisation and external measurements (loopn 2711-2798)
    std::vector<int32_t> ii, jj;
    std::vector<double> vv;
    if (icalcm == 1)
        for (int32_t v = 1; v <= c->nvar; v++)
            if (c->pre_weight[v - 1] != 0.0) { ii.push_back(v); jj.push_back(v); vv.push_back(c->pre_weight[v - 1]); }
    std::vector<double> vec_add(c->nall, 0.0);
    std::vector<unsigned long long> cnt_add(c->nvar, 0);
    double chi_extra = 0.0;
    int64_t ndf_extra = 0;
    if (c->opt.nregul != 0)
        for (int32_t v = 1; v <= c->nvar; v++) {
            int32_t t = c->var_to_tot[v - 1];
            vec_add[v - 1] -= c->par[t] * c->pre_weight[v - 1];
            chi_extra += c->pre_weight[v - 1] * c->par[t] * c->par[t];
        }
    for (auto &b : c->meas) {
        double weight = b.sigma > 0.0 ? 1.0 / (b.sigma * b.sigma) : 0.0;
        double dsum = -b.rhs;
        std::vector<int32_t> vix(b.labels.size());
        for (size_t q = 0; q < b.labels.size(); q++) {
            auto it = std::lower_bound(c->labels.begin(), c->labels.end(), b.labels[q]);
            int32_t t = (int32_t)(it - c->labels.begin());
            dsum += b.factors[q] * c->par[t];
            vix[q] = c->var_index[t];
        }
        for (size_t q = 0; q < b.labels.size(); q++) {
            if (b.factors[q] == 0.0) continue;
            if (vix[q] > 0) {
                vec_add[vix[q] - 1] -= weight * dsum * b.factors[q];
                cnt_add[vix[q] - 1] += 1;
                if (icalcm == 1)
                    for (size_t k = 0; k <= q; k++)
                        if (vix[k] > 0) { ii.push_back(vix[q]); jj.push_back(vix[k]); vv.push_back(weight * b.factors[q] * b.factors[k]); }
            }
        }
        chi_extra += weight * dsum * dsum;
        ndf_extra += 1;
    }
    // constraints (addcst, src/pede.f90:704-771): Lagrange rows + residuals
    for (int32_t j = 0; j < c->ncon; j++) {
        double rhs = c->crhs[j];
        for (auto &p : c->con_tot[j]) rhs = rhs - p.second * c->par[p.first];
        vec_add[c->nvar + j] = rhs;
        if (icalcm == 1 && c->opt.matsto != MPGPU_SPARSE)
            for (int64_t q = c->cptr[j]; q < c->cptr[j + 1]; q++) {
                ii.push_back(c->nvar + j + 1);
                jj.push_back(c->cidx[q]);
                vv.push_back(c->cval[q]);
            }
    }
    device_mupdat(c, ii, jj, vv);
    // fold host contributions into the device vector / counter
    std::vector<double> hv(c->nall);
    std::vector<unsigned long long> hc(c->nvar);
    c->d_vec.download(hv.data(), c->nall, s);
    c->d_cnt.download(hc.data(), c->nvar, s);
    HIPCHK(hipStreamSynchronize(s));
    for (int32_t i = 0; i < c->nall; i++) hv[i] += vec_add[i];
    for (int32_t i = 0; i < c->nvar; i++) hc[i] += cnt_add[i];
    c->d_vec.upload(hv.data(), c->nall, s);
    c->d_cnt.upload(hc.data(), c->nvar, s);
    c->last_vec = hv;
    c->last_cnt = hc;
    HIPCHK(hipStreamSynchronize(s));
    float k_ms = 0.f, t_ms = 0.f;
    HIPCHK(hipEventElapsedTime(&k_ms, e1, e2));
    HIPCHK(hipEventElapsedTime(&t_ms, e0, e2));
    auto sum_ms = [](const std::vector<std::pair<hipEvent_t, hipEvent_t>> &v) {
        double t = 0.0;
        for (const auto &p : v) {
            float m = 0.f;
            HIPCHK(hipEventElapsedTime(&m, p.first, p.second));
            t += m;
        }
        return t;
    };
    const double r_ms = sum_ms(t_rec), g_ms = sum_ms(t_gat);
    if (st) {
        std::memset(st, 0, sizeof(*st));
        st->chi2 = chi2 + chi_extra;
        st->ndf = ndfsum + ndf_extra;
        st->wndf = (double)st->ndf;
        for (int k = 0; k < 6; k++) st->nrej[k] = (int64_t)hist[k + 1];
        st->naccept = (int64_t)hist[0];
        st->nrec = c->nrec;
        st->ms_kernel = k_ms;
        st->ms_total = t_ms;
        st->ms_records = r_ms;
        st->n_records = (int64_t)t_rec.size();
        st->ms_gather = g_ms;
        st->n_gather = (int64_t)t_gat.size();
    }
    return 0;
    API_CATCH(c)
}

extern "C" int mpgpu_get_vector(const mpgpu_ctx *c, double *vec, int64_t *counter)
{
    if (!c || !c->built) return 30;
    if (vec) for (int32_t i = 0; i < c->nall; i++) vec[i] = i < (int32_t)c->last_vec.size() ? c->last_vec[i] : 0.0;
    if (counter) for (int32_t i = 0; i < c->nvar; i++) counter[i] = i < (int32_t)c->last_cnt.size() ? (int64_t)c->last_cnt[i] : 0;
    return 0;
}

extern "C" int mpgpu_get_matrix_packed(const mpgpu_ctx *c0, double *packed, int64_t len)
{
    mpgpu_ctx *c = const_cast<mpgpu_ctx *>(c0);
    if (!c || !c->built) return 30;
    API_TRY
    int64_t n = c->nall;
    if (len < n * (n + 1) / 2) throw MpError(24, "output too small");
    if (c->opt.matsto != MPGPU_SPARSE) {
        c->d_mat.download(packed, n * (n + 1) / 2, c->stream);
        HIPCHK(hipStreamSynchronize(c->stream));
        return 0;
    }
    host_blocks(c);
    std::vector<double> vals(c->nstored);
    c->d_mat.download(vals.data(), c->nstored, c->stream);
    HIPCHK(hipStreamSynchronize(c->stream));
    std::fill(packed, packed + n * (n + 1) / 2, 0.0);
    for (int64_t b = 0; b < c->nblock; b++) {
        int ga = c->brg[b], gb = c->bcol[b];
        int fa = c->gfirst[ga], fb = c->gfirst[gb];
        int nr = c->gfirst[ga + 1] - fa, nc = c->gfirst[gb + 1] - fb;
        for (int r = 0; r < nr; r++)
            for (int q = 0; q < nc; q++) {
                int64_t i = fa + r, j = fb + q;
                if (j > i) continue;
                packed[(i * i - i) / 2 + j - 1] = vals[c->boff[b] + (int64_t)r * nc + q];
            }
    }
    for (int32_t j = 0; j < c->ncon; j++)
        for (int64_t q = c->cptr[j]; q < c->cptr[j + 1]; q++) {
            int64_t i = c->nvar + j + 1, k = c->cidx[q];
            packed[(i * i - i) / 2 + k - 1] += c->cval[q];
        }
    return 0;
    API_CATCH(c)
}

extern "C" int mpgpu_mupdat(mpgpu_ctx *c, int32_t nel, const int32_t *i, const int32_t *j, const double *v)
{
    if (!c || !c->built) return 30;
    API_TRY
    std::vector<int32_t> ii(i, i + nel), jj(j, j + nel);
    std::vector<double> vv(v, v + nel);
    for (int32_t k = 0; k < nel; k++)
        if (std::max(ii[k], jj[k]) > c->nall) throw MpError(23, "Aborted, bad matrix index");
    device_mupdat(c, ii, jj, vv);
    return 0;
    API_CATCH(c)
}

// ---------------------------------------------------------------------------------------------
// solve
// ---------------------------------------------------------------------------------------------
namespace {

struct Work {
    DBuf<double> a, b, c2, d, e, f, g, h, x, tmp, part, scal;
};

void avprod_dev(mpgpu_ctx *c, const double *x, double *y)
{
    hipStream_t s = c->stream;
    if (c->opt.matsto == MPGPU_SPARSE) {
        l_block_symv(c->ngroup, c->d_brow.p, c->d_bcf.p, c->d_boff.p, c->d_gfirst.p, c->d_ccol_ptr.p, c->d_cc_off.p,
                     c->d_cc_fn.p, c->d_mat.p, x, y, c->nvar, c->nall, c->plan.max_grp, c->d_cc_blk.p, c->d_tb.p,
                     c->uniform_grp, c->d_bcol.p, c->nblock, c->spmv_variant, s);
        l_cons_symv(c->ncon, c->nvar, c->d_cptr.p, c->d_cidx.p, c->d_cval.p, x, y, s);
    } else {
        l_packed_symv(c->nall, c->d_mat.p, x, y, c->d_psymv.p, s);
    }
}

// the operator MINRES sees: A itself, or with eliminated constraints the leading nfit x nfit
// block of Q^T A Q (pede's avprod, src/pede.f90:5340-5358): x padded with zeros, Q x, A, Q^T
struct ElimWork {
    DBuf<double> e, f;
};
void op_dev(mpgpu_ctx *c, ElimWork &E, const double *x, double *y)
{
    if (!c->elim) { avprod_dev(c, x, y); return; }
    hipStream_t s = c->stream;
    const int32_t nv = c->nvar, nf = c->nfit;
    l_ql_pad(nf, nv, x, E.e.p, s);
    l_ql_apply(c->ql->dev, 0, E.e.p, s);
    avprod_dev(c, E.e.p, E.f.p);
    l_ql_apply(c->ql->dev, 1, E.f.p, s);
    l_copy(nf, E.f.p, y, s);
}

double dot_dev(mpgpu_ctx *c, int64_t n, const double *a, const double *b)
{
    l_dot(n, a, b, c->d_part.p, c->d_scal.p, c->stream);
    double r = 0.0;
    c->d_scal.download(&r, 1, c->stream);
    HIPCHK(hipStreamSynchronize(c->stream));
    return r;
}

// preconditioner state for MINRES
struct Precond {
    int kind = 0; // 0 none, 1 diagonal (+constraints), 2 band (+constraints)
    DBuf<double> cu, sm, scratch, band, kmat, h, Sf, Sb, tbuf, dtmp, Af, Ab;   // Af/Ab: scan levels of the carry
    DBuf<int> nrkd, india, hind, cb;
    DBuf<int64_t> koff;
    int n = 0, m = 0, w = 0, nlev = 0;
    bool spike = false;
};

// With eliminated constraints the preconditioner is built from pede's band of Q^T A Q (qlpssq,
// src/mpqldec.f90:711-945, called in mminrs, src/pede.f90:8294,8306): P = A V (one product per
// reflector), X = V^T P, the correction coefficients from X on the host (an m^3 recurrence),
// Y = 2 (P + V C^T), and the band (or diagonal) of A updated entry by entry.
void elim_band(mpgpu_ctx *c, int32_t mb, const int32_t *india, double *band)
{
    hipStream_t s = c->stream;
    const QlDev &q = c->ql->dev;
    const int32_t nv = c->nvar, m = q.m;
    DBuf<double> Vd, P, X, Cf, Sd;
    Vd.alloc((size_t)nv * m);
    P.alloc((size_t)nv * m);
    X.alloc((size_t)m * m);
    l_ql_vdense(q, Vd.p, s);
    for (int k = 0; k < m; k++) avprod_dev(c, Vd.p + (size_t)k * nv, P.p + (size_t)k * nv);
    l_ql_vtx_cols(q, P.p, nv, m, X.p, m, s);
    std::vector<double> hX((size_t)m * m), hC, hD;
    X.download(hX.data(), hX.size(), s);
    HIPCHK(hipStreamSynchronize(s));
    c->ql->band_coef(hX, hC, hD);
    std::vector<double> hS((size_t)m * m, 0.0);
    for (int k = 0; k < m; k++) hS[(size_t)k * m + k] = hD[k];
    Cf.upload(hC, s);
    Sd.upload(hS, s);
    l_ql_y(q, Cf.p, P.p, s);
    l_ql_band(q, c->nfit, mb, india, P.p, Sd.p, band, s);
    HIPCHK(hipStreamSynchronize(s));
}

void setup_precond(mpgpu_ctx *c, Precond &P)
{
    hipStream_t s = c->stream;
    P.n = c->elim ? c->nfit : c->nvar;
    P.m = c->ncon;
    DevModel M = model_of(c);
    if (c->opt.mbandw < 0) { P.kind = 0; return; }
    if (c->opt.mbandw == 0) {
        if (P.m > 1024) throw MpError(30, "diagonal preconditioner supports at most 1024 constraints");
        P.kind = 1;
        DBuf<double> diag;
        diag.alloc(P.n);
        l_get_diag(M, P.n, diag.p, s);
        if (c->elim) {
            std::vector<int32_t> one(P.n);
            for (int i = 0; i < P.n; i++) one[i] = i + 1;
            DBuf<int32_t> di;
            di.upload(one, s);
            elim_band(c, 1, di.p, diag.p);
        }
        P.cu.alloc(P.n);
        P.nrkd.alloc(1);
        P.sm.alloc(std::max(1, (P.m * P.m + P.m) / 2));
        P.scratch.alloc(P.n);
        l_precons(P.n, P.m, diag.p, P.cu.p, P.nrkd.p, c->d_cptr.p, c->d_cidx.p, c->d_cval.p, P.sm.p, P.scratch.p, s);
        HIPCHK(hipStreamSynchronize(s));
        return;
    }
    // band preconditioner (vmprep 7690-7700, equdecs)
    P.kind = 2;
    const int mb = c->opt.mbandw;
    if (mb > 64) throw MpError(30, "band preconditioner supports bandwidth <= 64");
    std::vector<int> india(P.n);
    for (int i = 1; i <= std::min(mb, P.n); i++) india[i - 1] = (i * i + i) / 2;
    for (int i = std::min(mb, P.n) + 1; i <= P.n; i++) india[i - 1] = india[i - 2] + mb;
    P.india.upload(india, s);
    P.band.alloc(india[P.n - 1]);
    l_get_band(M, P.n, mb, P.india.p, P.band.p, s);
    if (c->elim) elim_band(c, mb, P.india.p, P.band.p);
    P.nrkd.alloc(2);
    if (lltdec_parallel(P.n, mb) && !std::getenv("MPGPU_SEQUENTIAL_LLT")) {
        DBuf<double> L;
        DBuf<int8_t> bad;
        DBuf<int> flag;
        L.alloc((size_t)P.n * mb);
        bad.alloc(P.n);
        flag.alloc(3);
        l_lltdec_par(P.n, P.band.p, P.india.p, P.nrkd.p, c->opt.lprecm, mb, L.p, bad.p, flag.p, s);
    } else {
        l_lltdec(P.n, P.band.p, P.india.p, P.nrkd.p, c->opt.lprecm, mb, s);
    }
    if (P.m > 0) {
        // constraint blocks: merge overlapping variable ranges of constraints
        std::vector<int> first(P.m), last(P.m);
        for (int j = 0; j < P.m; j++) {
            first[j] = c->cidx[c->cptr[j]];
            last[j] = c->cidx[c->cptr[j + 1] - 1];
        }
        std::vector<int> cb(2 * P.m);
        std::vector<int64_t> koff(P.m + 1, 0);
        int j = 0;
        const int nextra = std::max(0, mb - 1);
        while (j < P.m) {
            int lo = first[j], hi = last[j], k = j + 1;
            while (k < P.m && first[k] <= hi) { hi = std::max(hi, last[k]); lo = std::min(lo, first[k]); k++; }
            for (int q = j; q < k; q++) { cb[2 * q] = lo - 1; cb[2 * q + 1] = hi - lo + 1 + nextra; }
            j = k;
        }
        for (int q = 0; q < P.m; q++) koff[q + 1] = koff[q] + cb[2 * q + 1];
        std::vector<double> kmat(koff[P.m], 0.0);
        for (int q = 0; q < P.m; q++)
            for (int64_t p = c->cptr[q]; p < c->cptr[q + 1]; p++) kmat[koff[q] + (c->cidx[p] - 1 - cb[2 * q])] += c->cval[p];
        P.cb.upload(cb, s);
        P.koff.upload(koff, s);
        P.kmat.upload(kmat, s);
        std::vector<int> hind(P.m);
        for (int i = 1; i <= P.m; i++) hind[i - 1] = (i * i + i) / 2;
        P.hind.upload(hind, s);
        P.h.alloc((P.m * P.m + P.m) / 2);
        l_band_k(P.m, P.n, P.band.p, P.india.p, P.cb.p, P.koff.p, P.kmat.p, mb, s);
        l_band_h(P.m, P.cb.p, P.koff.p, P.kmat.p, P.h.p, P.hind.p, P.nrkd.p + 1, s);
    }
    // chunk-parallel band solves (response matrices of chunks of rows)
    P.w = mb - 1;
    P.spike = P.w >= 1 && P.w <= 16;
    if (P.spike) {
        const int ch = spk_chunk_rows(), nch = (P.n + ch - 1) / ch;
        P.Sf.alloc((size_t)P.n * P.w);
        P.Sb.alloc((size_t)P.n * P.w);
        P.tbuf.alloc((size_t)nch * P.w);
        P.dtmp.alloc(std::max(1, P.m));
        DBuf<double> scr;
        scr.alloc((size_t)nch * P.w * ch);
        l_spk_build(P.n, P.w, P.band.p, P.india.p, P.Sf.p, P.Sb.p, scr.p, s);
        P.nlev = spk_scan_levels(P.n, P.w);
        if (P.nlev > 0) {
            const size_t lv = (size_t)nch * P.w * P.w * P.nlev;
            P.Af.alloc(lv);
            P.Ab.alloc(lv);
            l_spk_scan_build(P.n, P.w, P.Sf.p, P.Sb.p, P.Af.p, P.Ab.p, P.nlev, s);
        }
        HIPCHK(hipStreamSynchronize(s));
    }
    HIPCHK(hipStreamSynchronize(s));
}

void apply_precond(mpgpu_ctx *c, Precond &P, const double *x, double *y)
{
    hipStream_t s = c->stream;
    if (P.kind == 1) {
        l_presols(P.n, P.m, P.cu.p, c->d_cptr.p, c->d_cidx.p, c->d_cval.p, P.sm.p, x, y, s);
    } else if (P.kind == 2) {
        l_copy(P.n + P.m, x, y, s);
        if (P.spike)
            l_equslvs_spike(P.n, P.m, P.w, P.band.p, P.india.p, P.Sf.p, P.Sb.p, P.tbuf.p, P.cb.p, P.koff.p, P.kmat.p,
                            P.h.p, P.hind.p, P.dtmp.p, y, P.Af.p, P.Ab.p, P.nlev, s);
        else
            l_equslvs(P.n, P.m, P.band.p, P.india.p, P.cb.p, P.koff.p, P.kmat.p, P.h.p, P.hind.p, y, s);
    } else {
        l_copy(P.n + P.m, x, y, s);
    }
}

// the preconditioner of the current matrix, built on first use after the matrix changed
Precond &cached_precond(mpgpu_ctx *c)
{
    if (!c->pc || c->pc_bw != c->opt.mbandw || c->pc_lp != c->opt.lprecm) {
        c->pc.reset();
        auto P = std::make_shared<Precond>();
        setup_precond(c, *P);
        c->pc = P;
        c->pc_bw = c->opt.mbandw;
        c->pc_lp = c->opt.lprecm;
    }
    return *c->pc;
}

// MINRES, restating src/minresModule.f90:294-660 with device vectors and the scalar recurrence on the device
void minres_dev(mpgpu_ctx *c, const double *b_dev, double *x_dev, mpgpu_solve_stats *st)
{
    const int64_t n = c->elim ? c->nfit : c->nall;
    hipStream_t s = c->stream;
    ElimWork E;
    if (c->elim) { E.e.alloc(c->nvar); E.f.alloc(c->nvar); }
    Precond &P = cached_precond(c);
    const bool precon = P.kind != 0;
    DBuf<double> r1, r2, v, w, w1, w2, y;
    r1.alloc(n); r2.alloc(n); v.alloc(n); w.alloc(n); w1.alloc(n); w2.alloc(n); y.alloc(n);
    const double rtol = c->opt.mrestl;
    const int itnlim = c->opt.itnlim;
    int istop = 0, itn = 0;
    double Anorm = 0, Acond = 0, rnorm = 0, ynorm = 0, Arnorml = 0;
    HIPCHK(hipMemsetAsync(x_dev, 0, sizeof(double) * n, s));
    l_copy(n, b_dev, y.p, s);
    l_copy(n, b_dev, r1.p, s);
    if (precon) apply_precond(c, P, b_dev, y.p);
    double beta1 = dot_dev(c, n, b_dev, y.p);
    if (beta1 < 0.0) { istop = 8; goto done; }
    if (beta1 == 0.0) { istop = 0; goto done; }
    beta1 = std::sqrt(beta1);
    op_dev(c, E, y.p, w.p);
    Arnorml = std::sqrt(dot_dev(c, n, w.p, w.p));
    {
        // the iterations run with the scalar recurrence on the device (k_mr_step): no host round
        // trip per iteration; the host reads the stop code every few iterations
        MrState h{};
        h.beta1 = beta1;
        h.beta = beta1;
        h.oldb = 0.0;
        h.dbar = 0.0;
        h.epsln = 0.0;
        h.qrnorm = beta1;
        h.phibar = beta1;
        h.rhs1 = beta1;
        h.rhs2 = 0.0;
        h.tnorm2 = 0.0;
        h.ynorm2 = 0.0;
        h.cs = -1.0;
        h.sn = 0.0;
        h.Arnorml = Arnorml;
        h.rtol = rtol;
        h.itnlim = itnlim;
        DBuf<MrState> dst;
        dst.upload(&h, 1, s);
        MrState *d = dst.p;
        w.zero(s);
        w2.zero(s);
        l_copy(n, r1.p, r2.p, s);
        auto iteration = [&](int k) {
            l_mr_scale(n, d, y.p, v.p, s);
            op_dev(c, E, v.p, y.p);
            if (k >= 2) l_mr_axpy(n, d, 0, r1.p, y.p, s);
            l_dot(n, v.p, y.p, c->d_part.p, &d->alfa, s);
            l_mr_axpy(n, d, 1, r2.p, y.p, s);
            std::swap(r1.p, r2.p);          // r1 = r2 (old r2 buffer becomes r1)
            l_copy(n, y.p, r2.p, s);        // r2 = y
            if (precon) apply_precond(c, P, r2.p, y.p);
            l_dot(n, r2.p, y.p, c->d_part.p, &d->bnew, s);
            l_mr_step(d, s);
            l_mr_update(n, d, v.p, w.p, w1.p, w2.p, x_dev, s);
        };
        auto stopped = [&]() {
            dst.download(&h, 1, s);
            HIPCHK(hipStreamSynchronize(s));
            return h.istop != 0;
        };
        // the first iteration eagerly; then blocks of `check` iterations (an even number, so the
        // r1/r2 roles return to the same buffers) captured once into a HIP graph and replayed:
        // one launch per block instead of some twenty kernel launches per iteration
        const int check = 8;
        iteration(1);
        int k = 1;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        if (!stopped() && k < itnlim) {
            // MPGPU_NO_GRAPH: launch every iteration's kernels directly (for tools that do not
            // follow graph launches; rocprofv3's kernel trace has crashed inside hipGraphLaunch)
            static const bool no_graph = std::getenv("MPGPU_NO_GRAPH") != nullptr;
            bool graphed = !no_graph && hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) == hipSuccess;
            if (graphed) {
                for (int q = 1; q <= check; q++) iteration(k + q);
                graphed = hipStreamEndCapture(s, &graph) == hipSuccess && graph != nullptr &&
                          hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess;
                if (!graphed) (void)hipGetLastError();
            }
            for (;;) {
                if (graphed) HIPCHK(hipGraphLaunch(exec, s));
                else
                    for (int q = 1; q <= check; q++) iteration(k + q);
                k += check;
                if (stopped() || k >= itnlim + check) break;
            }
        }
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        istop = h.istop;
        itn = h.itn;
        Anorm = h.Anorm;
        Acond = h.Acond;
        rnorm = h.rnorm;
        ynorm = h.ynorm;
        Arnorml = h.Arnorml;
    }
done:
    HIPCHK(hipStreamSynchronize(s));
    if (st) {
        st->istop = istop;
        st->itn = itn;
        st->anorm = Anorm;
        st->acond = Acond;
        st->rnorm = rnorm;
        st->arnorm = Arnorml;
        st->ynorm = ynorm;
    }
}

// Givens reflection of MINRES-QLP (symortho, src/minresqlpModule.f90:1337-1380)
static void symortho(double a, double b, double &c, double &s, double &r)
{
    const double aa = std::fabs(a), ab = std::fabs(b);
    if (ab <= DBL_MIN) { s = 0.0; r = aa; c = (a == 0.0) ? 1.0 : a / aa; }
    else if (aa <= DBL_MIN) { c = 0.0; r = ab; s = b / ab; }
    else if (ab > aa) { double t = a / b; s = (b / ab) / std::sqrt(1.0 + t * t); c = s * t; r = b / s; }
    else { double t = b / a; c = (a / aa) / std::sqrt(1.0 + t * t); s = c * t; r = a / c; }
}

static double hyp(double a, double b, double c = 0.0) { return std::sqrt(a * a + b * b + c * c); }

// MINRES-QLP, restating src/minresqlpModule.f90:411-1330 as mminrsqlp calls it
// (src/pede.f90:8341-8437): rtol = mrestl, maxxnorm = n/sqrt(eps), trancond from mresmode
// (1: QR only, 2: QLP only, else mrestranscond), Acondlim and shift absent, symmetry checks on.
// Lanczos vectors live in HBM; the recurrences run on the host in FP64.
void minresqlp_dev(mpgpu_ctx *c, const double *b_dev, double *x_dev, mpgpu_solve_stats *st)
{
    const int64_t n = c->elim ? c->nfit : c->nall;
    hipStream_t s = c->stream;
    ElimWork E;
    if (c->elim) { E.e.alloc(c->nvar); E.f.alloc(c->nvar); }
    Precond &P = cached_precond(c);
    const bool precon = P.kind != 0;
    DBuf<double> r1, r2, v, w, wl, wl2, xl2, y;
    r1.alloc(n); r2.alloc(n); v.alloc(n); w.alloc(n); wl.alloc(n); wl2.alloc(n); xl2.alloc(n); y.alloc(n);
    const double eps = 2.220446049250313e-16, one = 1.0, zero = 0.0;
    const double EPSINV = std::pow(10.0, std::floor(std::log10(one / eps)));
    const double NORMMAX = std::pow(10.0, std::floor(std::log10(one / eps) / 2));
    const double rtol = c->opt.mrestl;
    const int itnlim = c->opt.itnlim;
    double trancond = c->opt.mrmode == 1 ? one / eps : (c->opt.mrmode == 2 ? one : c->opt.mrtcnd);
    trancond = std::min(trancond, NORMMAX);
    const double maxxnorm = std::min((double)n / std::sqrt(eps), one / eps);
    const double Acondlim = EPSINV, ieps = 0.1 / eps;
    auto nrm = [&](const double *a) { return std::sqrt(dot_dev(c, n, a, a)); };

    int istop = 0, itn = 0, QLPiter = 0;
    bool lastiter = false;
    double t1 = 0, t2 = 0, gminl = 0, gminl2 = 0, sc, t, z, epsa;
    double xnorm = 0, xl2norm = 0, Axnorm = 0, Anorm = 0, Acond = 1, pnorm = 0, relresl = 0, relAresl = 0, Arnorm = 0;
    HIPCHK(hipMemsetAsync(x_dev, 0, sizeof(double) * n, s));
    xl2.zero(s);
    l_copy(n, b_dev, y.p, s);
    l_copy(n, b_dev, r1.p, s);
    if (precon) apply_precond(c, P, b_dev, y.p);
    double beta1 = dot_dev(c, n, b_dev, y.p);
    if (beta1 < zero && nrm(y.p) > eps) istop = 11;
    if (beta1 == zero) istop = 3;
    beta1 = std::sqrt(beta1);
    if (precon) {
        apply_precond(c, P, y.p, r2.p);
        sc = dot_dev(c, n, y.p, y.p);
        t = dot_dev(c, n, r1.p, r2.p);
        z = std::fabs(sc - t);
        epsa = (std::fabs(sc) + eps) * std::pow(eps, 0.33333);
        if (z > epsa) istop = 10;
    }
    op_dev(c, E, y.p, w.p);
    op_dev(c, E, w.p, r2.p);
    sc = dot_dev(c, n, w.p, w.p);
    t = dot_dev(c, n, y.p, r2.p);
    z = std::fabs(sc - t);
    epsa = (std::fabs(sc) + eps) * std::pow(eps, 0.33333);
    if (z > epsa) istop = 9;

    double tau = 0, taul = 0, taul2 = 0, gmin = 0, beta = 0, betal = 0, betan = beta1, dbar = 0, dltan = 0, eta = 0,
           etal = 0, etal2 = 0, vepln = 0, veplnl = 0, veplnl2 = 0, eplnn = 0, epln = 0, gama = 0, gamal = 0, gamal2 = 0,
           gamal3 = 0, phi = beta1, cr1 = -1, sr1 = 0, cr2 = -1, sr2 = 0, cs = -1, sn = 0, ul4 = 0, ul3 = 0, ul2 = 0, ul = 0,
           u = 0, gamal_QLP = 0, vepln_QLP = 0, gama_QLP = 0, ul_QLP = 0, u_QLP = 0;
    double rnorm = betan, relres = rnorm / (Anorm * xnorm + beta1);
    l_copy(n, b_dev, r2.p, s);
    w.zero(s);
    wl.zero(s);
    while (istop <= 0) {
        itn++;
        betal = beta;
        beta = betan;
        l_scale_copy(n, one / beta, y.p, v.p, s);
        op_dev(c, E, v.p, y.p);
        if (itn >= 2) l_axpy_neg(n, beta / betal, r1.p, y.p, s);
        double alfa = dot_dev(c, n, v.p, y.p);
        l_axpy_neg(n, alfa / beta, r2.p, y.p, s);
        std::swap(r1.p, r2.p);          // r1 = r2
        l_copy(n, y.p, r2.p, s);        // r2 = y
        if (!precon) betan = nrm(y.p);
        else {
            apply_precond(c, P, r2.p, y.p);
            betan = dot_dev(c, n, r2.p, y.p);
            if (betan > zero) betan = std::sqrt(betan);
            else if (nrm(y.p) > eps) { istop = 11; break; }
        }
        pnorm = itn == 1 ? hyp(alfa, betan) : hyp(beta, alfa, betan);

        dbar = dltan;
        double dlta = cs * dbar + sn * alfa;
        epln = eplnn;
        double gbar = sn * dbar - cs * alfa;
        eplnn = sn * betan;
        dltan = -cs * betan;
        const double dlta_QLP = dlta;
        gamal3 = gamal2;
        gamal2 = gamal;
        gamal = gama;
        symortho(gbar, betan, cs, sn, gama);
        const double gama_tmp = gama;
        taul2 = taul;
        taul = tau;
        tau = cs * phi;
        phi = sn * phi;
        Axnorm = std::sqrt(Axnorm * Axnorm + tau * tau);
        if (itn > 2) {
            veplnl2 = veplnl;
            etal2 = etal;
            etal = eta;
            double dlta_tmp = sr2 * vepln - cr2 * dlta;
            veplnl = cr2 * vepln + sr2 * dlta;
            dlta = dlta_tmp;
            eta = sr2 * gama;
            gama = -cr2 * gama;
        }
        if (itn > 1) {
            double gtmp;
            symortho(gamal, dlta, cr1, sr1, gtmp);
            gamal = gtmp;
            vepln = sr1 * gama;
            gama = -cr1 * gama;
        }
        ul4 = ul3;
        ul3 = ul2;
        if (itn > 2) ul2 = (taul2 - etal2 * ul4 - veplnl2 * ul3) / gamal2;
        if (itn > 1) ul = (taul - etal * ul3 - veplnl * ul2) / gamal;
        const double xnorm_tmp = hyp(xl2norm, ul2, ul);
        if (std::fabs(gama) > eps) {
            u = (tau - eta * ul2 - vepln * ul) / gama;
            if (relAresl < relresl && hyp(xnorm_tmp, u) > maxxnorm) { u = zero; istop = 12; }
        } else {
            u = zero;
            istop = 14;
        }
        xl2norm = hyp(xl2norm, ul2);
        xnorm = hyp(xl2norm, ul, u);
        if (Acond < trancond && istop == 0 && QLPiter == 0) {          // MINRES updates
            const bool upd_x = xnorm < maxxnorm;
            l_qlp_minres(n, v.p, w.p, wl.p, wl2.p, x_dev, epln, dlta_QLP, gama_tmp > eps ? one / gama_tmp : 0.0, tau,
                         gama_tmp > eps, upd_x, s);
            if (!upd_x) { istop = 12; lastiter = true; }
        } else {                                                       // MINRES-QLP updates
            QLPiter++;
            QlpStep q{};
            q.first = QLPiter == 1;
            q.itn = itn;
            q.gamal3 = gamal3; q.veplnl2 = veplnl2; q.etal = etal; q.gamal_qlp = gamal_QLP; q.vepln_qlp = vepln_QLP;
            q.gama_qlp = gama_QLP; q.ul_qlp = ul_QLP; q.u_qlp = u_QLP; q.sr1 = sr1; q.cr1 = cr1; q.sr2 = sr2; q.cr2 = cr2;
            q.ul2 = ul2; q.ul = ul; q.u = u;
            l_qlp_step(n, v.p, w.p, wl.p, wl2.p, x_dev, xl2.p, q, s);
        }
        double gtmp = gamal;
        symortho(gtmp, eplnn, cr2, sr2, gamal);
        gamal_QLP = gtmp;
        vepln_QLP = vepln;
        gama_QLP = gama;
        ul_QLP = ul;
        u_QLP = u;
        const double abs_gama = std::fabs(gama);
        Anorm = std::max(std::max(Anorm, pnorm), std::max(gamal, abs_gama));
        if (itn == 1) { gmin = gama; gminl = gmin; }
        else { gminl2 = gminl; gminl = gmin; gmin = std::min(std::min(gminl2, gamal), abs_gama); }
        const double Acondl = Acond;
        Acond = Anorm / gmin;
        const double rnorml = rnorm;
        relresl = relres;
        if (istop != 14) rnorm = phi;
        relres = rnorm / (Anorm * xnorm + beta1);
        const double rootl = hyp(gbar, dltan);
        relAresl = rootl / Anorm;
        const double epsx = Anorm * xnorm * eps;
        if (istop == 0 || istop == 14) { t1 = one + relres; t2 = one + relAresl; }
        if (t1 <= one) istop = 5;
        else if (t2 <= one) istop = 7;
        else if (relres <= rtol) istop = 4;
        else if (relAresl <= rtol) istop = 6;
        else if (epsx >= beta1) istop = 2;
        else if (xnorm >= maxxnorm) istop = 12;
        else if (Acond >= Acondlim || Acond >= ieps) istop = 13;
        else if (itn >= itnlim) istop = 8;
        else if (betan < eps) istop = 1;
        if (istop != 0) {
            if (istop == 6 || istop == 7 || istop == 12 || istop == 13) lastiter = true;
            if (lastiter) { itn--; Acond = Acondl; rnorm = rnorml; relres = relresl; }
            op_dev(c, E, x_dev, r1.p);
            l_sub(n, b_dev, r1.p, r1.p, s);
            op_dev(c, E, r1.p, wl2.p);
            Arnorm = nrm(wl2.p);
            break;
        }
    }
    HIPCHK(hipStreamSynchronize(s));
    if (st) {
        st->istop = istop;
        st->itn = itn;
        st->anorm = Anorm;
        st->acond = Acond;
        st->rnorm = rnorm;
        st->arnorm = Arnorm;
        st->ynorm = xnorm;
    }
}

double rel_residual(mpgpu_ctx *c, const double *b_dev, const double *x_dev)
{
    const int64_t n = c->nall;
    DBuf<double> ax, r;
    ax.alloc(n);
    r.alloc(n);
    avprod_dev(c, x_dev, ax.p);
    l_sub(n, b_dev, ax.p, r.p, c->stream);
    double rr = dot_dev(c, n, r.p, r.p), bb = dot_dev(c, n, b_dev, b_dev);
    return bb > 0 ? std::sqrt(rr / bb) : std::sqrt(rr);
}

// MINRES / MINRES-QLP with eliminated constraints (mminrs, src/pede.f90:8270-8290 and
// 8318-8326; mminrsqlp 8370-8430): y = L^-T d from the constraint residuals d = rhs - C p (addcst,
// src/pede.f90:730-761, in the QL order), the reduced right-hand side is the leading part of
// Q^T (b - A Q [0; y]), and the solution is Q [x_fit; y].  rel_residual is that of the reduced
// system.
void elim_solve(mpgpu_ctx *c, const double *b_dev, double *x_dev, mpgpu_solve_stats *st, bool qlp)
{
    hipStream_t s = c->stream;
    const QlElim &q = *c->ql;
    const int32_t nv = c->nvar, nf = c->nfit, m = q.ncon;
    std::vector<double> d(m);
    for (int k = 0; k < m; k++) {
        const int j = q.order[k];
        double r = c->crhs[j];
        for (auto &p : c->con_tot[j]) r = r - p.second * c->par[p.first];
        d[k] = r;
    }
    const std::vector<double> yc = q.bsub(d);
    std::vector<double> he(nv, 0.0);
    for (int k = 0; k < m; k++) he[nf + k] = yc[k];
    DBuf<double> xe, bav, bred, xf;
    xe.upload(he, s);
    bav.alloc(nv);
    bred.alloc(nv);
    xf.alloc(nf);
    l_ql_apply(q.dev, 0, xe.p, s);
    avprod_dev(c, xe.p, bav.p);
    l_sub(nv, b_dev, bav.p, bred.p, s);
    l_ql_apply(q.dev, 1, bred.p, s);
    if (qlp) minresqlp_dev(c, bred.p, xf.p, st);
    else minres_dev(c, bred.p, xf.p, st);
    {
        ElimWork E;
        E.e.alloc(nv);
        E.f.alloc(nv);
        DBuf<double> ax, r;
        ax.alloc(nf);
        r.alloc(nf);
        op_dev(c, E, xf.p, ax.p);
        l_sub(nf, bred.p, ax.p, r.p, s);
        const double rr = dot_dev(c, nf, r.p, r.p), bb = dot_dev(c, nf, bred.p, bred.p);
        st->rel_residual = bb > 0 ? std::sqrt(rr / bb) : std::sqrt(rr);
    }
    xe.upload(he, s);
    l_copy(nf, xf.p, xe.p, s);
    l_ql_apply(q.dev, 0, xe.p, s);
    l_copy(nv, xe.p, x_dev, s);
    HIPCHK(hipStreamSynchronize(s));
}

} // namespace

// ---------------------------------------------------------------------------------------------
// dense direct solves: 'decomposition' (chdec2/chslv2 semantics), the LAPACK methods and the
// FP32 factor with FP64 refinement
// ---------------------------------------------------------------------------------------------
// Factor of the (augmented) parameter block for the dense methods: FP64 LDL^T with chdec2's
// semantics ('decomposition'), LAPACK's stop on the first non-positive pivot (fullLAPACK,
// unpackedLAPACK), or the FP32 LDL^T for the FP32 + FP64-refinement method (falls back to FP64
// when an FP32 pivot is not positive).
static void dense_factor(mpgpu_ctx *c, DenseFactor &F, bool want_fp32)
{
    hipStream_t s = c->stream;
    const int nv = c->nvar, m = c->ncon;
    const bool lapack = c->opt.metsol == MPGPU_LAPACK_PACKED || c->opt.metsol == MPGPU_LAPACK_UNPACKED;
    F.n = nv;
    F.metsol = c->opt.metsol;
    // pede's decomposition (chdec2) eliminates from the last parameter; LAPACK's dpptrf from the
    // first, and it stops at the first non-positive leading minor
    F.rev = !lapack;
    // With constraints the parameter block alone may be singular (modes fixed only by the
    // constraints, as pede's decomposition removes them by elimination).  The block is
    // augmented to A + rho C^T C and the right-hand side to b + rho C^T d: on the constraint
    // surface C x = d this changes neither x nor the multipliers, and the augmented block is
    // positive definite whenever the constrained problem is well posed.
    double cc_norm = 0.0;
    F.rho = 0.0;
    std::vector<double> hd(nv), cc(nv, 0.0);
    {
        DBuf<double> dg;
        dg.alloc(nv);
        l_get_diag(model_of(c), nv, dg.p, s);
        dg.download(hd.data(), nv, s);
        HIPCHK(hipStreamSynchronize(s));
    }
    if (m > 0) {
        for (int j = 0; j < m; j++) {
            double rn = 0.0;
            for (int64_t q = c->cptr[j]; q < c->cptr[j + 1]; q++) {
                cc[c->cidx[q] - 1] += c->cval[q] * c->cval[q];
                rn += c->cval[q] * c->cval[q];
            }
            cc_norm += rn;   // ||c_j c_j^T||_F = ||c_j||^2
        }
        double amax = 0.0, cmax = 0.0;
        for (int i = 0; i < nv; i++) { amax = std::max(amax, std::fabs(hd[i])); cmax = std::max(cmax, cc[i]); }
        F.rho = cmax > 0.0 ? (amax > 0.0 ? amax : 1.0) / cmax : 0.0;
    }
    const int64_t nn = (int64_t)nv * nv;
    auto pivots_to_host = [&](bool f32) {
        DBuf<double> d;
        d.alloc(nv);
        if (f32) ldl_pivots32(nv, F.a32.p, nv, d.p, s); else ldl_pivots(nv, F.a.p, nv, d.p, s);
        F.piv.assign(nv, 0.0);
        d.download(F.piv.data(), nv, s);
        HIPCHK(hipStreamSynchronize(s));
    };
    if (want_fp32) {
        // FP32 keeps 24 bits relative to each element: the matrix is scaled to unit diagonal
        // first, so that parameters of very different scale (shifts, rotations) lose the same
        // relative precision in the rounding
        std::vector<double> d(nv);
        for (int i = 0; i < nv; i++) {
            const double a = hd[i] + F.rho * cc[i];
            d[i] = a > 0.0 ? 1.0 / std::sqrt(a) : 1.0;
        }
        F.sc.upload(d, s);
        F.fp32 = true;
        F.a32.alloc(nn);
        F.w32.alloc(nv);
        F.work32.alloc(ldl_work_elems(nv));
        l_packed_to_full32(nv, c->d_mat.p, F.a32.p, nv, F.rev, F.sc.p, s);
        l_add_cct32(m, c->d_cptr.p, c->d_cidx.p, c->d_cval.p, F.rho, F.a32.p, nv, nv, F.rev, F.sc.p, s);
        HIPCHK(ldl_factor32(nv, F.a32.p, nv, F.w32.p, F.work32.p, s));
        pivots_to_host(true);
        bool ok = true;
        double evmax = 0.0, evmin = DBL_MAX;
        for (int k = 0; k < nv; k++) {
            const double d = F.piv[k];
            if (!(d > 0.0) || !std::isfinite(d)) { ok = false; break; }
            evmax = std::max(evmax, d);
            evmin = std::min(evmin, d);
        }
        if (ok) {
            F.nrank = nv;
            F.evmax = evmax;
            F.evmin = evmin;
            F.anorm = std::sqrt(l_packed_fro2(nv, c->d_mat.p, c->d_part.p, c->d_scal.p, s)) + F.rho * cc_norm;
            return;
        }
        // the FP32 factor is not positive definite: it cannot precondition CG
    }
    F.fp32 = false;
    F.a32.release(); F.w32.release(); F.work32.release();
    F.a.alloc(nn);
    F.w.alloc(nv);
    F.work.alloc(ldl_work_elems(nv));
    l_packed_to_full(nv, c->d_mat.p, F.a.p, nv, F.rev, s);
    l_add_cct(m, c->d_cptr.p, c->d_cidx.p, c->d_cval.p, F.rho, F.a.p, nv, nv, F.rev, s);
    HIPCHK(ldl_factor(nv, F.a.p, nv, F.w.p, F.work.p, s));
    pivots_to_host(false);
    int nrank = 0;
    double evmax = 0.0, evmin = 0.0;
    for (int k = 0; k < nv; k++) {
        const double d = F.piv[k];
        if (d > 0.0) {
            if (nrank == 0) evmax = evmin = d;
            evmax = std::max(evmax, d);
            evmin = std::min(evmin, d);
            nrank++;
        } else if (lapack) {
            // dpptrf: 'factorization ... failed at index', exit 29 (src/mplapack.f90:144-156)
            throw MpError(29, "Aborted, factorization of global matrix failed at index " + std::to_string(k + 1));
        }
    }
    F.nrank = nrank;
    F.evmax = evmax;
    F.evmin = evmin;
}

static void dense_solve(mpgpu_ctx *c, const double *b_dev, double *x_dev, mpgpu_solve_stats &loc)
{
    hipStream_t s = c->stream;
    const int64_t n = c->nall;
    const int nv = c->nvar, m = c->ncon;
    const bool lapack = c->opt.metsol == MPGPU_LAPACK_PACKED || c->opt.metsol == MPGPU_LAPACK_UNPACKED;
    // 'LAPACKwitherrors': the parameter covariance on the constraint surface, as pede's
    // inverse from the factor expanded by Q (src/pede.f90:9017-9120); with multipliers the
    // same matrix is diag(A^-1) - diag(W S^-1 W^T), W = A^-1 C^T, S = C W, and A may carry
    // the rho C^T C term (it cancels on the constraint surface)
    const bool want_cov = c->opt.lapack_errors && lapack;
    c->have_cov = false;
    std::vector<double> covcorr(want_cov ? nv : 0, 0.0);
    // The FP32 factor with refinement pays only for large blocks: below fp32_min_n parameters the
    // FP64 factor is the faster one on this GPU (measured: 3-16 ms against 20-140 ms for the
    // FP32 factor and its refinement steps at 500-5,000 parameters; 1.16x faster at 30,000)
    const bool by_size = c->opt.metsol == MPGPU_CHOLESKY_FP32_IR && nv < c->opt.fp32_min_n;
    if (!c->dfac || c->dfac->metsol != c->opt.metsol) {
        c->dfac.reset(new DenseFactor());
        dense_factor(c, *c->dfac, c->opt.metsol == MPGPU_CHOLESKY_FP32_IR && !by_size);
    }
    DenseFactor &F = *c->dfac;
    if (!F.fp32 && c->opt.metsol == MPGPU_CHOLESKY_FP32_IR) { loc.method = MPGPU_DECOMPOSITION; loc.istop = by_size ? 4 : 3; }
    loc.nrank = F.nrank;
    loc.evmax = F.evmax;
    loc.evmin = F.evmin;
    const double rho = F.rho;
    const double eps = std::numeric_limits<double>::epsilon();
    const double tol = c->opt.refine_tol > 0.0 ? c->opt.refine_tol : eps * std::sqrt((double)nv);
    const bool verbose = std::getenv("MPGPU_VERBOSE") != nullptr;

    // FP32-factor-preconditioned conjugate gradients on the (augmented) parameter block, nrhs
    // columns side by side with their own recurrences.  A column has converged when
    // ||r|| <= tol ||x|| ||A||_F (the backward-error test of LAPACK's mixed-precision Cholesky
    // solver DSPOSV, in 2-norms).  Returns false when a column did not converge within
    // refine_steps steps.
    auto refine = [&](double *v, int nrhs) -> bool {
        const int64_t len = (int64_t)nv * nrhs;
        DBuf<double> X, R, Z, P, Q, sc;
        X.alloc(len); R.alloc(len); Z.alloc(len); P.alloc(len); Q.alloc(len);
        sc.alloc(4 * (size_t)nrhs);
        std::vector<double> bb(nrhs), rz(nrhs), pq(nrhs), rr(nrhs), xx(nrhs), coef(nrhs);
        std::vector<char> active(nrhs, 1), conv(nrhs, 0);
        std::vector<double> hist;
        auto dots = [&](const double *A, const double *B, std::vector<double> &out) {
            l_col_dots(nv, nrhs, A, B, sc.p, s);
            sc.download(out.data(), nrhs, s);
            HIPCHK(hipStreamSynchronize(s));
        };
        l_zero(len, X.p, s);
        l_copy(len, v, R.p, s);
        dots(R.p, R.p, bb);
        for (int q = 0; q < nrhs; q++) if (!(bb[q] > 0.0)) { active[q] = 0; conv[q] = 1; }
        l_copy(len, R.p, Z.p, s);
        F.solve32(Z.p, nrhs, s);
        l_copy(len, Z.p, P.p, s);
        dots(R.p, Z.p, rz);
        const int maxit = std::max(1, c->opt.refine_steps);
        for (int it = 1; it <= maxit; it++) {
            int nact = 0;
            for (int q = 0; q < nrhs; q++) {
                if (!active[q]) continue;
                nact++;
                l_packed_symv(nv, c->d_mat.p, P.p + (int64_t)q * nv, Q.p + (int64_t)q * nv, c->d_psymv.p, s);
                l_cct_apply(m, c->d_cptr.p, c->d_cidx.p, c->d_cval.p, rho, P.p + (int64_t)q * nv, Q.p + (int64_t)q * nv, s);
            }
            if (nact == 0) break;
            dots(P.p, Q.p, pq);
            for (int q = 0; q < nrhs; q++) {
                coef[q] = 0.0;
                if (!active[q]) continue;
                if (!(pq[q] > 0.0)) { active[q] = 0; continue; }   // breakdown: not converged
                coef[q] = rz[q] / pq[q];
            }
            sc.upload(coef.data(), nrhs, s);
            l_col_axpy(nv, nrhs, sc.p, 1.0, P.p, X.p, s);     // x += alpha p
            l_col_axpy(nv, nrhs, sc.p, -1.0, Q.p, R.p, s);    // r -= alpha q
            loc.itn = std::max(loc.itn, it);
            dots(R.p, R.p, rr);
            dots(X.p, X.p, xx);
            int left = 0;
            double worst = 0.0;     // largest backward error of the columns still running
            for (int q = 0; q < nrhs; q++) {
                if (!active[q]) continue;
                const double be = std::sqrt(rr[q]) / (F.anorm * std::sqrt(xx[q]));
                if (verbose && q == 0)
                    std::fprintf(stderr, "mpgpu: refinement step %d: ||r|| / (||A||_F ||x||) = %.3e (tol %.3e), ||r||/||b|| = %.3e\n",
                                 it, be, tol, std::sqrt(rr[q] / bb[q]));
                if (be <= tol) { active[q] = 0; conv[q] = 1; }
                else { left++; worst = std::max(worst, be); }
            }
            if (left == 0 || it == maxit) break;
            // give up early when the rate of the last steps cannot reach tol within refine_steps
            // (an ill-conditioned matrix: the FP64 factor is then cheaper than the steps left)
            hist.push_back(worst);
            const int span = 4;
            if ((int)hist.size() > span) {
                const double rate = std::pow(hist.back() / hist[hist.size() - 1 - span], 1.0 / span);
                const double need = rate < 1.0 ? std::log(tol / hist.back()) / std::log(rate) : 1e30;
                if (it + need > maxit) break;
            }
            l_copy(len, R.p, Z.p, s);
            F.solve32(Z.p, nrhs, s);
            std::vector<double> rz2(nrhs);
            dots(R.p, Z.p, rz2);
            for (int q = 0; q < nrhs; q++) coef[q] = active[q] && rz[q] != 0.0 ? rz2[q] / rz[q] : 0.0;
            sc.upload(coef.data(), nrhs, s);
            l_col_xpby(nv, nrhs, Z.p, sc.p, P.p, s);           // p = z + beta p
            rz = rz2;
        }
        l_copy(len, X.p, v, s);
        for (int q = 0; q < nrhs; q++) if (!conv[q]) return false;
        return true;
    };
    // A v = rhs for nrhs columns: FP64 factor, or FP32-preconditioned CG with the FP64 factor
    // as the fallback when it does not converge (the FP64 factor then replaces the kept one)
    auto solveA = [&](double *v, int nrhs) {
        if (!F.fp32) { F.solve64(v, nrhs, s); return; }
        DBuf<double> keep;
        keep.alloc((size_t)nv * nrhs);
        l_copy((int64_t)nv * nrhs, v, keep.p, s);
        if (refine(v, nrhs)) { loc.istop = 1; return; }
        loc.method = MPGPU_DECOMPOSITION;
        loc.istop = 3;
        dense_factor(c, F, false);
        loc.nrank = F.nrank;
        loc.evmax = F.evmax;
        loc.evmin = F.evmin;
        l_copy((int64_t)nv * nrhs, keep.p, v, s);
        F.solve64(v, nrhs, s);
    };

    HIPCHK(hipMemcpyAsync(x_dev, b_dev, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    if (m == 0) {
        solveA(x_dev, 1);
    } else {
        // Z = A^-1 C^T (m columns), S = C Z, lambda = S^-1 (C A^-1 b_p - b_c), x = A^-1 (b_p - C^T lambda)
        std::vector<double> bp(n);
        HIPCHK(hipMemcpyAsync(bp.data(), b_dev, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        // right-hand sides: the m columns of C^T and b + rho C^T d, solved together
        std::vector<double> Z((size_t)(m + 1) * nv, 0.0);
        for (int j = 0; j < m; j++)
            for (int64_t q = c->cptr[j]; q < c->cptr[j + 1]; q++) Z[(size_t)j * nv + c->cidx[q] - 1] += c->cval[q];
        double *bt = Z.data() + (size_t)m * nv;
        for (int i = 0; i < nv; i++) bt[i] = bp[i];
        for (int j = 0; j < m; j++)
            for (int64_t q = c->cptr[j]; q < c->cptr[j + 1]; q++) bt[c->cidx[q] - 1] += rho * c->cval[q] * bp[nv + j];
        DBuf<double> zc;
        zc.upload(Z, s);
        solveA(zc.p, m + 1);
        zc.download(Z.data(), Z.size(), s);
        HIPCHK(hipStreamSynchronize(s));
        const double *y = Z.data() + (size_t)m * nv;
        // S (m x m) and rhs
        std::vector<double> S((size_t)m * m), g(m);
        for (int i = 0; i < m; i++) {
            for (int j = 0; j < m; j++) {
                double acc = 0.0;
                for (int64_t q = c->cptr[i]; q < c->cptr[i + 1]; q++) acc += c->cval[q] * Z[(size_t)j * nv + c->cidx[q] - 1];
                S[(size_t)i * m + j] = acc;
            }
            double acc = 0.0;
            for (int64_t q = c->cptr[i]; q < c->cptr[i + 1]; q++) acc += c->cval[q] * y[c->cidx[q] - 1];
            g[i] = acc - bp[nv + i];
        }
        if (want_cov) {
            // S = R R^T (host Cholesky, m small); diag(W S^-1 W^T)_i = |R^-1 w_i|^2
            std::vector<double> R((size_t)m * m, 0.0), yv(m);
            for (int j = 0; j < m; j++) {
                double d = S[(size_t)j * m + j];
                for (int k = 0; k < j; k++) d -= R[(size_t)j * m + k] * R[(size_t)j * m + k];
                if (!(d > 0.0)) throw MpError(29, "constraint matrix C A^-1 C^T not positive definite");
                R[(size_t)j * m + j] = std::sqrt(d);
                for (int i = j + 1; i < m; i++) {
                    double e = S[(size_t)i * m + j];
                    for (int k = 0; k < j; k++) e -= R[(size_t)i * m + k] * R[(size_t)j * m + k];
                    R[(size_t)i * m + j] = e / R[(size_t)j * m + j];
                }
            }
            for (int i = 0; i < nv; i++) {
                double acc2 = 0.0;
                for (int j = 0; j < m; j++) {
                    double e = Z[(size_t)j * nv + i];
                    for (int k = 0; k < j; k++) e -= R[(size_t)j * m + k] * yv[k];
                    yv[j] = e / R[(size_t)j * m + j];
                    acc2 += yv[j] * yv[j];
                }
                covcorr[i] = acc2;
            }
        }
        // dense symmetric solve of S lambda = g (Gaussian elimination with partial pivoting, m small)
        std::vector<double> Aug(S);
        for (int k = 0; k < m; k++) {
            int p = k;
            for (int i = k + 1; i < m; i++) if (std::fabs(Aug[(size_t)i * m + k]) > std::fabs(Aug[(size_t)p * m + k])) p = i;
            if (Aug[(size_t)p * m + k] == 0.0) throw MpError(29, "singular constraint system");
            if (p != k) { for (int j = 0; j < m; j++) std::swap(Aug[(size_t)k * m + j], Aug[(size_t)p * m + j]); std::swap(g[k], g[p]); }
            for (int i = k + 1; i < m; i++) {
                double f = Aug[(size_t)i * m + k] / Aug[(size_t)k * m + k];
                for (int j = k; j < m; j++) Aug[(size_t)i * m + j] -= f * Aug[(size_t)k * m + j];
                g[i] -= f * g[k];
            }
        }
        std::vector<double> lam(m);
        for (int i = m - 1; i >= 0; i--) {
            double acc = g[i];
            for (int j = i + 1; j < m; j++) acc -= Aug[(size_t)i * m + j] * lam[j];
            lam[i] = acc / Aug[(size_t)i * m + i];
        }
        std::vector<double> xs(n);
        for (int i = 0; i < nv; i++) {
            double acc = y[i];
            for (int j = 0; j < m; j++) acc -= Z[(size_t)j * nv + i] * lam[j];
            xs[i] = acc;
        }
        for (int j = 0; j < m; j++) xs[nv + j] = lam[j];
        HIPCHK(hipMemcpyAsync(x_dev, xs.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
        HIPCHK(hipStreamSynchronize(s));
    }
    loc.rel_residual = rel_residual(c, b_dev, x_dev);
    if (want_cov) {
        const int nbk = std::min(nv, 512);
        DBuf<double> E, dd, dg;
        E.alloc((size_t)nv * nbk);
        dd.alloc(nv);
        dg.alloc(nv);
        ldl_inv_diag(nv, F.a.p, nv, F.w.p, F.work.p, E.p, nbk, dd.p, s);
        l_get_diag(model_of(c), nv, dg.p, s);
        c->cov_diag.assign(nv, 0.0);
        c->diag_before.assign(n, 0.0);
        dd.download(c->cov_diag.data(), nv, s);
        dg.download(c->diag_before.data(), nv, s);
        HIPCHK(hipStreamSynchronize(s));
        if (F.rev) std::reverse(c->cov_diag.begin(), c->cov_diag.end());
        for (int i = 0; i < nv; i++) c->cov_diag[i] -= covcorr[i];
        c->have_cov = true;
    }
}

extern "C" int mpgpu_solve(mpgpu_ctx *c, const double *rhs, double *x, mpgpu_solve_stats *st)
{
    if (!c || !c->built) return fail(c, 30, "structure not built");
    API_TRY
    HIPCHK(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const int64_t n = c->nall;
    double t0 = now_ms();
    mpgpu_solve_stats loc{};
    loc.method = c->opt.metsol;
    DBuf<double> b, xd;
    b.alloc(n);
    xd.alloc(n);
    if (rhs) b.upload(rhs, n, s); else b.upload(c->last_vec.data(), n, s);
    if (c->elim && c->opt.metsol != MPGPU_MINRES && c->opt.metsol != MPGPU_MINRES_QLP)
        throw MpError(36, "structure was built for constraint elimination (MINRES methods); rebuild it for this method");
    if (c->opt.metsol == MPGPU_INVERSION && c->have_inverse) {
        // matrix already inverted by an earlier solve: x = A^-1 b (dbsvxl, src/mpnum.f90:1308)
        l_packed_symv((int)n, c->d_mat.p, b.p, xd.p, c->d_psymv.p, s);
        loc.nrank = -1;
        loc.rel_residual = -1.0;
    } else if (c->opt.metsol == MPGPU_INVERSION) {
        if (c->opt.matsto == MPGPU_SPARSE) throw MpError(36, "method inversion needs full matrix storage");
        // workspaceDiag = diagonal before inversion (minver, src/pede.f90:7859-7861)
        DBuf<double> dg;
        dg.alloc(n);
        l_get_diag(model_of(c), (int32_t)n, dg.p, s);
        c->diag_before.assign(n, 0.0);
        dg.download(c->diag_before.data(), n, s);
        // keep the matrix for the residual check, invert a copy in place of the stored one
        DBuf<double> keep;
        keep.alloc(c->nstored);
        HIPCHK(hipMemcpyAsync(keep.p, c->d_mat.p, sizeof(double) * c->nstored, hipMemcpyDeviceToDevice, s));
        DBuf<double> diag, vk;
        DBuf<int> used, info;
        diag.alloc(n);
        vk.alloc(n + 1);
        used.alloc(n);
        info.alloc(4);
        HIPCHK(hipMemcpyAsync(xd.p, b.p, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        l_gj_init((int)n, c->d_mat.p, diag.p, used.p, info.p, s);
        for (int64_t step = 0; step < n; step++) l_gj_step((int)n, c->d_mat.p, diag.p, used.p, info.p, vk.p, xd.p, s);
        l_gj_finish((int)n, c->d_mat.p, xd.p, used.p, info.p, s);
        int hinfo[4];
        info.download(hinfo, 4, s);
        HIPCHK(hipStreamSynchronize(s));
        loc.nrank = hinfo[1];
        // residual with the kept matrix: swap buffers
        std::swap(keep.p, c->d_mat.p);
        loc.rel_residual = rel_residual(c, b.p, xd.p);
        std::swap(keep.p, c->d_mat.p);
        c->have_inverse = true;
        c->pc.reset();
        c->dfac.reset();
    } else if (c->opt.metsol == MPGPU_MINRES) {
        if (c->elim) elim_solve(c, b.p, xd.p, &loc, false);
        else {
            minres_dev(c, b.p, xd.p, &loc);
            loc.rel_residual = rel_residual(c, b.p, xd.p);
        }
    } else if (c->opt.metsol == MPGPU_MINRES_QLP) {
        if (c->elim) elim_solve(c, b.p, xd.p, &loc, true);
        else {
            minresqlp_dev(c, b.p, xd.p, &loc);
            loc.rel_residual = rel_residual(c, b.p, xd.p);
        }
    } else if (c->opt.metsol == MPGPU_DECOMPOSITION || c->opt.metsol == MPGPU_CHOLESKY_FP32_IR ||
               c->opt.metsol == MPGPU_LAPACK_PACKED || c->opt.metsol == MPGPU_LAPACK_UNPACKED) {
        if (c->opt.matsto == MPGPU_SPARSE) throw MpError(36, "this method needs full matrix storage");
        dense_solve(c, b.p, xd.p, loc);
    } else {
        throw MpError(36, "Unsupported solution method");
    }
    xd.download(x, n, s);
    HIPCHK(hipStreamSynchronize(s));
    loc.ms = now_ms() - t0;
    if (st) *st = loc;
    return 0;
    API_CATCH(c)
}

extern "C" int mpgpu_avprod(mpgpu_ctx *c, const double *x, double *y)
{
    if (!c || !c->built) return 30;
    API_TRY
    if (c->have_inverse) throw MpError(30, "matrix holds its inverse after an inversion solve");
    DBuf<double> xd, yd;
    xd.upload(x, c->nall, c->stream);
    yd.alloc(c->nall);
    avprod_dev(c, xd.p, yd.p);
    yd.download(y, c->nall, c->stream);
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
    API_CATCH(c)
}

// y = A x with x, y in device memory, allocating nothing (full storage, nall <= 16384)
extern "C" int mpgpu_avprod_noalloc(mpgpu_ctx *c, const double *x_dev, double *y_dev)
{
    if (!c || !c->built) return 30;
    if (!x_dev || !y_dev) return fail(c, 30, "null device pointer");
    if (c->opt.matsto != MPGPU_FULL) return fail(c, 30, "the allocation-free product needs full matrix storage");
    if (c->have_inverse) return fail(c, 30, "matrix holds its inverse after an inversion solve");
    if (c->nall > noalloc_max_n())
        return fail(c, 30, "the allocation-free product is limited to " + std::to_string(noalloc_max_n()) +
                               " rows (matrix has " + std::to_string(c->nall) + ")");
    if (hipSetDevice(c->device) != hipSuccess) return fail(c, 30, "hipSetDevice failed");
    hipError_t e = noalloc_packed_symv(c->device, (int)c->nall, c->d_mat.p, x_dev, y_dev, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return fail(c, 30, std::string("HIP error: ") + hipGetErrorString(e));
    return 0;
}

extern "C" int mpgpu_get_errors(const mpgpu_ctx *c0, double *err, double *gcor)
{
    mpgpu_ctx *c = const_cast<mpgpu_ctx *>(c0);
    if (!c || !c->built) return 30;
    API_TRY
    if (!c->have_inverse && !c->have_cov)
        throw MpError(30, "errors need an inversion solve or a LAPACK-method solve with lapack_errors");
    std::vector<double> d(c->nvar);
    if (c->have_inverse) {
        DBuf<double> dg;
        dg.alloc(c->nvar);
        l_get_diag(model_of(c), c->nvar, dg.p, c->stream);
        dg.download(d.data(), c->nvar, c->stream);
        HIPCHK(hipStreamSynchronize(c->stream));
    } else {
        d = c->cov_diag;
    }
    // prtglo (src/pede.f90:4536-4546): err = sign * sqrt(|A^-1_ii|) in REAL(4)
    for (int32_t v = 0; v < c->nvar; v++) {
        double gmati = d[v];
        float e = std::sqrt(std::fabs((float)gmati));
        if (gmati < 0.0) e = -e;
        err[v] = e;
        if (gcor) {
            float g = -1.0f;
            double diag = c->diag_before[v];
            if (gmati * diag > 0.0) {
                double g2 = 1.0 - 1.0 / (gmati * diag);
                if (g2 >= 0.0 && g2 <= 1.0) g = (float)std::sqrt(g2);
            }
            gcor[v] = g;
        }
    }
    return 0;
    API_CATCH(c)
}

extern "C" int mpgpu_get_monitor(const mpgpu_ctx *c, int32_t *ngroups, int32_t *labels, float *mean_err, int64_t *hist)
{
    if (!c || !c->built) return 30;
    if (c->opt.monitor == 0 || !c->d_mon_hist.p) return fail(const_cast<mpgpu_ctx *>(c), 24, "residual monitoring is off");
    const int32_t ng = (int32_t)c->mon_tot.size();
    if (ngroups) *ngroups = ng;
    if (!labels) return 0;
    API_TRY
    HIPCHK(hipSetDevice(c->device));
    std::vector<unsigned int> h((size_t)ng * c->mon_bins);
    c->d_mon_hist.download(h.data(), h.size(), c->stream);
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int32_t g = 0; g < ng; g++) {
        labels[g] = c->labels[c->mon_tot[g]];
        if (mean_err) mean_err[g] = (float)c->mon_res[g];
    }
    if (hist)
        for (size_t k = 0; k < h.size(); k++) hist[k] = (int64_t)h[k];
    return 0;
    API_CATCH(const_cast<mpgpu_ctx *>(c))
}
