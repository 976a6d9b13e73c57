// Serving loop: alignment/calibration jobs of different sizes arrive over time, are admitted
// by their device-memory footprint under a budget, run concurrently on one GPU (one context
// per job on its worker's HIP stream, so the kernels of different jobs overlap on the device)
// and report their latency from arrival to result.  Submitting a job only queues it: the job
// refers to the caller's arrays (kept valid until the job is waited for), and the worker that
// takes it sizes it and waits until it fits the budget.
// No pede counterpart: pede runs one job per process (src/pede.f90:33-528).
#include "mpgpu_internal.h"
#include <algorithm>
#include <chrono>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <omp.h>
#include <thread>

namespace {

using clk = std::chrono::steady_clock;

struct Job {
    int64_t id = 0;
    mpgpu_options opt{};
    // the caller's arrays (valid until the job has been waited for)
    int64_t nrec = 0;
    const int64_t *rec_ptr = nullptr;
    const int32_t *inder = nullptr;
    const float *glder = nullptr;
    int32_t npar = 0;
    const int32_t *plab = nullptr;
    const double *pval = nullptr, *psig = nullptr;
    int32_t ncon = 0;
    const int64_t *cptr = nullptr;
    const int32_t *clab = nullptr;
    const double *cfac = nullptr, *crhs = nullptr;
    int64_t footprint = -1;   // device bytes, sized by the worker that takes the job
    double size_ms = 0;       // time of the sizing
    clk::time_point t_arrival, t_start, t_end;
    // results
    int status = -1;
    std::string err;
    std::vector<int32_t> labels;
    std::vector<double> corr;
    mpgpu_solve_stats sst{};
    double chi2 = 0;
    int64_t ndf = 0;
    bool done = false;
};

// device bytes a job needs: records (8 B/word + compiled tables), the matrix (full packed or
// an upper estimate of the group-block storage), the update buffer of the two-phase build
// (group-block storage: at most the lower triangle of each record's global words, 8 B each),
// vectors and solver work space
int64_t estimate_footprint(const Job &j)
{
    const int64_t nrec = j.nrec;
    const int64_t nw = nrec > 0 ? j.rec_ptr[nrec] - j.rec_ptr[0] : 0;
    int64_t rec = 12 * nw + 64 * nrec;
    // distinct labels of the global words: after the second zero-index word of a measurement
    // (measurement, locals, sigma, globals; src/mille.f90:27-59).  Two passes over the records
    // (the worker's OpenMP threads): their global words and label range, then a bitmap
    const int64_t *rp = j.rec_ptr;
    const int32_t *I = j.inder;
    int64_t upd = 0;
    int32_t lmin = INT32_MAX, lmax = INT32_MIN;
#pragma omp parallel for schedule(static, 256) reduction(+ : upd) reduction(min : lmin) reduction(max : lmax)
    for (int64_t r = 0; r < nrec; r++) {
        int zeros = 0;
        int64_t g = 0;
        for (int64_t w = rp[r] + 1; w < rp[r + 1]; w++) {
            const int32_t l = I[w];
            if (l == 0) { zeros++; continue; }
            if (zeros > 0 && zeros % 2 == 0) {
                g++;
                lmin = std::min(lmin, l);
                lmax = std::max(lmax, l);
            }
        }
        upd += 8 * g * (g + 1) / 2;
    }
    if (j.opt.matsto == MPGPU_FULL) upd = 0;
    upd = std::min<int64_t>(upd, (int64_t)128 << 30);
    int64_t nlab = 0;
    if (lmax >= lmin) {
        const int64_t span = (int64_t)lmax - lmin + 1;
        std::vector<uint64_t> bits((span + 63) / 64, 0);
        uint64_t *B = bits.data();
#pragma omp parallel for schedule(static, 256)
        for (int64_t r = 0; r < nrec; r++) {
            int zeros = 0;
            for (int64_t w = rp[r] + 1; w < rp[r + 1]; w++) {
                const int32_t l = I[w];
                if (l == 0) { zeros++; continue; }
                if (zeros > 0 && zeros % 2 == 0) {
                    const int64_t o = (int64_t)l - lmin;
                    const uint64_t m = 1ull << (o & 63);
                    if ((__atomic_load_n(&B[o >> 6], __ATOMIC_RELAXED) & m) == 0) __atomic_fetch_or(&B[o >> 6], m, __ATOMIC_RELAXED);
                }
            }
        }
        for (uint64_t b : bits) nlab += __builtin_popcountll(b);
    }
    int64_t nall = nlab + (int64_t)j.ncon;
    int64_t mat;
    if (j.opt.matsto == MPGPU_FULL) {
        // packed matrix, plus what the solve allocates beside it: the inversion keeps a copy
        // (1x packed); the FP64 factor is a full n x n FP64 matrix (2x packed); the FP32 factor
        // a full FP32 one (1x) and, when the refinement does not converge, the FP64 factor
        // as well (2x); the error diagonal of the LAPACK methods 512 unit columns
        const int64_t packed = 8 * nall * (nall + 1) / 2;
        int64_t mult = 2;
        switch (j.opt.metsol) {
        case MPGPU_DECOMPOSITION: case MPGPU_LAPACK_PACKED: case MPGPU_LAPACK_UNPACKED: mult = 3; break;
        case MPGPU_CHOLESKY_FP32_IR: mult = 4; break;
        default: mult = 2;
        }
        mat = packed * mult;
        if (j.opt.lapack_errors && (j.opt.metsol == MPGPU_LAPACK_PACKED || j.opt.metsol == MPGPU_LAPACK_UNPACKED))
            mat += 8 * nall * 512;
    } else {
        mat = std::min<int64_t>(8 * nall * (nall + 1) / 2, 8 * nw * 24);
    }
    return rec + mat + upd + 200 * nall + (64 << 20);
}

} // namespace

struct mpgpu_server {
    int device = 0;
    int64_t budget = 0, in_use = 0;
    int batch_max_params = 5000;   // (kept for the interface; jobs are run one per worker)
    std::mutex mu;
    std::condition_variable cv_work, cv_done, cv_budget;
    std::deque<std::shared_ptr<Job>> queue;
    std::map<int64_t, std::shared_ptr<Job>> jobs;
    std::vector<std::thread> workers;
    bool stop = false;
    int64_t next_id = 1;

    void run_job(Job &j, hipStream_t wstream)
    {
        j.t_start = clk::now();
        // MPGPU_VERBOSE: time of each step of the job on stderr
        static const bool verbose = std::getenv("MPGPU_VERBOSE") != nullptr;
        clk::time_point tp = j.t_start;
        double ms[6] = {0, 0, 0, 0, 0, 0};
        auto lap = [&](int k) {
            const clk::time_point t = clk::now();
            ms[k] = std::chrono::duration<double, std::milli>(t - tp).count();
            tp = t;
        };
        int rc = 0;
        mpgpu_ctx *c = nullptr;
        rc = mpgpu::create_context(&j.opt, device, wstream, &c);
        if (!rc) rc = mpgpu_add_records(c, j.nrec, j.rec_ptr, j.inder, j.glder, nullptr);
        if (!rc && j.npar > 0) rc = mpgpu_set_steering_parameters(c, j.npar, j.plab, j.pval, j.psig);
        if (!rc && j.ncon > 0) {
            std::vector<double> sig(j.ncon, 0.0);
            rc = mpgpu_add_equations(c, 0, j.ncon, j.cptr, j.clab, j.cfac, j.crhs, sig.data());
        }
        lap(0);
        int32_t ntot = 0, nvar = 0, ncon = 0;
        if (!rc) rc = mpgpu_build_structure(c, &ntot, &nvar, &ncon);
        lap(1);
        mpgpu_pass_stats ps{};
        if (!rc) rc = mpgpu_loopn(c, 1, 1, 0.0f, &ps);
        lap(2);
        std::vector<double> x((size_t)nvar + ncon);
        if (!rc) rc = mpgpu_solve(c, nullptr, x.data(), &j.sst);
        lap(3);
        if (verbose)
            std::fprintf(stderr, "job %lld: nvar %d nrec %lld: size %.1f create+copy %.1f structure %.1f pass %.1f solve %.1f ms "
                                 "(method %d, itn %d)\n",
                         (long long)j.id, nvar, (long long)j.nrec, j.size_ms, ms[0], ms[1], ms[2], ms[3],
                         j.sst.method, j.sst.itn);
        if (!rc) {
            std::vector<int32_t> lab(ntot), var(ntot), grp(ntot);
            std::vector<int64_t> cnt(ntot);
            mpgpu_get_labels(c, lab.data(), var.data(), cnt.data(), grp.data());
            j.labels.clear();
            j.corr.clear();
            for (int32_t t = 0; t < ntot; t++)
                if (var[t] > 0) { j.labels.push_back(lab[t]); j.corr.push_back(x[var[t] - 1]); }
            j.chi2 = ps.chi2;
            j.ndf = ps.ndf;
        } else if (c) {
            j.err = mpgpu_last_error(c);
        }
        if (c) mpgpu_destroy(c);
        j.status = rc;
        j.t_end = clk::now();
    }

    int omp_threads = 1;

    void worker()
    {
        (void)hipSetDevice(device);
        // the host parts of the workers' jobs share the host's cores instead of each taking all
        omp_set_num_threads(omp_threads);
        // one stream per worker for all its jobs (a stream made and destroyed per job costs
        // milliseconds); a job's context orders its work and frees its buffers on it
        hipStream_t wstream = nullptr;
        if (hipStreamCreateWithFlags(&wstream, hipStreamNonBlocking) != hipSuccess) wstream = nullptr;
        struct StreamGuard { hipStream_t s; ~StreamGuard() { if (s) { (void)hipStreamSynchronize(s); (void)hipStreamDestroy(s); } } } guard{wstream};
        for (;;) {
            std::shared_ptr<Job> j;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_work.wait(lk, [&] { return stop || !queue.empty(); });
                if (queue.empty()) return;   // stop, and nothing left
                j = queue.front();
                queue.pop_front();
            }
            // size the job (a scan of its words), then wait until it fits the budget; a job
            // alone may use more than the budget's remainder when nothing else runs
            const clk::time_point t_size = clk::now();
            j->footprint = estimate_footprint(*j);
            j->size_ms = std::chrono::duration<double, std::milli>(clk::now() - t_size).count();
            bool admitted = false;
            {
                std::unique_lock<std::mutex> lk(mu);
                if (j->footprint <= budget) {
                    cv_budget.wait(lk, [&] { return in_use + j->footprint <= budget || in_use == 0; });
                    in_use += j->footprint;
                    admitted = true;
                }
            }
            if (admitted) {
                run_job(*j, wstream);
            } else {
                j->t_start = j->t_end = clk::now();
                j->status = 30;
                j->err = "job larger than the memory budget";
            }
            {
                std::lock_guard<std::mutex> lk(mu);
                if (admitted) in_use -= j->footprint;
                j->done = true;
            }
            cv_done.notify_all();
            cv_budget.notify_all();
        }
    }
};

extern "C" {

int mpgpu_server_create(int32_t device, int32_t nworkers, int64_t budget_bytes, int32_t batch_max_params, mpgpu_server **out)
{
    if (!out || nworkers <= 0) return 30;
    auto *s = new mpgpu_server();
    s->device = device;
    s->budget = budget_bytes;
    s->batch_max_params = batch_max_params;
    // the host threads the process may use (OMP_NUM_THREADS when set: the cores a process on a
    // shared host is given, which may be far fewer than the processors it sees), shared out
    s->omp_threads = std::max(1, std::min(omp_get_num_procs(), omp_get_max_threads()) / nworkers);
    for (int i = 0; i < nworkers; i++) s->workers.emplace_back([s] { s->worker(); });
    *out = s;
    return 0;
}

int mpgpu_server_submit(mpgpu_server *s, const mpgpu_options *opt, int64_t nrec, const int64_t *rec_ptr, const int32_t *inder,
                        const float *glder, int32_t npar, const int32_t *plab, const double *pval, const double *psig,
                        int32_t ncon, const int64_t *cptr, const int32_t *clab, const double *cfac, const double *crhs,
                        int64_t *job_id)
{
    if (!s || !opt || nrec < 0) return 30;
    if (nrec > 0 && (!rec_ptr || !inder || !glder)) return 30;
    if (ncon > 0 && (!cptr || !clab || !cfac || !crhs)) return 30;
    auto j = std::make_shared<Job>();
    j->t_arrival = clk::now();
    j->opt = *opt;
    j->nrec = nrec;
    j->rec_ptr = rec_ptr;
    j->inder = inder;
    j->glder = glder;
    if (npar > 0) {
        j->npar = npar;
        j->plab = plab;
        j->pval = pval;
        j->psig = psig;
    }
    if (ncon > 0) {
        j->ncon = ncon;
        j->cptr = cptr;
        j->clab = clab;
        j->cfac = cfac;
        j->crhs = crhs;
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) return 30;
        j->id = s->next_id++;
        s->jobs[j->id] = j;
        s->queue.push_back(j);
    }
    s->cv_work.notify_one();
    if (job_id) *job_id = j->id;
    return 0;
}

// waits for a job; corr/labels may be NULL (then only statistics); capacity = size of arrays
int mpgpu_server_wait(mpgpu_server *s, int64_t job_id, int32_t capacity, int32_t *labels, double *corr, mpgpu_job_stats *st)
{
    if (!s) return 30;
    std::shared_ptr<Job> j;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->jobs.find(job_id);
        if (it == s->jobs.end()) return 30;
        j = it->second;
        s->cv_done.wait(lk, [&] { return j->done; });
        // another waiter on the same job may have removed it meanwhile
        auto again = s->jobs.find(job_id);
        if (again != s->jobs.end()) s->jobs.erase(again);
    }
    if (st) {
        std::memset(st, 0, sizeof(*st));
        st->status = j->status;
        st->nvar = (int32_t)j->labels.size();
        st->latency_ms = std::chrono::duration<double, std::milli>(j->t_end - j->t_arrival).count();
        st->queue_ms = std::chrono::duration<double, std::milli>(j->t_start - j->t_arrival).count();
        st->run_ms = std::chrono::duration<double, std::milli>(j->t_end - j->t_start).count();
        st->chi2 = j->chi2;
        st->ndf = j->ndf;
        st->istop = j->sst.istop;
        st->itn = j->sst.itn;
        st->rel_residual = j->sst.rel_residual;
        st->t_arrival_s = std::chrono::duration<double>(j->t_arrival.time_since_epoch()).count();
        st->t_done_s = std::chrono::duration<double>(j->t_end.time_since_epoch()).count();
        st->method = j->sst.method;
    }
    if (labels && corr) {
        int32_t n = std::min<int32_t>(capacity, (int32_t)j->labels.size());
        std::copy(j->labels.begin(), j->labels.begin() + n, labels);
        std::copy(j->corr.begin(), j->corr.begin() + n, corr);
    }
    return j->status;
}

int mpgpu_server_destroy(mpgpu_server *s)
{
    if (!s) return 0;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto &t : s->workers) t.join();
    delete s;
    return 0;
}

} // extern "C"
