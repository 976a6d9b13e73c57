/*
 * mpgpu.h -- C-ABI of the MI355X-native normal-matrix build and global solve for
 * Millepede-II pede.
 *
 * Flat ABI: plain pointers, explicit sizes, int status codes.  No C++ or torch types.
 * Status codes are the pede exit codes the reference would have stopped with
 * (doc/exitCodes.md); 0 = success.  The message of the last failure on a context is
 * returned by mpgpu_last_error().  Nothing here ever exits the process.
 *
 * Indexing follows the reference (src/mpmod.f90): labels are positive int32; "total index"
 * t = 1..ntot is the rank of a label in ascending label order (globalParLabelIndex(1,:),
 * src/mppar.f90:99,203-209); "variable index" v = 1..nvar (globalParLabelIndex(2,:), >0 for
 * variable parameters, -1 fixed by pre-sigma, -2 fixed by the entries cut, src/pede.f90:5956-5973);
 * Lagrange multipliers follow as nvar+1..nvar+ncon.  Packed symmetric matrices use pede's
 * lower-row-packed order, element (i,j), i>=j, at (i*i-i)/2+j (1-based; src/pede.f90:4204).
 *
 * Host buffers are owned by the caller; the library copies what it needs into device memory
 * it owns and keeps no host pointer after a call returns.  One context = one HIP stream;
 * calls on one context must be serialised by the caller; different contexts are independent.
 */
#ifndef MPGPU_H
#define MPGPU_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mpgpu_ctx mpgpu_ctx;

/* Solution methods: values of pede's metsol (src/mpmod.f90:35, src/pede.f90:10962-11016) */
enum {
    MPGPU_INVERSION = 1,        /* method inversion      -> minver  (src/pede.f90:7838)   */
    MPGPU_DECOMPOSITION = 3,    /* method decomposition  -> mchdec  (src/pede.f90:7951)   */
    MPGPU_MINRES = 4,           /* method (full|sparse)MINRES -> mminrs (src/pede.f90:8243) */
    MPGPU_MINRES_QLP = 5,       /* method (full|sparse)MINRES-QLP (src/pede.f90:8341)       */
    MPGPU_LAPACK_PACKED = 7,    /* method fullLAPACK     -> mdptrf (src/mplapack.f90:41): LL^T */
                                /* of the parameter block on the device, constraints by the */
                                /* Schur complement (same constrained solution)             */
    MPGPU_LAPACK_UNPACKED = 8,  /* method unpackedLAPACK -> mdutrf (src/mplapack.f90:187)    */
    MPGPU_CHOLESKY_FP32_IR = 11 /* no pede counterpart: FP32-MFMA Cholesky + FP64 refinement */
};

/* Matrix storage: pede's matsto (src/mpmod.f90:36) */
enum { MPGPU_FULL = 1, MPGPU_SPARSE = 2 };

/* Steering options used by the path; defaults are pede's (src/mpmod.f90:33-126).
 * Fill with mpgpu_default_options(). */
typedef struct mpgpu_options {
    int32_t metsol;          /* solution method (enum above)                          */
    int32_t matsto;          /* MPGPU_FULL or MPGPU_SPARSE                            */
    int32_t mbandw;          /* 'bandwidth': 0 diagonal, >0 band preconditioner, <0 none */
    int32_t lprecm;          /* 'bandwidth m 1': skyline (rank preserving) band        */
    int32_t lhuber;          /* 'outlierdownweighting n'                              */
    int32_t lfitbb;          /* local fit: check bordered-band structure (1)          */
    int32_t mreqenf;         /* 'entries f': min entries from binary files (25)       */
    int32_t mreqena;         /* 'entries f a': min entries from accepted fits (10)    */
    int32_t mcount;          /* 'countrecords': count per record (1) / equation (0)   */
    int32_t itnlim;          /* MINRES iteration limit (2000)                         */
    int32_t withmultipliers; /* constraints: Lagrange multipliers (1) or elimination by */
                             /* QL decomposition (0, pede's default; MINRES methods --  */
                             /* inversion and decomposition keep multipliers)           */
    int32_t mrmode;          /* MINRES-QLP mode                                       */
    float chicut;            /* 'chisqcut' first factor (0 = none)                    */
    float chirem;            /* 'chisqcut' second factor                              */
    float chhuge;            /* 'hugecut' (50)                                        */
    float dwcut;             /* 'dwfractioncut' (0)                                   */
    float cndlmx;            /* 'maxlocalcond' (100)                                  */
    float regula;            /* 'regularisation' factor (1)                           */
    double default_presigma; /* 'presigma' (0)                                        */
    double mrestl;           /* 'mrestol' (1e-6)                                      */
    double mrtcnd;           /* 'mrestranscond' (1e7)                                 */
    int32_t nregul;          /* regularisation flag                                   */
    int32_t refine_steps;    /* FP32+IR: maximum FP64 CG refinement steps (20)        */
    int32_t monitor;         /* residual monitoring of the passes: 0 off, 1 normalized */
                             /* residuals ('monitorresiduals'), 2 pulls ('monitorpulls') */
    int32_t monitor_bins;    /* bins per measurement group (100)                       */
    int32_t lapack_errors;   /* 'LAPACKwitherrors': parameter errors from the factor of */
                             /* the LAPACK methods (src/pede.f90:9017-9120, ilperr)     */
    int32_t pair_entries;    /* 'pairentries n': mreqpe, minimum number of entries of a */
                             /* pair of parameter groups for its matrix block to be      */
                             /* stored (src/pede.f90:10762-10772, 6564, 7036;            */
                             /* src/mpbits.f90); only 1 (every pair stored) is built:    */
                             /* mpgpu_build_structure returns 40 for larger values       */
    int32_t skip_empty_cons; /* 'skipemptycons' (iskpec): drop constraints without any  */
                             /* variable parameter; without it such a constraint stops  */
                             /* the run (src/pede.f90:1164-1190)                         */
    int32_t resolve_redundancy_cons; /* 'resolveredundancycons' (irslvrc): a constraint */
                             /* group (grpcon's grouping) with as many constraints as    */
                             /* parameters of its own fixes those parameters (-4) and is */
                             /* dropped (src/pede.f90:1015-1050, 5965-5966)              */
    double scale_errors[2];  /* 'scaleerrors f1 f2': errorScaleFactor, measurement      */
                             /* errors with (1) / without (2) global derivatives are     */
                             /* multiplied by it (src/mppar.f90:312-316); default 1, 1   */
    double refine_tol;       /* FP32+IR: stop when ||r|| <= refine_tol ||x|| ||A||_F     */
                             /* (0: eps sqrt(n), the test of LAPACK's DSPOSV)            */
    int32_t fp32_min_n;      /* FP32+IR: parameter blocks with fewer parameters use the  */
                             /* FP64 factor (reported as method 3, istop 4): below the   */
                             /* size where the FP32 factor saves more than the refinement */
                             /* costs (16384; 0: always try the FP32 factor)              */
    int32_t iteren;          /* iterated entries cut ('entries n1 n2 n3': n1*n3,        */
                             /* 'iterateentries n'; loop1i, src/pede.f90:6164-6265):     */
                             /* with counting per equation (mcount 0) and iteren above   */
                             /* mreqenf, a variable parameter with fewer than iteren     */
                             /* entries stays variable only with at least mreqenf        */
                             /* entries in equations without parameters fixed by the     */
                             /* entries cut, else it is fixed (-3); 0 = off              */
} mpgpu_options;

/* Statistics of one data pass (loopn) */
typedef struct mpgpu_pass_stats {
    double chi2;         /* Sum(chi^2) of the pass (getsums fvalue, src/pede.f90:11699)   */
    int64_t ndf;         /* Sum(ndf)                                                      */
    double wndf;         /* weighted Sum(ndf)                                             */
    int64_t nrej[6];     /* rejects: NaN, rank, cond(band), ndf<=0, huge chi2, chi2 cut   */
    int64_t naccept;     /* accepted local fits                                           */
    int64_t nrec;        /* records processed                                             */
    double ms_kernel;    /* device time of the record loop (all record and gather kernels) */
    double ms_total;     /* device time of the whole pass                                 */
    double ms_records;   /* device time of the record-kernel launches (HIP events each)  */
    int64_t n_records;   /* number of record-kernel launches                              */
    double ms_gather;    /* device time of the block-gather launches (group-block storage) */
    int64_t n_gather;    /* number of block-gather launches                               */
} mpgpu_pass_stats;

/* Statistics of one solve */
typedef struct mpgpu_solve_stats {
    int32_t method;      /* method actually used                                          */
    int32_t nrank;       /* rank found (inversion / decomposition: positive pivots)       */
    int32_t istop;       /* MINRES stop code (src/minresModule.f90:591-600); FP32+IR: 1   */
                         /* refinement converged, 3 not within refine_steps (the FP64   */
                         /* factor then solved the system: method = 3), 4 block smaller  */
                         /* than fp32_min_n (solved by the FP64 factor: method = 3)      */
    int32_t itn;         /* MINRES iterations / refinement steps                          */
    double anorm, acond, rnorm, arnorm, ynorm;
    double rel_residual; /* ||A x - b|| / ||b||, evaluated in FP64 on the device           */
    double ms;           /* wall time of the solve (device-synchronised)                  */
    double evmax, evmin; /* decomposition: largest / smallest positive pivot of LDL^T       */
                         /* (chdec2, src/mpnum.f90:917-925)                                */
} mpgpu_solve_stats;

/* Counts of the structure build, as pede's first and second data loops print them in
 * millepede.log (loop1, src/pede.f90:6124-6141, counters of prepareRecords,
 * src/mppar.f90:345-351; loop2, src/pede.f90:6683-6838 and 7271-7320) */
typedef struct mpgpu_record_summary {
    int64_t nrec;          /* NREC: records                                                  */
    int64_t neq;           /* nbEqRead: equations (measurements)                             */
    int64_t neq_glob;      /* nbEqWithGlobPar: equations with global derivatives             */
    int64_t nglob_deriv;   /* nbGlobDeriv: global derivatives (words, zeros included)        */
    int32_t fp64_payload;  /* 1 if any record carries FP64 values (pede's NRECD > 0)         */
    int32_t ntot;          /* nTotalGlobalPar                                                */
    int32_t nvar;          /* nVarGlobalPar                                                  */
    int32_t ngroup_tot;    /* nTotalParGroups: groups over all parameters                    */
    int32_t ngroup_var;    /* nVarParGroups: groups with variable parameters                 */
    int32_t ncon;          /* nConstraints: constraints kept (empty / redundant ones dropped) */
    int32_t max_glob_rec;  /* nGlobalMaxPerRec: most distinct variable parameters in a record */
    int32_t max_loc_rec;   /* nLocalMaxPerRec: largest local parameter index in a record     */
    int32_t max_eq_rec;    /* nEqMaxPerRec: most equations in a record                       */
    int64_t nempty_rec;    /* nEmptyRecords: records without variable parameters             */
} mpgpu_record_summary;

/* ---- context ---------------------------------------------------------------------------- */
void mpgpu_default_options(mpgpu_options *opt);
/* create a context bound to HIP device `device`; no pede counterpart (pede's state is mpmod) */
int mpgpu_create(const mpgpu_options *opt, int32_t device, mpgpu_ctx **out);
int mpgpu_destroy(mpgpu_ctx *ctx);
const char *mpgpu_last_error(const mpgpu_ctx *ctx);
int mpgpu_set_options(mpgpu_ctx *ctx, const mpgpu_options *opt);

/* ---- input: Mille records (replaces readFromBinary + read_mille, src/pede.f90:1714-2075) --
 * Records are appended in order.  Record r occupies words [rec_ptr[r], rec_ptr[r+1]) of
 * inder/glder exactly as they are stored in a Mille record (word 1 = (0.0, error count),
 * src/mille.f90:27-59).  Exactly one of glder_f (FP32 payload) / glder_d (FP64) is non-NULL.
 * Labels in global words stay raw labels.  The records are in device memory when the call
 * returns; the caller's arrays may be reused. */
int mpgpu_add_records(mpgpu_ctx *ctx, int64_t nrec, const int64_t *rec_ptr, const int32_t *inder,
                      const float *glder_f, const double *glder_d);
/* Read a whole Mille binary file (C binary, or Fortran sequential when fortran != 0; plain or
 * gzip-compressed; ROOT files are not supported) and append its records: each record is read
 * into a pinned staging buffer, and a full buffer is copied to the device while the next one
 * fills.  On an error (18) the file's records are dropped and earlier ones stay. */
int mpgpu_read_mille_file(mpgpu_ctx *ctx, const char *path, int32_t fortran, int64_t *nrec_out);

/* ---- parameters from steering (Parameter lists, src/pede.f90:5925-5950) --------------------
 * labels[k], start value[k], presigma[k] (0 = none, <0 fixed, >0 pre-sigma). */
int mpgpu_set_steering_parameters(mpgpu_ctx *ctx, int32_t n, const int32_t *labels, const double *values,
                                  const double *presigma);
/* Constraint or Measurement blocks (src/pede.f90:10916-11075).  Block c has rhs[c], sigma[c]
 * and entries [ptr[c], ptr[c+1]) of (labels, factors).  kind: 0 constraint, 1 measurement. */
int mpgpu_add_equations(mpgpu_ctx *ctx, int32_t kind, int32_t nblock, const int64_t *ptr, const int32_t *labels,
                        const double *factors, const double *rhs, const double *sigma);

/* ---- structure (loop1 + loop2, src/pede.f90:5777-7527; src/mppar.f90; src/mpbits.f90) ----
 * Builds label -> index maps, entry counts, variable/fixed flags, parameter groups and the
 * sparsity structure of the global matrix; compiles the records into device layout. */
int mpgpu_build_structure(mpgpu_ctx *ctx, int32_t *ntot_out, int32_t *nvar_out, int32_t *ncon_out);
/* globalParLabelIndex(1:2,:), globalParLabelCounter, group start labels (ntot each) */
int mpgpu_get_labels(const mpgpu_ctx *ctx, int32_t *labels, int32_t *var_index, int64_t *counters,
                     int32_t *group_first_label);
/* stored elements of the global matrix (matrix words), number of group blocks */
int mpgpu_matrix_info(const mpgpu_ctx *ctx, int64_t *stored_words, int64_t *nblocks, int32_t *ngroups);

/* group-block storage of the matrix (sparse storage; the information of pede's
 * sparseMatrixOffsets / sparseMatrixColumns, src/mpbits.f90:1220-1350, in contiguous blocks):
 * gfirst[ngroups+1] first variable index (1-based) of each parameter group, brow[ngroups+1]
 * first block of each block row, bcol[nblocks] column group of each block (ascending within a
 * row, column group <= row group), boff[nblocks] offset of the row-major block in the stored
 * matrix words; a diagonal block uses its lower triangle.  Sizes from mpgpu_matrix_info. */
int mpgpu_get_block_structure(const mpgpu_ctx *ctx, int32_t *gfirst, int64_t *brow, int32_t *bcol, int64_t *boff);
/* the choices of the library's autotuner for this structure (no counterpart in the reference:
 * pede has no tunable kernels): the group-block product variants timed at the structure build,
 * with their median times, and the one chosen; NUL-terminated text of at most len bytes */
int mpgpu_get_tuning(const mpgpu_ctx *ctx, char *text, int32_t len);
/* the counts pede's loop1 / loop2 write to millepede.log (after mpgpu_build_structure) */
int mpgpu_get_record_summary(const mpgpu_ctx *ctx, mpgpu_record_summary *out);

/* ---- state ------------------------------------------------------------------------------- */
int mpgpu_set_parameters(mpgpu_ctx *ctx, const double *par_by_total_index);   /* globalParameter */
int mpgpu_get_parameters(const mpgpu_ctx *ctx, double *par_by_total_index);

/* ---- data pass: loopn + loopbf (src/pede.f90:2479-2885, 3387-4442) -------------------------
 * icalcm = 1 builds matrix + vector + chi2, 0 vector + chi2 only.  nloopn is pede's pass
 * counter (1 for the first pass: no down-weighting), chicut the cut factor in force. Adds the
 * pre-sigma diagonal (src/pede.f90:2711-2718) and external measurements, then the Lagrange
 * rows and constraint residuals (addcst, src/pede.f90:704-771). */
int mpgpu_loopn(mpgpu_ctx *ctx, int32_t icalcm, int32_t nloopn, float chicut, mpgpu_pass_stats *st);
/* globalVector (nall) and globalCounter (nvar) of the last pass */
int mpgpu_get_vector(const mpgpu_ctx *ctx, double *vec, int64_t *counter);
/* whole matrix as packed nall(nall+1)/2 FP64 (tests; small nall only) */
int mpgpu_get_matrix_packed(const mpgpu_ctx *ctx, double *packed, int64_t len);
/* mupdat: add v[k] to element (i[k], j[k]) (1-based, any order; src/pede.f90:3134-3204) */
int mpgpu_mupdat(mpgpu_ctx *ctx, int32_t nel, const int32_t *i, const int32_t *j, const double *v);

/* ---- global solve (xloopn metsol dispatch, src/pede.f90:8822-8855) -------------------------
 * Solves A x = rhs for the matrix of the last icalcm=1 pass.  rhs may be NULL (use the
 * globalVector of the last pass, as xloopn copies it into globalCorrections).  x: nall. */
int mpgpu_solve(mpgpu_ctx *ctx, const double *rhs, double *x, mpgpu_solve_stats *st);
/* y = A x (avprod, src/pede.f90:5331-5358), x,y host arrays of nall */
int mpgpu_avprod(mpgpu_ctx *ctx, const double *x, double *y);
/* y = A x like mpgpu_avprod, with x and y (nall each) in device memory, allocating no device
 * memory: no counterpart in the reference; for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops, harnesses that forbid it).  Full storage only, nall <=
 * 16384 (else 30).  The product's partial sums use a fixed ~19 MB device array per GPU, held
 * until the process ends and shared by every call on that GPU: calls that overlap on one GPU
 * race.  Returns after the product is complete. */
int mpgpu_avprod_noalloc(mpgpu_ctx *ctx, const double *x_dev, double *y_dev);
/* parameter errors sqrt|A^-1_ii| and global correlations (prtglo, src/pede.f90:4538-4546);
 * after an inversion solve, or a LAPACK-method solve with lapack_errors set (then from the
 * covariance of the constrained parameters, src/pede.f90:9017-9120).  err, gcor: nvar
 * (gcor may be NULL) */
int mpgpu_get_errors(const mpgpu_ctx *ctx, double *err, double *gcor);
/* residual monitoring (src/pede.f90:3955-3980 in loopbf, 6700-6710 and 7075-7085 in loop2):
 * histograms of the last data pass, one per measurement group -- the measurements with local
 * and global derivatives, grouped by the label of their first global derivative, in label order
 * -- of the normalized residuals (or pulls) of the first local-fit iteration, monitor_bins bins
 * of width 0.1 centred on zero; mean_err is the group's mean measurement error (summed in
 * record order, rounded to REAL(4) as pede prints it).  With labels == NULL only *ngroups is set.
 * Requires options.monitor. */
int mpgpu_get_monitor(const mpgpu_ctx *ctx, int32_t *ngroups, int32_t *labels, float *mean_err, int64_t *hist);

/* ---- complete pede-like run: steering file -> millepede.res (src/pede.f90:33-528) -------- */
typedef struct mpgpu_run_stats {
    int32_t ntot, nvar, ncon, iterations;
    int64_t nrec;
    double final_chi2;
    int64_t final_ndf;
    double ms_structure, ms_passes, ms_solves;
    int32_t exit_code;
} mpgpu_run_stats;
int mpgpu_run_steering(const char *steering_file, const char *output_dir, int32_t device, mpgpu_run_stats *st);

/* ---- serving: concurrent jobs on one GPU (no pede counterpart; pede runs one job) ---------
 * Jobs are admitted by their estimated device-memory footprint under budget_bytes and run by
 * nworkers workers, each with its own context and HIP stream, so the kernels of different
 * jobs overlap on the device.  A job is: options, FP32 Mille records, optional Parameter
 * lines (label, start value, pre-sigma) and Lagrange constraints (CSR: cptr[ncon+1]).
 * mpgpu_server_submit only queues the job: its arrays are not copied and must stay valid and
 * unchanged until mpgpu_server_wait has returned for it; the worker that takes the job sizes
 * it (status 30 when it exceeds the whole budget) and starts it when it fits.  Every job is
 * solved in its own context: batch_max_params is accepted and not used. */
typedef struct mpgpu_server mpgpu_server;
typedef struct mpgpu_job_stats {
    int32_t status, nvar;
    double latency_ms, queue_ms, run_ms; /* arrival -> result, arrival -> start, start -> result */
    double chi2;
    int64_t ndf;
    int32_t istop, itn;
    double rel_residual;
    double t_arrival_s, t_done_s;        /* the steady (CLOCK_MONOTONIC) clock in seconds at   */
                                         /* arrival (entry of submit) and at the result       */
    int32_t method, reserved;            /* solution method that ran (mpgpu_solve_stats)      */
} mpgpu_job_stats;
int mpgpu_server_create(int32_t device, int32_t nworkers, int64_t budget_bytes, int32_t batch_max_params,
                        mpgpu_server **out);
int mpgpu_server_submit(mpgpu_server *s, const mpgpu_options *opt, int64_t nrec, const int64_t *rec_ptr,
                        const int32_t *inder, const float *glder, int32_t npar, const int32_t *plab, const double *pval,
                        const double *psig, int32_t ncon, const int64_t *cptr, const int32_t *clab, const double *cfac,
                        const double *crhs, int64_t *job_id);
/* blocks until the job is done; returns its status; fills up to capacity (label, correction)
 * pairs of its variable parameters */
int mpgpu_server_wait(mpgpu_server *s, int64_t job_id, int32_t capacity, int32_t *labels, double *corr,
                      mpgpu_job_stats *st);
int mpgpu_server_destroy(mpgpu_server *s);

#ifdef __cplusplus
}
#endif
#endif
