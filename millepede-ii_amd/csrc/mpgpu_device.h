// Device-side data model shared by the HIP kernels and the host code of libmpgpu.
//
// Records are compiled once (loop2 time) into a structure-of-arrays layout that keeps the
// Mille words as they are (values as FP32 or FP64, indices as int32) and adds per-equation
// and per-record offsets, so that the per-pass kernel never scans for zero markers again.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mpgpu {

// ---- compiled records (one batch) -----------------------------------------------------
struct DevRecords {
    int64_t nrec;
    const int64_t *rec_eq;   // [nrec+1] equation range of each record
    const int64_t *rec_gl;   // [nrec+1] range in gl_tix (distinct labels of the record, sorted)
    const int64_t *rec_pb;   // [nrec+1] range in pair_blk (sparse storage only)
    const int64_t *eq_sl;    // word index of the measurement word (the local words follow)
    const uint32_t *eq_n;    // number of local words | number of global words << 16
    const float *vf;         // word values (FP32 payload) or nullptr
    const double *vd;        // word values (FP64 payload) or nullptr
    const int32_t *widx;     // local words: local index; global words: position in gl list
    const int32_t *gl_tix;   // total index (0-based) of each distinct label of a record
    const int32_t *pair_blk; // block id per (segment_i >= segment_j) pair of a record
    const float *rec_w;      // per-record (file) weight, nullptr = 1
    const int64_t *order;    // processing order of the records (nullptr = file order)
    const int32_t *band;     // per record: 0 full local fit, else bordered-band structure
                             // (mbnd+1) | mbdr << 10 | mside << 20 (record kernel skips these)
    const int64_t *rec_word; // [nrec+1] word range of each record
    const int64_t *rec_cb;   // group-block storage: offset of the record's update in cbuf (-1: none)
    const int32_t *rec_cn;   //   number of update values of the record
    double *cbuf;            //   update buffer of the current batch (two-phase matrix build)
    int32_t max_gl, max_eq, max_nalc, max_var, max_seg;
};

// ---- global model -----------------------------------------------------------------------
struct DevModel {
    const double *par;        // globalParameter by total index
    const int32_t *var;       // total index -> variable index (1-based) or <= 0
    int32_t nvar;             // variable parameters
    int32_t sparse;           // 0: packed full, 1: group-block sparse
    double *mat;              // packed (nall(nall+1)/2) or block values
    const int32_t *group_of;  // [nvar+1] variable index (1-based) -> group (0-based)
    const int32_t *grp_tot;   // [ntot] total index -> group of the variable (-1: fixed)
    const int32_t *gfirst;    // [ngroup+1] first variable index (1-based) of each group
    const int64_t *brow;      // [ngroup+1] block-row pointers (lower triangle incl. diagonal)
    const int32_t *bcol;      // [nblock] column group of each block (ascending in a row)
    const int64_t *boff;      // [nblock] offset of each block's values (row-major r x c)
    double *vec;              // globalVector (nall)
    unsigned long long *cnt;  // globalCounter (nvar)
};

struct PassOpts {
    int32_t icalcm, nloopn, lhuber, lfitbb;
    float chicut, chhuge, dwcut, cndlmx;
    double esc[2];   // errorScaleFactor: measurement errors with / without global derivatives
};

// residual monitoring (pede's measHists): per total index the measurement group of a
// measurement whose first global derivative it is (-1: none), the group's mean error, and
// bins counts per group
struct MonCtx {
    const int32_t *grp;
    const double *res;
    unsigned int *hist;
    int32_t bins;
    int32_t mode;   // 0 off, 1 normalized residuals, 2 pulls
};

// per-record results of one pass
struct RecOut {
    double *chi2;    // contribution to Sum(chi^2)
    int32_t *ndf;    // contribution to Sum(ndf)
    int8_t *status;  // 0 accepted, 1..6 pede reject classes, 7 no local parameters,
                     // 8 needs bordered-band local fit, 9 exceeds kernel limits, 20 bad record
};

// LDS plan of the record kernel (all counts in elements)
struct LdsPlan {
    int32_t max_gl, max_eq, nalc_max, max_var, max_pair;
    int32_t qe;      // equations kept per record variable (1..8)
    int32_t gh;      // (unused)
    int32_t stage;   // 1: group-block storage (update-buffer row descriptors in LDS)
    int32_t max_w;   // words of the largest record (staged in LDS)
    int32_t max_grp; // parameters of the largest group
    int32_t bytes;
};

// LDS plan of the bordered-band record kernel
struct BandPlan {
    int32_t max_gl, max_eq, nalc_max, max_var, max_pair, mp1_max, nbdr_max, max_sup;
    int64_t vinv_stride;     // doubles of global scratch per workgroup (packed local inverse)
    int32_t bytes;
};

// MINRES recurrence state in device memory (src/minresModule.f90:294-660): the dot products of an
// iteration land here, one thread advances the scalar recurrence and the vector kernels read
// their coefficients from it; once istop is set the remaining launched iterations do nothing
struct MrState {
    double beta1, beta, oldb, alfa, bnew, dbar, epsln, oldeps, qrnorm, phibar, rhs1, rhs2, tnorm2, ynorm2, cs, sn;
    double gmin, gmax, Anorm, Acond, rnorm, ynorm, Arnorml, delta, denom, phi, rtol;
    int itn, itnlim, istop, upd;
};

// scalars of one MINRES-QLP vector step (k_qlp_step)
struct QlpStep {
    int first, itn;
    double gamal3, veplnl2, etal, gamal_qlp, vepln_qlp, gama_qlp, ul_qlp, u_qlp, sr1, cr1, sr2, cr2, ul2, ul, u;
};

} // namespace mpgpu
