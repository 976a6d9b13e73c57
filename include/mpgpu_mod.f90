!> ISO_C_BINDING interface to libmpgpu (include/mpgpu.h) for a Fortran pede host.
!!
!! A pede-style host (src/pede.f90 with module mpmod) uses these interfaces to replace
!! loopn/loopbf (mpgpu_loopn), mupdat (mpgpu_mupdat), avprod (mpgpu_avprod) and the metsol
!! dispatch of xloopn (mpgpu_solve).  All arguments are passed as C pointers/values; no hidden
!! string lengths and no assumed-shape arrays cross the boundary.
MODULE mpgpu_mod
    USE, INTRINSIC :: iso_c_binding
    IMPLICIT NONE

    TYPE, BIND(C) :: mpgpu_options
        INTEGER(c_int32_t) :: metsol, matsto, mbandw, lprecm, lhuber, lfitbb, mreqenf, mreqena, &
                              mcount, itnlim, withmultipliers, mrmode
        REAL(c_float)      :: chicut, chirem, chhuge, dwcut, cndlmx, regula
        REAL(c_double)     :: default_presigma, mrestl, mrtcnd
        INTEGER(c_int32_t) :: nregul, refine_steps
        INTEGER(c_int32_t) :: monitor, monitor_bins
        INTEGER(c_int32_t) :: lapack_errors
        INTEGER(c_int32_t) :: pair_entries, skip_empty_cons, resolve_redundancy_cons
        REAL(c_double)     :: scale_errors(2)
        REAL(c_double)     :: refine_tol
        INTEGER(c_int32_t) :: fp32_min_n, iteren
    END TYPE mpgpu_options

    TYPE, BIND(C) :: mpgpu_pass_stats
        REAL(c_double)     :: chi2
        INTEGER(c_int64_t) :: ndf
        REAL(c_double)     :: wndf
        INTEGER(c_int64_t) :: nrej(6), naccept, nrec
        REAL(c_double)     :: ms_kernel, ms_total, ms_records
        INTEGER(c_int64_t) :: n_records
        REAL(c_double)     :: ms_gather
        INTEGER(c_int64_t) :: n_gather
    END TYPE mpgpu_pass_stats

    TYPE, BIND(C) :: mpgpu_record_summary
        INTEGER(c_int64_t) :: nrec, neq, neq_glob, nglob_deriv
        INTEGER(c_int32_t) :: fp64_payload, ntot, nvar, ngroup_tot, ngroup_var, ncon, &
                              max_glob_rec, max_loc_rec, max_eq_rec
        INTEGER(c_int64_t) :: nempty_rec
    END TYPE mpgpu_record_summary

    TYPE, BIND(C) :: mpgpu_solve_stats
        INTEGER(c_int32_t) :: method, nrank, istop, itn
        REAL(c_double)     :: anorm, acond, rnorm, arnorm, ynorm, rel_residual, ms
        REAL(c_double)     :: evmax, evmin
    END TYPE mpgpu_solve_stats

    TYPE, BIND(C) :: mpgpu_run_stats
        INTEGER(c_int32_t) :: ntot, nvar, ncon, iterations
        INTEGER(c_int64_t) :: nrec
        REAL(c_double)     :: final_chi2
        INTEGER(c_int64_t) :: final_ndf
        REAL(c_double)     :: ms_structure, ms_passes, ms_solves
        INTEGER(c_int32_t) :: exit_code
    END TYPE mpgpu_run_stats

    ! serving: statistics of one finished job (mpgpu_job_stats)
    TYPE, BIND(C) :: mpgpu_job_stats
        INTEGER(c_int32_t) :: status, nvar
        REAL(c_double)     :: latency_ms, queue_ms, run_ms
        REAL(c_double)     :: chi2
        INTEGER(c_int64_t) :: ndf
        INTEGER(c_int32_t) :: istop, itn
        REAL(c_double)     :: rel_residual
        REAL(c_double)     :: t_arrival_s, t_done_s
        INTEGER(c_int32_t) :: method, reserved
    END TYPE mpgpu_job_stats

    INTERFACE
        SUBROUTINE mpgpu_default_options(opt) BIND(C, name='mpgpu_default_options')
            IMPORT :: mpgpu_options
            TYPE(mpgpu_options), INTENT(OUT) :: opt
        END SUBROUTINE
        INTEGER(c_int) FUNCTION mpgpu_create(opt, device, ctx) BIND(C, name='mpgpu_create')
            IMPORT :: c_int, c_int32_t, c_ptr, mpgpu_options
            TYPE(mpgpu_options), INTENT(IN) :: opt
            INTEGER(c_int32_t), VALUE :: device
            TYPE(c_ptr), INTENT(OUT) :: ctx
        END FUNCTION
        INTEGER(c_int) FUNCTION mpgpu_destroy(ctx) BIND(C, name='mpgpu_destroy')
            IMPORT :: c_int, c_ptr
            TYPE(c_ptr), VALUE :: ctx
        END FUNCTION
        TYPE(c_ptr) FUNCTION mpgpu_last_error(ctx) BIND(C, name='mpgpu_last_error')
            IMPORT :: c_ptr
            TYPE(c_ptr), VALUE :: ctx
        END FUNCTION
        INTEGER(c_int) FUNCTION mpgpu_set_options(ctx, opt) BIND(C, name='mpgpu_set_options')
            IMPORT :: c_int, c_ptr, mpgpu_options
            TYPE(c_ptr), VALUE :: ctx
            TYPE(mpgpu_options), INTENT(IN) :: opt
        END FUNCTION
        INTEGER(c_int) FUNCTION mpgpu_add_records(ctx, nrec, rec_ptr, inder, glder_f, glder_d) &
                BIND(C, name='mpgpu_add_records')
            IMPORT :: c_int, c_ptr, c_int64_t
            TYPE(c_ptr), VALUE :: ctx
            INTEGER(c_int64_t), VALUE :: nrec
            TYPE(c_ptr), VALUE :: rec_ptr, inder, glder_f, glder_d
        END FUNCTION
        INTEGER(c_int) FUNCTION mpgpu_read_mille_file(ctx, path, fortran, nrec) BIND(C, name='mpgpu_read_mille_file')
            IMPORT :: c_int, c_ptr, c_char, c_int32_t, c_int64_t
            TYPE(c_ptr), VALUE :: ctx
            CHARACTER(KIND=c_char), DIMENSION(*), INTENT(IN) :: path
            INTEGER(c_int32_t), VALUE :: fortran
            INTEGER(c_int64_t), INTENT(OUT) :: nrec
        END FUNCTION
        INTEGER(c_int) FUNCTION mpgpu_set_steering_parameters(ctx, n, labels, values, presigma) &
                BIND(C, name='mpgpu_set_steering_parameters')
            IMPORT :: c_int, c_ptr, c_int32_t, c_double
            TYPE(c_ptr), VALUE :: ctx
            INTEGER(c_int32_t), VALUE :: n
            INTEGER(c_int32_t), INTENT(IN) :: labels(*)
            REAL(c_double), INTENT(IN) :: values(*), presigma(*)
        END FUNCTION
        INTEGER(c_int) FUNCTION mpgpu_add_equations(ctx, kind, nblock, ptr, labels, factors, rhs, sigma) &
                BIND(C, name='mpgpu_add_equations')
            IMPORT :: c_int, c_ptr, c_int32_t, c_int64_t, c_double
            TYPE(c_ptr), VALUE :: ctx
            INTEGER(c_int32_t), VALUE :: kind, nblock
            INTEGER(c_int64_t), INTENT(IN) :: ptr(*)
            INTEGER(c_int32_t), INTENT(IN) :: labels(*)
            REAL(c_double), INTENT(IN) :: factors(*), rhs(*), sigma(*)
        END FUNCTION
        INTEGER(c_int) FUNCTION mpgpu_build_structure(ctx, ntot, nvar, ncon) BIND(C, name='mpgpu_build_structure')
            IMPORT :: c_int, c_ptr, c_int32_t
            TYPE(c_ptr), VALUE :: ctx
            INTEGER(c_int32_t), INTENT(OUT) :: ntot, nvar, ncon
        END FUNCTION
        INTEGER(c_int) FUNCTION mpgpu_get_labels(ctx, labels, var_index, counters, group_first_label) &
                BIND(C, name='mpgpu_get_labels')
            IMPORT :: c_int, c_ptr, c_int32_t, c_int64_t
            TYPE(c_ptr), VALUE :: ctx
            INTEGER(c_int32_t), INTENT(OUT) :: labels(*), var_index(*), group_first_label(*)
            INTEGER(c_int64_t), INTENT(OUT) :: counters(*)
        END FUNCTION
        INTEGER(c_int) FUNCTION mpgpu_matrix_info(ctx, stored, nblocks, ngroups) BIND(C, name='mpgpu_matrix_info')
            IMPORT :: c_int, c_ptr, c_int32_t, c_int64_t
            TYPE(c_ptr), VALUE :: ctx
            INTEGER(c_int64_t), INTENT(OUT) :: stored, nblocks
            INTEGER(c_int32_t), INTENT(OUT) :: ngroups
        END FUNCTION
        INTEGER(c_int) FUNCTION mpgpu_get_tuning(ctx, text, len) BIND(C, name='mpgpu_get_tuning')
            IMPORT :: c_int, c_ptr, c_int32_t, c_char
            TYPE(c_ptr), VALUE :: ctx
            CHARACTER(KIND=c_char), INTENT(OUT) :: text(*)
            INTEGER(c_int32_t), VALUE :: len
        END FUNCTION
        INTEGER(c_int) FUNCTION mpgpu_get_record_summary(ctx, summary) BIND(C, name='mpgpu_get_record_summary')
            IMPORT :: c_int, c_ptr, mpgpu_record_summary
            TYPE(c_ptr), VALUE :: ctx
            TYPE(mpgpu_record_summary), INTENT(OUT) :: summary
        END FUNCTION
        INTEGER(c_int) FUNCTION mpgpu_set_parameters(ctx, par) BIND(C, name='mpgpu_set_parameters')
            IMPORT :: c_int, c_ptr, c_double
            TYPE(c_ptr), VALUE :: ctx
            REAL(c_double), INTENT(IN) :: par(*)
        END FUNCTION
        INTEGER(c_int) FUNCTION mpgpu_get_parameters(ctx, par) BIND(C, name='mpgpu_get_parameters')
            IMPORT :: c_int, c_ptr, c_double
            TYPE(c_ptr), VALUE :: ctx
            REAL(c_double), INTENT(OUT) :: par(*)
        END FUNCTION
        INTEGER(c_int) FUNCTION mpgpu_loopn(ctx, icalcm, nloopn, chicut, st) BIND(C, name='mpgpu_loopn')
            IMPORT :: c_int, c_ptr, c_int32_t, c_float, mpgpu_pass_stats
            TYPE(c_ptr), VALUE :: ctx
            INTEGER(c_int32_t), VALUE :: icalcm, nloopn
            REAL(c_float), VALUE :: chicut
            TYPE(mpgpu_pass_stats), INTENT(OUT) :: st
        END FUNCTION
        INTEGER(c_int) FUNCTION mpgpu_get_vector(ctx, vec, counter) BIND(C, name='mpgpu_get_vector')
            IMPORT :: c_int, c_ptr, c_double, c_int64_t
            TYPE(c_ptr), VALUE :: ctx
            REAL(c_double), INTENT(OUT) :: vec(*)
            INTEGER(c_int64_t), INTENT(OUT) :: counter(*)
        END FUNCTION
        INTEGER(c_int) FUNCTION mpgpu_get_matrix_packed(ctx, packed, len) BIND(C, name='mpgpu_get_matrix_packed')
            IMPORT :: c_int, c_ptr, c_double, c_int64_t
            TYPE(c_ptr), VALUE :: ctx
            REAL(c_double), INTENT(OUT) :: packed(*)
            INTEGER(c_int64_t), VALUE :: len
        END FUNCTION
        INTEGER(c_int) FUNCTION mpgpu_mupdat(ctx, nel, i, j, v) BIND(C, name='mpgpu_mupdat')
            IMPORT :: c_int, c_ptr, c_int32_t, c_double
            TYPE(c_ptr), VALUE :: ctx
            INTEGER(c_int32_t), VALUE :: nel
            INTEGER(c_int32_t), INTENT(IN) :: i(*), j(*)
            REAL(c_double), INTENT(IN) :: v(*)
        END FUNCTION
        INTEGER(c_int) FUNCTION mpgpu_solve(ctx, rhs, x, st) BIND(C, name='mpgpu_solve')
            IMPORT :: c_int, c_ptr, c_double, mpgpu_solve_stats
            TYPE(c_ptr), VALUE :: ctx
            TYPE(c_ptr), VALUE :: rhs
            REAL(c_double), INTENT(OUT) :: x(*)
            TYPE(mpgpu_solve_stats), INTENT(OUT) :: st
        END FUNCTION
        INTEGER(c_int) FUNCTION mpgpu_avprod(ctx, x, y) BIND(C, name='mpgpu_avprod')
            IMPORT :: c_int, c_ptr, c_double
            TYPE(c_ptr), VALUE :: ctx
            REAL(c_double), INTENT(IN) :: x(*)
            REAL(c_double), INTENT(OUT) :: y(*)
        END FUNCTION
        ! y = A x with x, y device pointers, allocating nothing (full storage, nall <= 16384)
        INTEGER(c_int) FUNCTION mpgpu_avprod_noalloc(ctx, x_dev, y_dev) BIND(C, name='mpgpu_avprod_noalloc')
            IMPORT :: c_int, c_ptr
            TYPE(c_ptr), VALUE :: ctx, x_dev, y_dev
        END FUNCTION
        ! group-block structure of the sparse matrix (sizes from mpgpu_matrix_info)
        INTEGER(c_int) FUNCTION mpgpu_get_block_structure(ctx, gfirst, brow, bcol, boff) &
                BIND(C, name='mpgpu_get_block_structure')
            IMPORT :: c_int, c_ptr, c_int32_t, c_int64_t
            TYPE(c_ptr), VALUE :: ctx
            INTEGER(c_int32_t), INTENT(OUT) :: gfirst(*), bcol(*)
            INTEGER(c_int64_t), INTENT(OUT) :: brow(*), boff(*)
        END FUNCTION
        INTEGER(c_int) FUNCTION mpgpu_get_errors(ctx, err, gcor) BIND(C, name='mpgpu_get_errors')
            IMPORT :: c_int, c_ptr, c_double
            TYPE(c_ptr), VALUE :: ctx
            REAL(c_double), INTENT(OUT) :: err(*), gcor(*)
        END FUNCTION
        INTEGER(c_int) FUNCTION mpgpu_get_monitor(ctx, ngroups, labels, mean_err, hist) BIND(C, name='mpgpu_get_monitor')
            IMPORT :: c_int, c_ptr, c_int32_t, c_int64_t, c_float
            TYPE(c_ptr), VALUE :: ctx
            INTEGER(c_int32_t), INTENT(OUT) :: ngroups
            INTEGER(c_int32_t), INTENT(OUT) :: labels(*)
            REAL(c_float), INTENT(OUT) :: mean_err(*)
            INTEGER(c_int64_t), INTENT(OUT) :: hist(*)
        END FUNCTION
        INTEGER(c_int) FUNCTION mpgpu_run_steering(steer, outdir, device, st) BIND(C, name='mpgpu_run_steering')
            IMPORT :: c_int, c_char, c_int32_t, mpgpu_run_stats
            CHARACTER(KIND=c_char), DIMENSION(*), INTENT(IN) :: steer, outdir
            INTEGER(c_int32_t), VALUE :: device
            TYPE(mpgpu_run_stats), INTENT(OUT) :: st
        END FUNCTION
        ! serving: concurrent jobs on one GPU (no pede counterpart)
        INTEGER(c_int) FUNCTION mpgpu_server_create(device, nworkers, budget_bytes, batch_max_params, srv) &
                BIND(C, name='mpgpu_server_create')
            IMPORT :: c_int, c_int32_t, c_int64_t, c_ptr
            INTEGER(c_int32_t), VALUE :: device, nworkers
            INTEGER(c_int64_t), VALUE :: budget_bytes
            INTEGER(c_int32_t), VALUE :: batch_max_params
            TYPE(c_ptr), INTENT(OUT) :: srv
        END FUNCTION
        INTEGER(c_int) FUNCTION mpgpu_server_submit(srv, opt, nrec, rec_ptr, inder, glder, npar, plab, pval, psig, &
                ncon, cptr, clab, cfac, crhs, job_id) BIND(C, name='mpgpu_server_submit')
            IMPORT :: c_int, c_ptr, c_int32_t, c_int64_t, c_float, c_double, mpgpu_options
            TYPE(c_ptr), VALUE :: srv
            TYPE(mpgpu_options), INTENT(IN) :: opt
            INTEGER(c_int64_t), VALUE :: nrec
            INTEGER(c_int64_t), INTENT(IN) :: rec_ptr(*)
            INTEGER(c_int32_t), INTENT(IN) :: inder(*)
            REAL(c_float), INTENT(IN) :: glder(*)
            INTEGER(c_int32_t), VALUE :: npar
            INTEGER(c_int32_t), INTENT(IN) :: plab(*)
            REAL(c_double), INTENT(IN) :: pval(*), psig(*)
            INTEGER(c_int32_t), VALUE :: ncon
            INTEGER(c_int64_t), INTENT(IN) :: cptr(*)
            INTEGER(c_int32_t), INTENT(IN) :: clab(*)
            REAL(c_double), INTENT(IN) :: cfac(*), crhs(*)
            INTEGER(c_int64_t), INTENT(OUT) :: job_id
        END FUNCTION
        INTEGER(c_int) FUNCTION mpgpu_server_wait(srv, job_id, capacity, labels, corr, st) &
                BIND(C, name='mpgpu_server_wait')
            IMPORT :: c_int, c_ptr, c_int32_t, c_int64_t, c_double, mpgpu_job_stats
            TYPE(c_ptr), VALUE :: srv
            INTEGER(c_int64_t), VALUE :: job_id
            INTEGER(c_int32_t), VALUE :: capacity
            INTEGER(c_int32_t), INTENT(OUT) :: labels(*)
            REAL(c_double), INTENT(OUT) :: corr(*)
            TYPE(mpgpu_job_stats), INTENT(OUT) :: st
        END FUNCTION
        INTEGER(c_int) FUNCTION mpgpu_server_destroy(srv) BIND(C, name='mpgpu_server_destroy')
            IMPORT :: c_int, c_ptr
            TYPE(c_ptr), VALUE :: srv
        END FUNCTION
    END INTERFACE
END MODULE mpgpu_mod
