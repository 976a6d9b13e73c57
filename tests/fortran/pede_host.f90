!> A pede-style Fortran host that runs one matrix pass and one solve on the GPU through the
!! ISO_C_BINDING module include/mpgpu_mod.f90 (test of the Fortran side of the boundary).
!!
!! usage: pede_host <C Mille file> <metsol> <matsto> <mreqenf>
!! prints one line "nrec ntot nvar ncon chi2 ndf method istop rel_residual", then one line
!! "label correction" per variable parameter.
PROGRAM pede_host
    USE, INTRINSIC :: iso_c_binding
    USE mpgpu_mod
    IMPLICIT NONE
    TYPE(mpgpu_options) :: opt
    TYPE(mpgpu_pass_stats) :: pst
    TYPE(mpgpu_solve_stats) :: sst
    TYPE(c_ptr) :: ctx
    INTEGER(c_int) :: rc
    INTEGER(c_int32_t) :: ntot, nvar, ncon, i
    INTEGER(c_int64_t) :: nrec
    CHARACTER(len=1024) :: path, arg
    INTEGER :: metsol, matsto, mreqenf
    REAL(c_double), ALLOCATABLE :: x(:)
    INTEGER(c_int32_t), ALLOCATABLE :: labels(:), var_index(:), group_first(:)
    INTEGER(c_int64_t), ALLOCATABLE :: counters(:)

    CALL get_command_argument(1, path)
    CALL get_command_argument(2, arg)
    READ(arg, *) metsol
    CALL get_command_argument(3, arg)
    READ(arg, *) matsto
    CALL get_command_argument(4, arg)
    READ(arg, *) mreqenf

    CALL mpgpu_default_options(opt)
    opt%metsol = INT(metsol, c_int32_t)
    opt%matsto = INT(matsto, c_int32_t)
    opt%mreqenf = INT(mreqenf, c_int32_t)
    opt%mcount = 1_c_int32_t
    rc = mpgpu_create(opt, 0_c_int32_t, ctx)
    CALL check(rc, 'mpgpu_create')
    rc = mpgpu_read_mille_file(ctx, TRIM(path)//c_null_char, 0_c_int32_t, nrec)
    CALL check(rc, 'mpgpu_read_mille_file')
    rc = mpgpu_build_structure(ctx, ntot, nvar, ncon)              ! loop1 / loop2
    CALL check(rc, 'mpgpu_build_structure')
    rc = mpgpu_loopn(ctx, 1_c_int32_t, 1_c_int32_t, 0.0_c_float, pst) ! loopn + loopbf
    CALL check(rc, 'mpgpu_loopn')
    ALLOCATE(x(nvar + ncon))
    rc = mpgpu_solve(ctx, c_null_ptr, x, sst)                        ! the metsol dispatch
    CALL check(rc, 'mpgpu_solve')
    ALLOCATE(labels(ntot), var_index(ntot), group_first(ntot), counters(ntot))
    rc = mpgpu_get_labels(ctx, labels, var_index, counters, group_first)
    CALL check(rc, 'mpgpu_get_labels')
    WRITE(*, '(I0,1X,I0,1X,I0,1X,I0,1X,ES24.16,1X,I0,1X,I0,1X,I0,1X,ES12.4)') nrec, ntot, nvar, ncon, &
        pst%chi2, pst%ndf, sst%method, sst%istop, sst%rel_residual
    DO i = 1, ntot
        IF (var_index(i) > 0) WRITE(*, '(I0,1X,ES24.16)') labels(i), x(var_index(i))
    END DO
    rc = mpgpu_destroy(ctx)
CONTAINS
    SUBROUTINE check(code, what)
        INTEGER(c_int), INTENT(IN) :: code
        CHARACTER(len=*), INTENT(IN) :: what
        IF (code /= 0) THEN
            WRITE(*, '(A,A,I0)') what, ' failed with exit code ', code
            STOP 1
        END IF
    END SUBROUTINE check
END PROGRAM pede_host
