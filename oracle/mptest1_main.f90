! Driver for the reference's wire-chamber generator (subroutine mptest, src/mptest1.f90:81).
! Writes mp2tst.bin (Fortran sequential Mille records), mp2con.txt and mp2str.txt into the
! current directory.  Usage: mptest1_gen [nEvents]   (default 10000, as pede's -n default)
PROGRAM mptest1_gen
    IMPLICIT NONE
    INTEGER :: n, ios
    CHARACTER(LEN=32) :: arg
    n = 10000
    IF (COMMAND_ARGUMENT_COUNT() >= 1) THEN
        CALL GET_COMMAND_ARGUMENT(1, arg)
        READ(arg, *, IOSTAT=ios) n
    END IF
    CALL mptest(n)
END PROGRAM mptest1_gen
