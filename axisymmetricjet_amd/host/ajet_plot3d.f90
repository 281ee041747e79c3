! PLOT3D grid and solution files in the layout the reference reads and writes: Fortran
! sequential unformatted records with 4-byte length markers, multi-grid ('m'), whole ('w'),
! double precision ('d'), NDIM = 3 with ND(1,3) = 1, grids with IBLANK
! (ModPLOT3D_IO.f90:123-294 Read_Grid, :434-537 Read_Soln, :539-599 Write_Soln).  The files
! are parsed through stream access so that the record markers are checked, as p3d_detect_
! does when it sniffs a file (plot3d_format.c:16-163).
MODULE ajet_plot3d
  USE, INTRINSIC :: ISO_FORTRAN_ENV, ONLY : INT32, INT64, REAL64
  IMPLICIT NONE
  PRIVATE
  PUBLIC :: p3d_read_grid, p3d_read_soln, p3d_write_soln

CONTAINS

  SUBROUTINE fail(msg, fname)
    CHARACTER(LEN=*), INTENT(IN) :: msg, fname
    WRITE (*, '(A,A,A)') 'PLOT3D: ', msg, ' '//TRIM(fname)
    STOP 1
  END SUBROUTINE fail

  ! record marker pair around a payload of nbytes, checked
  SUBROUTINE open_record(u, nbytes, fname, pos)
    INTEGER, INTENT(IN) :: u
    INTEGER(INT64), INTENT(IN) :: nbytes
    CHARACTER(LEN=*), INTENT(IN) :: fname
    INTEGER(INT64), INTENT(INOUT) :: pos
    INTEGER(INT32) :: m
    INTEGER :: ios
    READ (u, POS=pos, IOSTAT=ios) m
    IF (ios /= 0) CALL fail('file ends inside a record:', fname)
    IF (INT(m, INT64) /= nbytes) CALL fail('unexpected record length in', fname)
    pos = pos + 4
  END SUBROUTINE open_record

  SUBROUTINE close_record(u, nbytes, fname, pos)
    INTEGER, INTENT(IN) :: u
    INTEGER(INT64), INTENT(IN) :: nbytes
    CHARACTER(LEN=*), INTENT(IN) :: fname
    INTEGER(INT64), INTENT(INOUT) :: pos
    INTEGER(INT32) :: m
    INTEGER :: ios
    READ (u, POS=pos + nbytes, IOSTAT=ios) m
    IF (ios /= 0) CALL fail('file ends inside a record:', fname)
    IF (INT(m, INT64) /= nbytes) CALL fail('record marker mismatch in', fname)
    pos = pos + nbytes + 4
  END SUBROUTINE close_record

  SUBROUTINE read_header(u, fname, pos, nz, nr)
    INTEGER, INTENT(IN) :: u
    CHARACTER(LEN=*), INTENT(IN) :: fname
    INTEGER(INT64), INTENT(INOUT) :: pos
    INTEGER, INTENT(OUT) :: nz, nr
    INTEGER(INT32) :: ngrid, nd(3)
    INTEGER :: ios
    CALL open_record(u, 4_INT64, fname, pos)
    READ (u, POS=pos, IOSTAT=ios) ngrid
    IF (ios /= 0) CALL fail('file ends inside a record:', fname)
    CALL close_record(u, 4_INT64, fname, pos)
    IF (ngrid /= 1) CALL fail('only single-block files are supported:', fname)
    CALL open_record(u, 12_INT64, fname, pos)
    READ (u, POS=pos, IOSTAT=ios) nd
    IF (ios /= 0) CALL fail('file ends inside a record:', fname)
    CALL close_record(u, 12_INT64, fname, pos)
    IF (nd(3) /= 1) CALL fail('expected ND(1,3) = 1 in', fname)
    nz = nd(1)
    nr = nd(2)
  END SUBROUTINE read_header

  ! Grid: z, r (the x and y coordinates) and IBLANK (all 1 when the file has none).
  SUBROUTINE p3d_read_grid(fname, nz, nr, z, r, iblank)
    CHARACTER(LEN=*), INTENT(IN) :: fname
    INTEGER, INTENT(OUT) :: nz, nr
    REAL(REAL64), ALLOCATABLE, INTENT(OUT) :: z(:,:), r(:,:)
    INTEGER(INT32), ALLOCATABLE, INTENT(OUT) :: iblank(:,:)
    REAL(REAL64), ALLOCATABLE :: w(:,:)
    INTEGER :: u, ios
    INTEGER(INT64) :: pos, npts, nb
    INTEGER(INT32) :: m
    LOGICAL :: ex
    INQUIRE (FILE=fname, EXIST=ex)
    IF (.NOT. ex) CALL fail('File is not readable:', fname)
    OPEN (NEWUNIT=u, FILE=fname, ACCESS='stream', FORM='unformatted', STATUS='old', IOSTAT=ios)
    IF (ios /= 0) CALL fail('cannot open', fname)
    pos = 1
    CALL read_header(u, fname, pos, nz, nr)
    npts = INT(nz, INT64) * nr
    READ (u, POS=pos, IOSTAT=ios) m
    IF (ios /= 0) CALL fail('file ends inside a record:', fname)
    nb = INT(m, INT64)
    ALLOCATE (z(nz, nr), r(nz, nr), w(nz, nr), iblank(nz, nr))
    IF (nb == 3 * npts * 8 + npts * 4) THEN
      pos = pos + 4
      READ (u, POS=pos, IOSTAT=ios) z, r, w, iblank
      IF (ios /= 0) CALL fail('file ends inside a record:', fname)
      CALL close_record(u, nb, fname, pos)
    ELSE IF (nb == 3 * npts * 8) THEN
      pos = pos + 4
      READ (u, POS=pos, IOSTAT=ios) z, r, w
      IF (ios /= 0) CALL fail('file ends inside a record:', fname)
      CALL close_record(u, nb, fname, pos)
      iblank = 1
    ELSE
      CALL fail('not a whole double-precision 3-D grid:', fname)
    END IF
    CLOSE (u)
  END SUBROUTINE p3d_read_grid

  ! Solution: tau(1:4) and the five variables (here the primitives rho, Vr, Vtheta, Vz, T).
  SUBROUTINE p3d_read_soln(fname, nz, nr, q, tau)
    CHARACTER(LEN=*), INTENT(IN) :: fname
    INTEGER, INTENT(OUT) :: nz, nr
    REAL(REAL64), ALLOCATABLE, INTENT(OUT) :: q(:,:,:)
    REAL(REAL64), INTENT(OUT) :: tau(4)
    INTEGER :: u, ios
    INTEGER(INT64) :: pos, nb
    LOGICAL :: ex
    INQUIRE (FILE=fname, EXIST=ex)
    IF (.NOT. ex) CALL fail('File is not readable:', fname)
    OPEN (NEWUNIT=u, FILE=fname, ACCESS='stream', FORM='unformatted', STATUS='old', IOSTAT=ios)
    IF (ios /= 0) CALL fail('cannot open', fname)
    pos = 1
    CALL read_header(u, fname, pos, nz, nr)
    CALL open_record(u, 32_INT64, fname, pos)
    READ (u, POS=pos, IOSTAT=ios) tau
    IF (ios /= 0) CALL fail('file ends inside a record:', fname)
    CALL close_record(u, 32_INT64, fname, pos)
    ALLOCATE (q(nz, nr, 5))
    nb = INT(nz, INT64) * nr * 5 * 8
    CALL open_record(u, nb, fname, pos)
    READ (u, POS=pos, IOSTAT=ios) q
    IF (ios /= 0) CALL fail('file ends inside a record:', fname)
    CALL close_record(u, nb, fname, pos)
    CLOSE (u)
  END SUBROUTINE p3d_read_soln

  ! The file Write_Soln writes for NDIM = 3, 'm', 'w', 'd' (ModPLOT3D_IO.f90:556-575).
  SUBROUTINE p3d_write_soln(fname, nz, nr, q, tau)
    CHARACTER(LEN=*), INTENT(IN) :: fname
    INTEGER, INTENT(IN) :: nz, nr
    REAL(REAL64), INTENT(IN) :: q(nz, nr, 5)
    REAL(REAL64), INTENT(IN) :: tau(4)
    INTEGER :: u
    INTEGER(INT32) :: nb
    OPEN (NEWUNIT=u, FILE=fname, ACCESS='stream', FORM='unformatted', STATUS='replace')
    WRITE (u) 4_INT32, 1_INT32, 4_INT32
    WRITE (u) 12_INT32, INT(nz, INT32), INT(nr, INT32), 1_INT32, 12_INT32
    WRITE (u) 32_INT32, tau, 32_INT32
    nb = INT(INT(nz, INT64) * nr * 5 * 8, INT32)
    WRITE (u) nb, q, nb
    CLOSE (u)
  END SUBROUTINE p3d_write_soln

END MODULE ajet_plot3d
