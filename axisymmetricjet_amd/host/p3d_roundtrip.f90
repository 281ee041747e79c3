! Reads a PLOT3D grid and solution with the host's reader and writes the solution back with
! its writer (used by the CPU tests to check the file format against the Python writer).
!   p3d_roundtrip <grid> <soln-in> <soln-out>
PROGRAM p3d_roundtrip
  USE, INTRINSIC :: ISO_FORTRAN_ENV, ONLY : INT32, REAL64
  USE ajet_plot3d
  IMPLICIT NONE
  CHARACTER(LEN=512) :: g, si, so
  INTEGER :: nz, nr, nz2, nr2
  REAL(REAL64), ALLOCATABLE :: z(:,:), r(:,:), q(:,:,:)
  INTEGER(INT32), ALLOCATABLE :: ib(:,:)
  REAL(REAL64) :: tau(4)
  CALL GET_COMMAND_ARGUMENT(1, g)
  CALL GET_COMMAND_ARGUMENT(2, si)
  CALL GET_COMMAND_ARGUMENT(3, so)
  CALL p3d_read_grid(TRIM(g), nz, nr, z, r, ib)
  CALL p3d_read_soln(TRIM(si), nz2, nr2, q, tau)
  IF (nz /= nz2 .OR. nr /= nr2) STOP 2
  WRITE (*, '(I0,1X,I0,1X,I0,3(1X,ES24.16))') nz, nr, SUM(ib), SUM(z), SUM(r), SUM(q)
  CALL p3d_write_soln(TRIM(so), nz, nr, q, tau)
END PROGRAM p3d_roundtrip
