! Fortran host of the MI355X RK4 step.  It keeps the reference's driver structure and entry
! points -- PROGRAM ModSolver, ComputeTimeStep and RK4 as argument-less procedures over module
! state (ModSolver.f90:19-125, ModInputOutput.f90:389, ModRungeKutta.f90:5) -- the bc-file
! input (ModPatch.f90:30-49) and the PLOT3D restart/output files (ModInputOutput.f90:201-263).
! The arithmetic of the step runs in libajet.so on the GPU; the state stays resident in HBM
! and only output steps copy it back.

! Run-time configuration and state shared by the entry points (the part of ModuleVariables the
! driver needs).  Defaults are the reference's constants.
MODULE ModuleVariables
  USE, INTRINSIC :: ISO_C_BINDING
  USE, INTRINSIC :: ISO_FORTRAN_ENV, ONLY : INT32, INT64, REAL64
  USE ajet_capi
  IMPLICIT NONE
  CHARACTER(LEN=256) :: GRID_FILENAME = 'axijet_mesh.xyz'
  CHARACTER(LEN=256) :: INITIAL_CONDITION = 'Restart.q'
  CHARACTER(LEN=256) :: TARGET_FILE = 'axijet_soln.q'
  CHARACTER(LEN=256) :: bc_file = 'bc.dat'
  CHARACTER(LEN=256) :: out_prefix = ''
  REAL(REAL64) :: Re = 1000.0D0, Pr = 0.72D0, CFL = 0.5D0, gam = 1.4D0
  REAL(REAL64) :: C_mu = 0.0D0, C_beta = 0.0D0, C_kappa = 0.0D0
  INTEGER :: ShockCapturing = 0, do_viscous_terms = 1
  ! ControlForcing = 1 adds the feedback forcing of ModControlForcing.f90 read from forcing_file
  INTEGER :: ControlForcing = 0
  CHARACTER(LEN=256) :: forcing_file = 'Vz_Control_Vr_Feedback.dat'
  ! LearnedLAD = 1: the learned LAD coefficient (ajet_set_learned_lad) with the network of
  ! mlp_file, a raw little-endian copy of the ajet_mlp struct
  INTEGER :: LearnedLAD = 0, mlp_refresh = 12
  CHARACTER(LEN=256) :: mlp_file = 'learned_lad.bin'
  REAL(REAL64) :: sensor_threshold = 0.1D0, ramp_min = 1.0D-8
  INTEGER :: noutput = 200
  INTEGER(INT64) :: no_iterations = 10000
  INTEGER(INT64) :: iteration, iteration_start
  REAL(REAL64) :: time, time_start, delt, tau(4)
  INTEGER :: Nz, Nr, nlogged = 0
  REAL(REAL64), ALLOCATABLE :: zg(:,:), rg(:,:), pv(:,:,:), pv_target(:,:,:), X0rho(:,:)
  INTEGER(INT32), ALLOCATABLE :: iblank(:,:)
  TYPE(ajet_bc_t), ALLOCATABLE :: bclines(:)
  TYPE(C_PTR) :: ctx = C_NULL_PTR
END MODULE ModuleVariables

MODULE ModDecomp
  USE ModuleVariables
  IMPLICIT NONE
CONTAINS
  ! StopTheCode (ModDecomp.f90:109-119): release the device and stop.
  SUBROUTINE StopTheCode
    INTEGER(C_INT) :: rc
    IF (C_ASSOCIATED(ctx)) rc = ajet_destroy(ctx)
    ctx = C_NULL_PTR
    STOP
  END SUBROUTINE StopTheCode

  SUBROUTINE check(rc, what)
    INTEGER(C_INT), INTENT(IN) :: rc
    CHARACTER(LEN=*), INTENT(IN) :: what
    IF (rc == AJET_EBLOWUP) THEN
      PRINT *, "Time step blow up"
      CALL StopTheCode
    ELSE IF (rc /= AJET_OK) THEN
      PRINT *, 'ajet: ', what, ' failed (', rc, '): ', TRIM(ajet_last_error(ctx))
      CALL StopTheCode
    END IF
  END SUBROUTINE check
END MODULE ModDecomp

! Reading the bc file (ModPatch.f90:30-49): 9 integers per line until the first line that
! does not read; the patches themselves are built by the library (ajet_build_patches rules).
MODULE ModPatch
  USE ModuleVariables
  IMPLICIT NONE
CONTAINS
  SUBROUTINE CreateAllPatches
    INTEGER :: u, reason, n, k
    INTEGER(INT64) :: v(9)
    OPEN (NEWUNIT=u, FILE=TRIM(bc_file), STATUS='old')
    n = 0
    DO
      READ (u, *, IOSTAT=reason) v
      IF (reason /= 0) EXIT
      n = n + 1
    END DO
    ALLOCATE (bclines(MAX(n, 1)))
    REWIND (u)
    DO k = 1, n
      READ (u, *) v
      bclines(k) = ajet_bc_t(v(1), v(2), v(3), v(4), v(5), v(6), v(7))
    END DO
    CLOSE (u)
  END SUBROUTINE CreateAllPatches
END MODULE ModPatch

MODULE ModInputOutput
  USE ModuleVariables
  USE ModDecomp
  USE ajet_plot3d
  IMPLICIT NONE
CONTAINS
  SUBROUTINE GridRead
    CALL p3d_read_grid(TRIM(GRID_FILENAME), Nz, Nr, zg, rg, iblank)
    PRINT *, "Read grid done"
  END SUBROUTINE GridRead

  ! SolnRead (ModInputOutput.f90:91-148): the target first, then the initial condition whose
  ! tau carries the iteration (tau(1)) and the time (tau(4)).
  SUBROUTINE SolnRead
    INTEGER :: n1, n2
    REAL(REAL64) :: tdum(4)
    CALL p3d_read_soln(TRIM(TARGET_FILE), n1, n2, pv_target, tdum)
    IF (n1 /= Nz .OR. n2 /= Nr) CALL fail_size(TARGET_FILE)
    CALL p3d_read_soln(TRIM(INITIAL_CONDITION), n1, n2, pv, tau)
    IF (n1 /= Nz .OR. n2 /= Nr) CALL fail_size(INITIAL_CONDITION)
    ALLOCATE (X0rho(Nz, Nr))
    X0rho = pv(:, :, 1)
    PRINT *, "DONE READING SOLUTION"
  END SUBROUTINE SolnRead

  ! The feedback matrix and its Gaussian window, in the layout ControlForcingTerms reads
  ! (ModControlForcing.f90:16-36): forcing_mat(i, j) row by row, then z0, r0, lz, lr.
  SUBROUTINE ForcingRead(f)
    TYPE(ajet_forcing_t), INTENT(OUT) :: f
    REAL(REAL64) :: fm(5, 5)
    INTEGER :: u, i, j
    OPEN (NEWUNIT=u, FILE=TRIM(forcing_file), STATUS='old')
    DO i = 1, 5
      DO j = 1, 5
        READ (u, *) fm(i, j)
      END DO
    END DO
    READ (u, *) f%z0
    READ (u, *) f%r0
    READ (u, *) f%lz
    READ (u, *) f%lr
    CLOSE (u)
    DO i = 1, 5
      DO j = 1, 5
        PRINT *, i, j, fm(i, j)
      END DO
    END DO
    PRINT *, f%z0, f%r0, f%lz, f%lr
    f%F = TRANSPOSE(fm)
    f%amp = 0.0D0
    f%strouhal = 0.0D0
    f%uj = 0.0D0
    f%t0 = 0.0D0
  END SUBROUTINE ForcingRead

  ! The network of the learned LAD coefficient from mlp_file (the bytes of one ajet_mlp).
  SUBROUTINE MlpRead(m)
    TYPE(ajet_mlp_t), INTENT(OUT) :: m
    INTEGER :: u, ios
    OPEN (NEWUNIT=u, FILE=TRIM(mlp_file), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='old', IOSTAT=ios)
    IF (ios /= 0) THEN
      PRINT *, 'cannot open ', TRIM(mlp_file)
      CALL StopTheCode
    END IF
    READ (u, IOSTAT=ios) m
    IF (ios /= 0) THEN
      PRINT *, 'cannot read the network from ', TRIM(mlp_file)
      CALL StopTheCode
    END IF
    CLOSE (u)
  END SUBROUTINE MlpRead

  SUBROUTINE fail_size(f)
    CHARACTER(LEN=*), INTENT(IN) :: f
    PRINT *, 'solution size differs from the grid: ', TRIM(f)
    CALL StopTheCode
  END SUBROUTINE fail_size

  ! The time-step lines RK4 prints on rank 0 (ModRungeKutta.f90:152-157), for the steps
  ! taken since the last call; the time steps stay on the device between outputs.
  SUBROUTINE PrintSteps
    REAL(REAL64), ALLOCATABLE, TARGET :: dts(:)
    REAL(REAL64), TARGET :: t
    INTEGER(C_INT) :: n
    INTEGER :: k
    REAL(REAL64) :: tt
    n = ajet_get_dt_log(ctx, C_NULL_PTR, 0_C_INT32_T, C_LOC(t))
    IF (n < 0) CALL check(n, 'ajet_get_dt_log')
    ! the device log keeps the first AJET_DTLOG_CAP time steps
    IF (n > AJET_DTLOG_CAP .AND. nlogged < AJET_DTLOG_CAP) &
      PRINT *, 'time-step log full: no Iteration lines past ', iteration_start + AJET_DTLOG_CAP
    n = MIN(n, AJET_DTLOG_CAP)
    IF (n <= nlogged) RETURN
    ALLOCATE (dts(n))
    n = ajet_get_dt_log(ctx, C_LOC(dts), INT(n, C_INT32_T), C_NULL_PTR)
    tt = time_start
    DO k = 1, nlogged
      tt = tt + dts(k)
    END DO
    DO k = nlogged + 1, n
      tt = tt + dts(k)
      WRITE (*, 201) iteration_start + k, dts(k), tt
    END DO
    nlogged = n
    delt = dts(n)
201 FORMAT("Iteration=",I0,2X, "dt=",E17.9, 2X, "time=",E17.9)
  END SUBROUTINE PrintSteps

  ! WriteOutput (ModInputOutput.f90:201-263): 'RocFlo-CM.########.q' with tau(1) = iteration
  ! and, after the first step, tau(4) = time.
  SUBROUTINE WriteOutput
    CHARACTER(LEN=512) :: fname
    REAL(REAL64), TARGET :: t
    INTEGER(C_INT) :: rc
    CALL check(ajet_get_state(ctx, pv), 'ajet_get_state')
    rc = ajet_get_dt_log(ctx, C_NULL_PTR, 0_C_INT32_T, C_LOC(t))
    IF (rc < 0) CALL check(rc, 'ajet_get_dt_log')
    time = time_start + t
    tau(1) = REAL(iteration, REAL64)
    IF (iteration /= iteration_start) tau(4) = time
    PRINT *, "TOTAL TIME=", tau(1), iteration, iteration_start, tau(4)
    WRITE (fname, '(A,A10,I8.8,A2)') TRIM(out_prefix), 'RocFlo-CM.', iteration, '.q'
    CALL p3d_write_soln(TRIM(fname), Nz, Nr, pv, tau)
    PRINT *, "Density min/max", MINVAL(pv(:, :, 1)), MAXVAL(X0rho)
  END SUBROUTINE WriteOutput

  ! ComputeTimeStep (ModInputOutput.f90:389-438): the CFL reduction runs on the device and
  ! the time step stays there for RK4.
  SUBROUTINE ComputeTimeStep
    CALL check(ajet_compute_dt(ctx, C_NULL_PTR), 'ComputeTimeStep')
  END SUBROUTINE ComputeTimeStep
END MODULE ModInputOutput

MODULE ModRungeKutta
  USE ModuleVariables
  USE ModDecomp
  USE ModInputOutput
  IMPLICIT NONE
CONTAINS
  ! RK4 (ModRungeKutta.f90:5-171): the four stages, the axis fix and the final update on the
  ! GPU; output every noutput iterations.
  SUBROUTINE RK4
    CALL check(ajet_rk4_step(ctx), 'RK4')
    IF (MOD(iteration, INT(noutput, INT64)) == 0) THEN
      CALL PrintSteps
      PRINT *, "Calling output"
      CALL WriteOutput
    END IF
  END SUBROUTINE RK4
END MODULE ModRungeKutta

PROGRAM ModSolver
  USE ModuleVariables
  USE ModDecomp
  USE ModPatch
  USE ModInputOutput
  USE ModRungeKutta
  IMPLICIT NONE
  CHARACTER(LEN=256) :: nml
  INTEGER :: u, ios
  TYPE(ajet_params_t) :: p
  TYPE(ajet_forcing_t) :: fc
  TYPE(ajet_mlp_t) :: mlp
  NAMELIST /ajet_input/ GRID_FILENAME, INITIAL_CONDITION, TARGET_FILE, bc_file, out_prefix, Re, Pr, CFL, &
                        ShockCapturing, C_mu, C_beta, C_kappa, do_viscous_terms, no_iterations, noutput, &
                        ControlForcing, forcing_file, LearnedLAD, mlp_file, sensor_threshold, ramp_min, &
                        mlp_refresh

  ! run-time inputs the reference compiles in (ModuleVariables.f90:6-9, 69, 73, 152, 164, 177;
  ! ModSolver.f90:116), read from a namelist file (default ajet.nml) when present
  nml = 'ajet.nml'
  IF (COMMAND_ARGUMENT_COUNT() >= 1) CALL GET_COMMAND_ARGUMENT(1, nml)
  OPEN (NEWUNIT=u, FILE=TRIM(nml), STATUS='old', IOSTAT=ios)
  IF (ios == 0) THEN
    READ (u, NML=ajet_input, IOSTAT=ios)
    IF (ios /= 0) THEN
      PRINT *, 'cannot read namelist ajet_input from ', TRIM(nml)
      STOP 1
    END IF
    CLOSE (u)
  END IF

  CALL GridRead
  CALL CreateAllPatches
  CALL ajet_params_default(p, INT(Nz, C_INT32_T), INT(Nr, C_INT32_T))
  p%Re = Re
  p%Pr = Pr
  p%CFL = CFL
  p%shock_capturing = ShockCapturing
  p%viscous = do_viscous_terms
  p%C_mu = C_mu
  p%C_beta = C_beta
  p%C_kappa = C_kappa
  IF (ControlForcing == 1) p%forcing = AJET_FORCING_FEEDBACK
  CALL check(ajet_create(p, zg, rg, iblank, bclines, INT(SIZE(bclines), C_INT32_T), ctx), 'ajet_create')
  IF (ControlForcing == 1) THEN
    CALL ForcingRead(fc)
    CALL check(ajet_set_forcing(ctx, fc), 'ajet_set_forcing')
  END IF
  IF (LearnedLAD == 1) THEN
    CALL MlpRead(mlp)
    CALL check(ajet_set_learned_lad(ctx, mlp, sensor_threshold, ramp_min, INT(mlp_refresh, C_INT32_T)), &
               'ajet_set_learned_lad')
  END IF
  CALL SolnRead
  CALL check(ajet_set_state(ctx, pv, pv_target), 'ajet_set_state')

  iteration = INT(tau(1), INT64)
  iteration_start = iteration
  time_start = tau(4)
  time = time_start
  CALL WriteOutput

  ! The reference also calls ComputeTimeStep once here (ModSolver.f90:112); the loop's first
  ! call recomputes the same value, so it is not repeated.
  DO iteration = iteration_start + 1, iteration_start + no_iterations
    CALL ComputeTimeStep
    CALL RK4
  END DO
  iteration = iteration_start + no_iterations
  CALL PrintSteps
  CALL StopTheCode
END PROGRAM ModSolver
