! ORACLE -- test infrastructure only.  A driver of this project's own that links the
! reference solver's modules, compiled unmodified from /root/reference by oracle/build_ref.sh,
! and calls them in the order of PROGRAM ModSolver (ModSolver.f90:19-125).  It sets the
! run-time module variables the reference exposes (Re, Pr, CFL, dims, ShockCapturing, file
! names, noutput) instead of editing the reference, and dumps raw FP64 fixtures.
!
! Usage: ref_harness <namelist file>
!   &harness h_grid, h_ic, h_target, h_bc, h_re, h_pr, h_cfl, h_shock, h_mode, h_nsteps,
!            h_out, h_cmu, h_cbeta, h_ckappa /
!   mode = 'rhs'   : AssignValues + ComputeRHS of the initial condition -> <out>rhs.bin,
!                    plus tables.bin, metrics.bin and, with shock=1, the LAD stages
!   mode = 'steps' : nsteps iterations of {ComputeTimeStep; RK4} -> <out>state.bin, <out>dt.bin
!   mode = 'ladrhs': ComputeRHS with the artificial coefficients C*Re*filtered*ramp injected
!                    into mu, beta, kappa (C_mu, C_beta, C_kappa are PARAMETERs = 0 in the
!                    reference, so the scaling is restated here) -> <out>ladrhs.bin
!   mode = 'forcingrhs': ComputeRHS, then the reference's ControlForcingTerms (its switch
!                    ControlForcing is a PARAMETER = 0, so the call is made here, after the
!                    BCs instead of before them) with Vz_Control_Vr_Feedback.dat from the
!                    working directory -> <out>forcingrhs.bin
!   mode = 'write' : the reference's WriteOutput of the initial condition (the output
!                    ModSolver.f90:108 writes before the time loop) -> RocFlo-CM.########.q in
!                    the working directory
!   h_write = .true. with mode 'steps': WriteOutput after the last step as well (a restart
!                    file: a later run with h_ic pointing at it is the reference's restart)
PROGRAM ref_harness
  USE ModuleVariables
  USE ModInputOutput
  USE ModAllocate
  USE ModAssign
  USE ModDecomp
  USE ModCommunicate
  USE ModDeriv
  USE ModTGfilter
  USE ModOperator
  USE ModRungeKutta
  USE ModPatch
  USE ModPlot3D_IO
  USE ModShockCapturing
  USE ModControlForcing
  USE ModRHS
  IMPLICIT NONE

  CHARACTER(LEN=200) :: h_grid, h_ic, h_target, h_bc, h_out, nmlfile
  CHARACTER(LEN=16) :: h_mode
  DOUBLE PRECISION :: h_re, h_pr, h_cfl, h_cmu, h_cbeta, h_ckappa
  INTEGER :: h_shock, h_nsteps, u, s
  LOGICAL :: h_write
  DOUBLE PRECISION, ALLOCATABLE :: dts(:), scratch(:,:,:)
  NAMELIST /harness/ h_grid, h_ic, h_target, h_bc, h_re, h_pr, h_cfl, h_shock, h_mode, h_nsteps, &
                     h_out, h_cmu, h_cbeta, h_ckappa, h_write

  h_re = 1000.0D0; h_pr = 0.72D0; h_cfl = 0.5D0; h_shock = 0; h_nsteps = 1; h_mode = 'rhs'
  h_write = .FALSE.
  h_cmu = 0.0D0; h_cbeta = 0.0D0; h_ckappa = 0.0D0
  CALL get_command_argument(1, nmlfile)
  OPEN(NEWUNIT=u, FILE=TRIM(nmlfile), STATUS='old')
  READ(u, NML=harness)
  CLOSE(u)

  GRID_FILENAME = h_grid
  INITIAL_CONDITION = h_ic(1:80)
  TARGET_FILE = h_target
  bc_file = h_bc
  Re = h_re
  Pr = h_pr
  CFL = h_cfl
  ShockCapturing = h_shock
  dims(1) = 1
  dims(2) = 1
  noutput = HUGE(noutput)

  ! --- ModSolver.f90:19-106 ---
  CALL MPI_INIT(ierr)
  CALL MPI_COMM_SIZE(MPI_COMM_WORLD, numtasks, ierr)
  CALL MPI_COMM_RANK(MPI_COMM_WORLD, rank, ierr)
  CALL GridRead
  CALL CreateDecomposition
  CALL DistributeGridDims
  CALL DistributeGridPoints
  CALL GatherStartEndIndices
  CALL AllocateVariables
  CALL AssignGridCoordinates
  CALL CreateAllPatches
  CALL SBP_operator
  CALL ProcessorCommunication(zvec)
  CALL ComputeMetric(zvec,1,dzdxi)
  CALL ProcessorCommunication(rvec)
  CALL ComputeMetric(rvec,2,drdeta)
  DO i = 1, Nz
    DO j = 1, Nr
      dxidz(i,j,1) = 1.0D0/dzdxi(i,j,1)
      detadr(i,j,1) = 1.0D0/drdeta(i,j,1)
    ENDDO
  ENDDO
  CALL ProcessorCommunication(dxidz)
  CALL ComputeMetric(dxidz,1,ddxidxidz)
  CALL ProcessorCommunication(detadr)
  CALL ComputeMetric(detadr,2,ddetadetadr)
  CALL SolnRead
  CALL AssignValues
  iteration = tau(1)
  iteration_start = tau(1)
  time_start = tau(4)
  time = time_start

  IF (TRIM(h_mode) == 'rhs') THEN
    CALL dump3(TRIM(h_out)//'tables.bin')
    CALL dump_metrics(TRIM(h_out)//'metrics.bin')
    CALL ComputeRHS
    CALL dump_state(rhs, TRIM(h_out)//'rhs.bin')
    IF (ShockCapturing == 1) CALL dump_lad_stages(TRIM(h_out)//'lad.bin')
  ELSEIF (TRIM(h_mode) == 'ladrhs') THEN
    CALL inject_lad()
    CALL ComputeRHS
    CALL dump_state(rhs, TRIM(h_out)//'ladrhs.bin')
  ELSEIF (TRIM(h_mode) == 'forcingrhs') THEN
    CALL ComputeRHS
    iteration = iteration_start + 1
    CALL ControlForcingTerms
    CALL dump_state(rhs, TRIM(h_out)//'forcingrhs.bin')
  ELSEIF (TRIM(h_mode) == 'write') THEN
    CALL WriteOutput
  ELSEIF (TRIM(h_mode) == 'steps') THEN
    ALLOCATE(dts(h_nsteps))
    CALL ComputeTimeStep
    DO s = 1, h_nsteps
      iteration = iteration_start + s
      CALL ComputeTimeStep
      dts(s) = delt
      CALL RK4
    ENDDO
    CALL dump_state(pv, TRIM(h_out)//'state.bin')
    IF (h_write) CALL WriteOutput
    OPEN(NEWUNIT=u, FILE=TRIM(h_out)//'dt.bin', ACCESS='stream', FORM='unformatted', STATUS='replace')
    WRITE(u) dts
    WRITE(u) time
    CLOSE(u)
  ENDIF
  CALL MPI_FINALIZE(ierr)

CONTAINS

  SUBROUTINE dump_state(a, fname)
    DOUBLE PRECISION :: a(:,:,:,:)
    CHARACTER(LEN=*) :: fname
    INTEGER :: uu
    OPEN(NEWUNIT=uu, FILE=fname, ACCESS='stream', FORM='unformatted', STATUS='replace')
    WRITE(uu) a(1:Nz,1:Nr,1,1:5)
    CLOSE(uu)
  END SUBROUTINE dump_state

  SUBROUTINE dump3(fname)
    CHARACTER(LEN=*) :: fname
    INTEGER :: uu
    OPEN(NEWUNIT=uu, FILE=fname, ACCESS='stream', FORM='unformatted', STATUS='replace')
    WRITE(uu) SBP1, SBP2, TGfilt
    CLOSE(uu)
  END SUBROUTINE dump3

  SUBROUTINE dump_metrics(fname)
    CHARACTER(LEN=*) :: fname
    INTEGER :: uu
    OPEN(NEWUNIT=uu, FILE=fname, ACCESS='stream', FORM='unformatted', STATUS='replace')
    WRITE(uu) dxidz(1:Nz,1:Nr,1), detadr(1:Nz,1:Nr,1), ddxidxidz(1:Nz,1:Nr,1), ddetadetadr(1:Nz,1:Nr,1)
    CLOSE(uu)
  END SUBROUTINE dump_metrics

  ! After a ComputeRHS with ShockCapturing=1: mustarQuantity etc. hold the z-filtered
  ! quantities (filtered in place, ModShockCapturing.f90:63-75).  Filter them in r with the
  ! reference's own PerformTGfilter and dump both.
  SUBROUTINE dump_lad_stages(fname)
    CHARACTER(LEN=*) :: fname
    INTEGER :: uu
    ALLOCATE(scratch(-3:Nz+4,-3:Nr+4,1))
    OPEN(NEWUNIT=uu, FILE=fname, ACCESS='stream', FORM='unformatted', STATUS='replace')
    WRITE(uu) mustarQuantity(1:Nz,1:Nr,1), betastarQuantity(1:Nz,1:Nr,1), kappastarQuantity(1:Nz,1:Nr,1)
    scratch = 0.0D0
    CALL PerformTGfilter(mustarQuantity,2,scratch)
    WRITE(uu) scratch(1:Nz,1:Nr,1)
    scratch = 0.0D0
    CALL PerformTGfilter(betastarQuantity,2,scratch)
    WRITE(uu) scratch(1:Nz,1:Nr,1)
    scratch = 0.0D0
    CALL PerformTGfilter(kappastarQuantity,2,scratch)
    WRITE(uu) scratch(1:Nz,1:Nr,1)
    WRITE(uu) d4MgnStrRatedxi4(1:Nz,1:Nr,1), d4MgnStrRatedeta4(1:Nz,1:Nr,1), d4Tvecdxi4(1:Nz,1:Nr,1), &
              d4Tvecdeta4(1:Nz,1:Nr,1), d4dilvecdxi4(1:Nz,1:Nr,1), d4dilvecdeta4(1:Nz,1:Nr,1), &
              MgnStrRate(1:Nz,1:Nr,1), dilvec(1:Nz,1:Nr,1)
    CLOSE(uu)
    DEALLOCATE(scratch)
  END SUBROUTINE dump_lad_stages

  ! Run ComputeRHS once (with ShockCapturing=1) to get the filtered quantities, scale them
  ! as ModShockCapturing.f90:66,72,79,85-88 do with the given constants, and add them to
  ! the transport coefficients as UpdateTransportCoeffs (ModShockCapturing.f90:248-251) does.
  ! The following ComputeRHS then adds zero (the reference's C are 0) to these values.
  SUBROUTINE inject_lad()
    DOUBLE PRECISION :: ramp
    DOUBLE PRECISION, ALLOCATABLE :: ms(:,:), bs(:,:), ks(:,:)
    ShockCapturing = 1
    CALL ComputeRHS
    ALLOCATE(scratch(-3:Nz+4,-3:Nr+4,1), ms(Nz,Nr), bs(Nz,Nr), ks(Nz,Nr))
    scratch = 0.0D0
    CALL PerformTGfilter(mustarQuantity,2,scratch)
    ms = h_cmu*Re*scratch(1:Nz,1:Nr,1)
    scratch = 0.0D0
    CALL PerformTGfilter(betastarQuantity,2,scratch)
    bs = h_cbeta*Re*scratch(1:Nz,1:Nr,1)
    scratch = 0.0D0
    CALL PerformTGfilter(kappastarQuantity,2,scratch)
    ks = h_ckappa*Re*Pr*scratch(1:Nz,1:Nr,1)
    CALL AssignValues
    DO i = 1, Nz
      DO j = 1, Nr
        ramp = ( 1.0D0+tanh( 5.0D0*(zvec(i,j,1)-10.0D0) ) )/2.0D0
        muvec(i,j,1) = muvec(i,j,1) + ms(i,j)*ramp
        betavec(i,j,1) = betavec(i,j,1) + bs(i,j)*ramp
        lambdavec(i,j,1) = betavec(i,j,1) - (2.0D0/3.0D0)*muvec(i,j,1)
        kappavec(i,j,1) = kappavec(i,j,1) + ks(i,j)*ramp
      ENDDO
    ENDDO
    DEALLOCATE(scratch, ms, bs, ks)
  END SUBROUTINE inject_lad

END PROGRAM ref_harness
