! ISO_C_BINDING mirror of include/ajet_capi.h: the derived types are bind(C) copies of the C
! structs and every entry point is an interface to the gfx950 library libajet.so.
MODULE ajet_capi
  USE, INTRINSIC :: ISO_C_BINDING
  IMPLICIT NONE

  INTEGER(C_INT), PARAMETER :: AJET_OK = 0, AJET_EBLOWUP = -5, AJET_ENONFINITE = -6
  INTEGER(C_INT), PARAMETER :: AJET_DTLOG_CAP = 1048576  ! entries the device time-step log holds

  ! ajet_params (solver constants; ModuleVariables.f90:12-13, 66, 71, 73, 152, 164, 177)
  TYPE, BIND(C) :: ajet_params_t
    INTEGER(C_INT32_T) :: nz, nr, viscous, shock_capturing, forcing, reserved
    REAL(C_DOUBLE) :: gam, Re, Pr, CFL, C_mu, C_beta, C_kappa
    REAL(C_DOUBLE) :: sponge_start_top, sponge_A_top, sponge_end_left, sponge_A_left
  END TYPE ajet_params_t

  ! one bc-file line: grid, bc, bc_dir, is, ie, js, je (ModPatch.f90:49)
  TYPE, BIND(C) :: ajet_bc_t
    INTEGER(C_INT64_T) :: grid, bc, bc_dir, i_start, i_end, j_start, j_end
  END TYPE ajet_bc_t

  INTEGER(C_INT32_T), PARAMETER :: AJET_FORCING_FEEDBACK = 1, AJET_FORCING_EXCITATION = 2

  ! ajet_forcing: F(k, e) is the C member F[e][k], the feedback weight of variable k in
  ! equation e (the transpose of the reference's forcing_mat(e, k), ModControlForcing.f90:20-27)
  TYPE, BIND(C) :: ajet_forcing_t
    REAL(C_DOUBLE) :: F(5, 5)
    REAL(C_DOUBLE) :: z0, r0, lz, lr
    REAL(C_DOUBLE) :: amp, strouhal, uj, t0
  END TYPE ajet_forcing_t

  ! ajet_mlp: the learned LAD network (include/ajet_capi.h).  The C arrays w1[128][104] etc. are
  ! row major, so their Fortran mirrors take the indices in the other order: w1(k, o) is the
  ! weight of input k in output o.
  TYPE, BIND(C) :: ajet_mlp_t
    REAL(C_FLOAT) :: w1(104, 128), b1(128)
    REAL(C_FLOAT) :: w2(128, 128), b2(128)
    REAL(C_FLOAT) :: w3(128, 128), b3(128)
    REAL(C_FLOAT) :: w4(128, 2), b4(2)
    REAL(C_FLOAT) :: shift(104), scale(104)
    REAL(C_FLOAT) :: out_scale(2)
  END TYPE ajet_mlp_t

  INTERFACE
    SUBROUTINE ajet_params_default(p, nz, nr) BIND(C, name='ajet_params_default')
      IMPORT :: ajet_params_t, C_INT32_T
      TYPE(ajet_params_t), INTENT(OUT) :: p
      INTEGER(C_INT32_T), VALUE :: nz, nr
    END SUBROUTINE ajet_params_default

    INTEGER(C_INT) FUNCTION ajet_create(p, z, r, iblank, lines, nlines, ctx) BIND(C, name='ajet_create')
      IMPORT :: ajet_params_t, ajet_bc_t, C_INT, C_DOUBLE, C_INT32_T, C_PTR
      TYPE(ajet_params_t), INTENT(IN) :: p
      REAL(C_DOUBLE), INTENT(IN) :: z(*), r(*)
      INTEGER(C_INT32_T), INTENT(IN) :: iblank(*)
      TYPE(ajet_bc_t), INTENT(IN) :: lines(*)
      INTEGER(C_INT32_T), VALUE :: nlines
      TYPE(C_PTR), INTENT(OUT) :: ctx
    END FUNCTION ajet_create

    INTEGER(C_INT) FUNCTION ajet_destroy(ctx) BIND(C, name='ajet_destroy')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
    END FUNCTION ajet_destroy

    TYPE(C_PTR) FUNCTION ajet_last_error_c(ctx) BIND(C, name='ajet_last_error')
      IMPORT :: C_PTR
      TYPE(C_PTR), VALUE :: ctx
    END FUNCTION ajet_last_error_c

    INTEGER(C_INT) FUNCTION ajet_set_state(ctx, pv, pv_target) BIND(C, name='ajet_set_state')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      REAL(C_DOUBLE), INTENT(IN) :: pv(*), pv_target(*)
    END FUNCTION ajet_set_state

    INTEGER(C_INT) FUNCTION ajet_get_state(ctx, pv) BIND(C, name='ajet_get_state')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      REAL(C_DOUBLE), INTENT(OUT) :: pv(*)
    END FUNCTION ajet_get_state

    ! dt: C_NULL_PTR keeps the time step on the device (no synchronisation)
    INTEGER(C_INT) FUNCTION ajet_compute_dt(ctx, dt) BIND(C, name='ajet_compute_dt')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx, dt
    END FUNCTION ajet_compute_dt

    INTEGER(C_INT) FUNCTION ajet_rk4_step(ctx) BIND(C, name='ajet_rk4_step')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
    END FUNCTION ajet_rk4_step

    INTEGER(C_INT) FUNCTION ajet_set_forcing(ctx, f) BIND(C, name='ajet_set_forcing')
      IMPORT :: C_INT, C_PTR, ajet_forcing_t
      TYPE(C_PTR), VALUE :: ctx
      TYPE(ajet_forcing_t), INTENT(IN) :: f
    END FUNCTION ajet_set_forcing

    INTEGER(C_INT) FUNCTION ajet_set_learned_lad(ctx, model, sensor_threshold, ramp_min, refresh) &
        BIND(C, name='ajet_set_learned_lad')
      IMPORT :: C_INT, C_PTR, C_DOUBLE, C_INT32_T, ajet_mlp_t
      TYPE(C_PTR), VALUE :: ctx
      TYPE(ajet_mlp_t), INTENT(IN) :: model
      REAL(C_DOUBLE), VALUE :: sensor_threshold, ramp_min
      INTEGER(C_INT32_T), VALUE :: refresh
    END FUNCTION ajet_set_learned_lad

    INTEGER(C_INT) FUNCTION ajet_get_dt_log(ctx, dts, n, time) BIND(C, name='ajet_get_dt_log')
      IMPORT :: C_INT, C_PTR, C_INT32_T
      TYPE(C_PTR), VALUE :: ctx, dts, time
      INTEGER(C_INT32_T), VALUE :: n
    END FUNCTION ajet_get_dt_log
  END INTERFACE

CONTAINS

  ! ajet_last_error as a Fortran string
  FUNCTION ajet_last_error(ctx) RESULT(msg)
    TYPE(C_PTR), INTENT(IN) :: ctx
    CHARACTER(LEN=256) :: msg
    CHARACTER(KIND=C_CHAR), POINTER :: s(:)
    INTEGER :: k
    TYPE(C_PTR) :: p
    msg = ' '
    p = ajet_last_error_c(ctx)
    IF (.NOT. C_ASSOCIATED(p)) RETURN
    CALL C_F_POINTER(p, s, [256])
    DO k = 1, 256
      IF (s(k) == C_NULL_CHAR) EXIT
      msg(k:k) = s(k)
    END DO
  END FUNCTION ajet_last_error

END MODULE ajet_capi
