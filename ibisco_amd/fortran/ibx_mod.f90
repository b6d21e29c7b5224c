!> iso_c_binding interface to libibx.so (include/ibx.h): the per-step force and
!> integration path of IBIsCO on the GPU, bound from Fortran.  Arrays cross the
!> boundary as plain pointers; indices on the C side are 0-based.
MODULE IBX_MOD
  USE, INTRINSIC :: ISO_C_BINDING
  IMPLICIT NONE

  INTEGER(C_INT), PARAMETER :: IBX_OK = 0, IBX_E_INVALID = -1, IBX_E_LIST_OVERFLOW = -2, &
       IBX_E_TABLE = -3, IBX_E_DEVICE = -4, IBX_E_NOMEM = -5, IBX_E_IO = -6

  INTERFACE
     ! set-up from an input directory in the reference's formats (control, interaction, coordinate, ...)
     INTEGER(C_INT) FUNCTION ibx_load_dir(dir, ctx) BIND(C, NAME='ibx_load_dir')
       IMPORT :: C_INT, C_CHAR, C_PTR
       CHARACTER(KIND=C_CHAR), DIMENSION(*), INTENT(IN) :: dir
       TYPE(C_PTR), INTENT(OUT) :: ctx
     END FUNCTION ibx_load_dir

     SUBROUTINE ibx_destroy(ctx) BIND(C, NAME='ibx_destroy')
       IMPORT :: C_PTR
       TYPE(C_PTR), VALUE :: ctx
     END SUBROUTINE ibx_destroy

     TYPE(C_PTR) FUNCTION ibx_last_error(ctx) BIND(C, NAME='ibx_last_error')
       IMPORT :: C_PTR
       TYPE(C_PTR), VALUE :: ctx
     END FUNCTION ibx_last_error

     ! nsteps of NEW_LOOP, numbered step0+1 .. step0+nsteps
     INTEGER(C_INT) FUNCTION ibx_step(ctx, nsteps, step0) BIND(C, NAME='ibx_step')
       IMPORT :: C_INT, C_PTR
       TYPE(C_PTR), VALUE :: ctx
       INTEGER(C_INT), VALUE :: nsteps, step0
     END FUNCTION ibx_step

     INTEGER(C_INT) FUNCTION ibx_step_graph(ctx, nsteps, step0) BIND(C, NAME='ibx_step_graph')
       IMPORT :: C_INT, C_PTR
       TYPE(C_PTR), VALUE :: ctx
       INTEGER(C_INT), VALUE :: nsteps, step0
     END FUNCTION ibx_step_graph

     INTEGER(C_INT) FUNCTION ibx_compute_forces(ctx, rebuild) BIND(C, NAME='ibx_compute_forces')
       IMPORT :: C_INT, C_PTR
       TYPE(C_PTR), VALUE :: ctx
       INTEGER(C_INT), VALUE :: rebuild
     END FUNCTION ibx_compute_forces

     ! obs(1:3) V_NB, (4:6) V_BOND, (7:9) V_ANGLE, (10:12) V_TORSION, (13:15) V_OOP,
     ! (16:21) PT11 PT22 PT33 PT12 PT13 PT23, (22:23) EK, (24) TEMP, (25) TEMP_ATOM, (26) TEMP_BEAD
     INTEGER(C_INT) FUNCTION ibx_get_observables(ctx, sys, obs) BIND(C, NAME='ibx_get_observables')
       IMPORT :: C_INT, C_PTR, C_DOUBLE
       TYPE(C_PTR), VALUE :: ctx
       INTEGER(C_INT), VALUE :: sys
       REAL(C_DOUBLE), INTENT(OUT) :: obs(32)
     END FUNCTION ibx_get_observables

     ! window sums of obs(1:26) since the last reset (AVERAGE.f90), obs(29) = steps summed
     INTEGER(C_INT) FUNCTION ibx_get_averages(ctx, sys, sums, reset) BIND(C, NAME='ibx_get_averages')
       IMPORT :: C_INT, C_PTR, C_DOUBLE
       TYPE(C_PTR), VALUE :: ctx
       INTEGER(C_INT), VALUE :: sys, reset
       REAL(C_DOUBLE), INTENT(OUT) :: sums(32)
     END FUNCTION ibx_get_averages

     ! one TRZ frame of system sys (WRITETRJ.F90 format), frame 1 creating the file; time = step*DT + initime
     INTEGER(C_INT) FUNCTION ibx_write_trz(ctx, sys, path, title, frame, ntrj, step, initime) &
          BIND(C, NAME='ibx_write_trz')
       IMPORT :: C_INT, C_PTR, C_CHAR, C_DOUBLE
       TYPE(C_PTR), VALUE :: ctx
       INTEGER(C_INT), VALUE :: sys
       CHARACTER(KIND=C_CHAR), DIMENSION(*), INTENT(IN) :: path, title
       INTEGER(C_INT), VALUE :: frame, ntrj, step
       REAL(C_DOUBLE), VALUE :: initime
     END FUNCTION ibx_write_trz

     ! s(3, NITEMS) unwrapped positions (SXYZ), v(3, NATOMS) reduced velocities (VXYZ)
     INTEGER(C_INT) FUNCTION ibx_get_state(ctx, sys, s, v) BIND(C, NAME='ibx_get_state')
       IMPORT :: C_INT, C_PTR, C_DOUBLE
       TYPE(C_PTR), VALUE :: ctx
       INTEGER(C_INT), VALUE :: sys
       REAL(C_DOUBLE), INTENT(OUT) :: s(*), v(*)
     END FUNCTION ibx_get_state

     INTEGER(C_INT) FUNCTION ibx_set_state(ctx, sys, s, v) BIND(C, NAME='ibx_set_state')
       IMPORT :: C_INT, C_PTR, C_DOUBLE
       TYPE(C_PTR), VALUE :: ctx
       INTEGER(C_INT), VALUE :: sys
       REAL(C_DOUBLE), INTENT(IN) :: s(*), v(*)
     END FUNCTION ibx_set_state

     ! sizes: natoms, nvs, list capacity, padded slots, device bytes
     INTEGER(C_INT) FUNCTION ibx_get_info(ctx, sys, info) BIND(C, NAME='ibx_get_info')
       IMPORT :: C_INT, C_PTR, C_INT64_T
       TYPE(C_PTR), VALUE :: ctx
       INTEGER(C_INT), VALUE :: sys
       INTEGER(C_INT64_T), INTENT(OUT) :: info(8)
     END FUNCTION ibx_get_info

     ! multiple time step: nonbonded forces every k steps with weight k, bonded every step
     INTEGER(C_INT) FUNCTION ibx_set_respa(ctx, k) BIND(C, NAME='ibx_set_respa')
       IMPORT :: C_INT, C_PTR
       TYPE(C_PTR), VALUE :: ctx
       INTEGER(C_INT), VALUE :: k
     END FUNCTION ibx_set_respa

     ! list-skin autotuner on (1) or off (0)
     INTEGER(C_INT) FUNCTION ibx_set_skin_tuning(ctx, on) BIND(C, NAME='ibx_set_skin_tuning')
       IMPORT :: C_INT, C_PTR
       TYPE(C_PTR), VALUE :: ctx
       INTEGER(C_INT), VALUE :: on
     END FUNCTION ibx_set_skin_tuning
  END INTERFACE

CONTAINS

  !> The text of ibx_last_error as a Fortran string.
  FUNCTION ibx_error_text(ctx) RESULT(msg)
    TYPE(C_PTR), INTENT(IN) :: ctx
    CHARACTER(LEN=256) :: msg
    TYPE(C_PTR) :: p
    CHARACTER(KIND=C_CHAR), POINTER :: c(:)
    INTEGER :: k
    msg = ' '
    p = ibx_last_error(ctx)
    IF (.NOT. C_ASSOCIATED(p)) RETURN
    CALL C_F_POINTER(p, c, [256])
    DO k = 1, 256
       IF (c(k) == C_NULL_CHAR) EXIT
       msg(k:k) = c(k)
    END DO
  END FUNCTION ibx_error_text

END MODULE IBX_MOD
