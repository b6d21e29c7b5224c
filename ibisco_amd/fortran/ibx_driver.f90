!> A step driver in the shape of NEW_LOOP.F90:19-156 over libibx.so: reads the
!> input directory (the reference's control/interaction/coordinate files), runs
!> NSTEP steps in blocks of NTRJ on the GPU and prints, per block, the
!> observables OUTPUT.F90 reports (energies in kJ/mol, temperature in K).
!>
!>   ibx_driver <input dir> <nsteps> <block>
PROGRAM IBX_DRIVER
  USE, INTRINSIC :: ISO_C_BINDING
  USE IBX_MOD
  IMPLICIT NONE
  ! UNIT.f90:9-23
  REAL(8), PARAMETER :: KB = 1.380658D-23, NA = 6.0221367D23
  REAL(8), PARAMETER :: CONV = NA * KB * 300.0D0 / 1000.0D0, TEMPSCALE = 300.0D0
  TYPE(C_PTR) :: ctx
  CHARACTER(LEN=1024) :: dir, arg
  INTEGER :: nstep, ntrj, step, n, rc
  REAL(C_DOUBLE) :: obs(32)
  REAL(8) :: epot, ekin

  CALL GET_COMMAND_ARGUMENT(1, dir)
  CALL GET_COMMAND_ARGUMENT(2, arg)
  READ(arg, *) nstep
  CALL GET_COMMAND_ARGUMENT(3, arg)
  READ(arg, *) ntrj

  rc = ibx_load_dir(TRIM(dir) // C_NULL_CHAR, ctx)
  IF (rc /= IBX_OK) THEN
     WRITE(*, '(A,I0,2A)') 'ibx_load_dir failed (', rc, '): ', TRIM(ibx_error_text(C_NULL_PTR))
     STOP 1
  END IF
  WRITE(*, '(A)') '# step  Etot  Epot  Ekin (kJ/mol)  T (K)  V_NB  V_BOND  V_ANGLE  V_TORSION'
  step = 0
  DO WHILE (step < nstep)
     n = MIN(ntrj, nstep - step)
     rc = ibx_step_graph(ctx, n, step)
     IF (rc /= IBX_OK) THEN
        WRITE(*, '(A,I0,2A)') 'ibx_step_graph failed (', rc, '): ', TRIM(ibx_error_text(ctx))
        CALL ibx_destroy(ctx)
        STOP 1
     END IF
     step = step + n
     rc = ibx_get_observables(ctx, 0, obs)
     epot = SUM(obs(1:15))
     ekin = obs(22) + obs(23)
     WRITE(*, '(I8,8ES20.11)') step, (epot + ekin) * CONV, epot * CONV, ekin * CONV, obs(24) * TEMPSCALE, &
          SUM(obs(1:3)) * CONV, SUM(obs(4:6)) * CONV, SUM(obs(7:9)) * CONV, SUM(obs(10:12)) * CONV
  END DO
  CALL ibx_destroy(ctx)
END PROGRAM IBX_DRIVER
