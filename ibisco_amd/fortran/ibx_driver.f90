!> A step driver in the shape of NEW_LOOP.F90:19-156 over libibx.so: reads the
!> input directory (the reference's control/interaction/coordinate files), runs
!> NSTEP steps in blocks of NTRJ on the GPU and prints, per block, the
!> observables OUTPUT.F90 reports (energies in kJ/mol, temperature in K); with
!> a fourth argument, also appends a TRZ frame per block to that file, as
!> WRITETRJ.F90 does every NTRJ steps (title: the coordinate file's first word).
!>
!>   ibx_driver <input dir> <nsteps> <block> [trajectory file]
PROGRAM IBX_DRIVER
  USE, INTRINSIC :: ISO_C_BINDING
  USE IBX_MOD
  IMPLICIT NONE
  ! UNIT.f90:9-23
  REAL(8), PARAMETER :: KB = 1.380658D-23, NA = 6.0221367D23
  REAL(8), PARAMETER :: CONV = NA * KB * 300.0D0 / 1000.0D0, TEMPSCALE = 300.0D0
  TYPE(C_PTR) :: ctx
  CHARACTER(LEN=1024) :: dir, arg, trz
  CHARACTER(LEN=80) :: title
  INTEGER :: nstep, ntrj, step, n, rc, frame
  REAL(C_DOUBLE) :: obs(32)
  REAL(8) :: epot, ekin

  CALL GET_COMMAND_ARGUMENT(1, dir)
  CALL GET_COMMAND_ARGUMENT(2, arg)
  READ(arg, *) nstep
  CALL GET_COMMAND_ARGUMENT(3, arg)
  READ(arg, *) ntrj
  trz = ''
  IF (COMMAND_ARGUMENT_COUNT() >= 4) CALL GET_COMMAND_ARGUMENT(4, trz)
  title = 'ibx'
  IF (LEN_TRIM(trz) > 0) THEN   ! RDCOOR.f90:55-58: the title is the first word of the coordinate file
     OPEN(31, FILE=TRIM(dir) // '/coordinate', STATUS='OLD', ACTION='READ')
     READ(31, *) title
     CLOSE(31)
  END IF
  frame = 0

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
     IF (LEN_TRIM(trz) > 0) THEN
        frame = frame + 1
        rc = ibx_write_trz(ctx, 0, TRIM(trz) // C_NULL_CHAR, TRIM(title) // C_NULL_CHAR, frame, ntrj, step, &
             0.0_C_DOUBLE)
        IF (rc /= IBX_OK) THEN
           WRITE(*, '(A,I0)') 'ibx_write_trz failed: ', rc
           CALL ibx_destroy(ctx)
           STOP 1
        END IF
     END IF
     rc = ibx_get_observables(ctx, 0, obs)
     epot = SUM(obs(1:15))
     ekin = obs(22) + obs(23)
     WRITE(*, '(I8,8ES20.11)') step, (epot + ekin) * CONV, epot * CONV, ekin * CONV, obs(24) * TEMPSCALE, &
          SUM(obs(1:3)) * CONV, SUM(obs(4:6)) * CONV, SUM(obs(7:9)) * CONV, SUM(obs(10:12)) * CONV
  END DO
  CALL ibx_destroy(ctx)
END PROGRAM IBX_DRIVER
