! Driver for the reference's own subroutines, compiled from the reference
! sources where they lie (recipe: oracle/Makefile.ref, outputs in oracle/_ref/).
!
! TEST INFRASTRUCTURE ONLY.  It reproduces the set-up sequence of the
! reference's main program (main.F90:107-234) without its file output, then
!   force        one NEW_FORCE, dump forces, energies, virial, lists (ref_dump.bin)
!   run N        N steps in NEW_LOOP order (MTS_FORCE when mts>0, SCALEBP for NPT), dump per-step
!                observables (ref_obs.bin), the final state (ref_state.bin) and box (ref_box.bin)
!   time N       the same N steps timed with OMP_GET_WTIME (prints LOOP_SECONDS)
!   trz N        N steps, then the reference's WRITETRJ twice at that state (s-md.trj with two frames,
!                config.xyz, restart), and the values it wrote them from (ref_trz_vals.bin)
! For systems above 20000 items the exclusion lists are built by an O(N)
! routine of this file (same lists as BUILD_CONNECTIVITY.f90, whose NITEMS^2
! LOGICAL matrix would not fit in memory).  Without virtual sites the force
! phase runs NEW_FORCE.F90's sequence from this file (FORCE_NOVS) because the
! zero-trip OpenMP loop of DISTRIBUTE_VSFORCE.f90:21-38 raises SIGFPE in the
! amdflang OpenMP runtime; with virtual sites NEW_FORCE itself is called.
PROGRAM REF_HARNESS
  USE MODULEPARSING
  USE MTS
  USE VAR
  USE OMP_LIB
  IMPLICIT NONE
  INTEGER :: I, A, NRUN, K, NARG
  CHARACTER(len=32) :: MODE, ARG
  REAL*8 :: T0, T1

  NARG = COMMAND_ARGUMENT_COUNT()
  MODE = 'force'
  NRUN = 0
  IF (NARG >= 1) CALL GET_COMMAND_ARGUMENT(1, MODE)
  IF (NARG >= 2) THEN
     CALL GET_COMMAND_ARGUMENT(2, ARG)
     READ(ARG, *) NRUN
  END IF

  OPEN (1, FILE='ERROR')
  ISTOP = 0
  CALL UNIT()
  NVIRTA = 0
  CALL RDCONTROL()
  IF (ISTOP == 1) STOP 'Failed in RDCONTROL'
  CALL ALLOCATEVAR()
  CALL RDINTERACT()
  IF (ISTOP == 1) STOP 'Failed in RDINTERACT'
  CALL RDCOOR()
  IF (ISTOP == 1) STOP 'Failed in RDCOOR'
  CALL ALLOCATEVAR2()
  NCOARSE = 0
  IF (IBRDESCR .EQ. 0) THEN
     CALL RDVIRTUAL()
     IF (ISTOP .eq. 1) STOP 'Failed at RDVIRTUAL'
     CALL MAKE_LISTS()
     IF (ISTOP .eq. 1) STOP 'Failed at MAKE_LISTS'
     CALL VIRTUAL_DEF()
  ELSE
     NUMATOMS = NATOMS
     NUMBEADS = 0
     ALLOCATE(ATOM(NUMATOMS))
     DO I = 1, NATOMS
        ATOM(I) = I
     END DO
  END IF
  CALL SHIFT()
  MAXNAB_ATOM = 1000
  ALLOCATE(LIST_ATOM(MAXNAB_ATOM, NUMATOMS), CELL_ATOM(NUMATOMS), LCLIST_ATOM(NUMATOMS))
  IF (IBRDESCR .eq. 0) THEN
     MAXNAB_BEAD = 1000
     ALLOCATE(LIST_BEAD(MAXNAB_BEAD, NCOARSE), CELL_BEAD(NCOARSE), LCLIST_BEAD(NCOARSE))
  END IF
  ALLOCATE(NNEBS(NITEMS))
  CALL FTABLE()

  EK = 0.0
  MKTEMP_ATOM = 2.0 / REAL(3.0 * (NUMATOMS - 1))
  MKTEMP_BEAD = 2.0 / REAL(3.0 * (NUMBEADS - 1))
  TOTMASS = 0.0
  DO A = 1, NUMATOMS
     I = ATOM(A)
     EK(1) = EK(1) + MASS(ITYPE(I)) * (VXYZ(1,I)**2.0 + VXYZ(2,I)**2.0 + VXYZ(3,I)**2.0)
     TOTMASS = TOTMASS + MASS(ITYPE(I))
  END DO
  DO A = 1, NUMBEADS
     I = BEAD(A)
     EK(2) = EK(2) + MASS(ITYPE(I)) * (VXYZ(1,I)**2.0 + VXYZ(2,I)**2.0 + VXYZ(3,I)**2.0)
     TOTMASS = TOTMASS + MASS(ITYPE(I))
  END DO
  EK = 0.5 * EK
  TEMP = SUM(EK) * MKTEMP
  TEMP_ATOM = EK(1) * MKTEMP_ATOM
  TEMP_BEAD = EK(2) * MKTEMP_BEAD

  ! SETLIS.f90:17-24 zeroes the per-atom term counts only when bond types exist;
  ! a system without bonds (twoatombox) would otherwise read them uninitialised
  NIJK = 0; NOANGLEIJK = 0; NIJKL = 0; FNIJKL = 0; NOOPIJKL = 0
  CALL SETLIS()
  IF (ISTOP == 1) STOP 'Failed in SETLIS'
  IF (IBRDESCR .eq. 0) THEN
     CALL RDVIRTANGLES()
     IF (ISTOP .eq. 1) STOP 'Failed in RDVIRTANGLES'
  END IF
  CONNECTIONS = 0
  CONNECTED_TO = 0
  IF (NITEMS <= 20000) THEN
     CALL BUILD_CONNECTIVITY(NUMATOMS, ATOM, NONBEXC_ATOM)
     IF (IBRDESCR .eq. 0) CALL BUILD_CONNECTIVITY(NCOARSE, BEAD, NONBEXC_BEAD)
  ELSE
     CALL HARNESS_CONNECTIVITY(NUMATOMS, ATOM, NONBEXC_ATOM)
     IF (IBRDESCR .eq. 0) CALL HARNESS_CONNECTIVITY(NCOARSE, BEAD, NONBEXC_BEAD)
  END IF
  IF (IBRDESCR .eq. 0) THEN
     CALL RDVIRTBONDS()
     IF (ISTOP .EQ. 1) STOP 'Failed in RDVIRTBONDS'
  END IF
  CALL MAPS(MAP_ATOM, MAPSIZE_ATOM, NCELLX_ATOM, NCELLY_ATOM, NCELLZ_ATOM)
  CALL LINKS(HEAD_ATOM, MAXNUMCELL_ATOM, ATOM, NUMATOMS, CELL_ATOM, &
       NCELLX_ATOM, NCELLY_ATOM, NCELLZ_ATOM, LCLIST_ATOM)
  IF (IBRDESCR .eq. 0) THEN
     CALL MAPS(MAP_BEAD, MAPSIZE_BEAD, NCELLX_BEAD, NCELLY_BEAD, NCELLZ_BEAD)
     CALL LINKS(HEAD_BEAD, MAXNUMCELL_BEAD, BEAD, NCOARSE, CELL_BEAD, &
          NCELLX_BEAD, NCELLY_BEAD, NCELLZ_BEAD, LCLIST_BEAD)
  END IF
  CALL UPDATE_NEIGHBOURLIST()

  IF (NMTS > 0) THEN
     ALLOCATE(MTS_FXYZ(3, 3, NCOARSE))
     ALLOCATE(MTS_PT11(3), MTS_PT22(3), MTS_PT33(3), MTS_PT12(3), MTS_PT13(3), MTS_PT23(3))
     ALLOCATE(MTS_V_NB(3, 3))
     MTS_FXYZ = 0.0
     MTS_PT11 = 0.0; MTS_PT22 = 0.0; MTS_PT33 = 0.0
     MTS_PT12 = 0.0; MTS_PT13 = 0.0; MTS_PT23 = 0.0
     MTS_V_NB = 0.0
  END IF

  IF (MODE == 'force') THEN
     CALL FORCE()
     CALL DUMP_FORCE()
  ELSE IF (MODE == 'run') THEN
     OPEN(21, FILE='ref_obs.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
     DO K = 1, NRUN
        STEP = K
        CALL ONE_STEP(STEP)
        WRITE(21) V_NB, V_BOND, V_ANGLE, V_TORSION, V_OOP, &
             DBLE(PT11), DBLE(PT22), DBLE(PT33), DBLE(PT12), DBLE(PT13), DBLE(PT23), EK, DBLE(TEMP)
     END DO
     CLOSE(21)
     OPEN(22, FILE='ref_state.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
     WRITE(22) NITEMS, NATOMS
     WRITE(22) SXYZ, VXYZ, FXYZ
     CLOSE(22)
     OPEN(24, FILE='ref_box.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
     WRITE(24) DBLE(BOX(1)), DBLE(BOX(2)), DBLE(BOX(3))
     CLOSE(24)
  ELSE IF (MODE == 'trz') THEN
     DO K = 1, NRUN
        STEP = K
        CALL ONE_STEP(STEP)
     END DO
     OPEN(113, FILE='s-md.trj', FORM='UNFORMATTED', ACCESS='SEQUENTIAL', STATUS='REPLACE')
     NFRAME = 0
     CALL WRITETRJ(NRUN)
     CALL WRITETRJ(NRUN)
     CLOSE(113)
     OPEN(23, FILE='ref_trz_vals.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
     WRITE(23) NATOMS, NTRJ
     WRITE(23) DBLE(PT11), DBLE(PT22), DBLE(PT33), DBLE(PT12), DBLE(PT13), DBLE(PT23)
     WRITE(23) SUM(V_BOND) + SUM(V_ANGLE) + SUM(V_TORSION) + SUM(V_OOP) + SUM(V_NB), SUM(EK), DBLE(MKTEMP)
     WRITE(23) DBLE(BOX(1)), DBLE(BOX(2)), DBLE(BOX(3)), DBLE(DT), DBLE(INITIME)
     WRITE(23) SXYZ(:, 1:NATOMS), VXYZ(:, 1:NATOMS)
     CLOSE(23)
  ELSE IF (MODE == 'time') THEN
     T0 = OMP_GET_WTIME()
     DO K = 1, NRUN
        STEP = K
        CALL ONE_STEP(STEP)
     END DO
     T1 = OMP_GET_WTIME()
     WRITE(*, '(A,ES16.8,A,I0,A,I0,A,I0)') 'LOOP_SECONDS ', T1 - T0, ' STEPS ', NRUN, &
          ' ITEMS ', NATOMS, ' THREADS ', OMP_GET_MAX_THREADS()
  END IF

CONTAINS

  ! Step body in the order of NEW_LOOP.F90:66-122 / MTS_LOOP.f90:26-44 (no output).
  SUBROUTINE ONE_STEP(ST)
    INTEGER, INTENT(IN) :: ST
    CALL SHIFT()
    IF (IBRDESCR .eq. 0) CALL VIRTUAL_DEF()
    IF (MOD(ST, NUPDATE) .eq. 0) CALL UPDATE_NEIGHBOURLIST()
    IF (NMTS > 0) THEN
       CALL MTS_FORCE(ST)
    ELSE
       CALL FORCE()
    END IF
    CALL MOVE()
    IF (MOD(ST, HALT_DRIFT) .eq. 0) CALL MOMENTUM()
    IF (ENSEMBLE .eq. 2) CALL SCALEBP(ST)
  END SUBROUTINE ONE_STEP

  SUBROUTINE FORCE()
    IF (NVIRTA > 0) THEN
       CALL NEW_FORCE()
    ELSE
       CALL FORCE_NOVS()
    END IF
  END SUBROUTINE FORCE

  ! NEW_FORCE.F90:17-80 for a system without virtual sites.
  SUBROUTINE FORCE_NOVS()
    INTEGER :: II
    FXYZ = 0.0
    FXYZNB = 0.0
    PT11 = 0.0; PT22 = 0.0; PT33 = 0.0; PT12 = 0.0; PT13 = 0.0; PT23 = 0.0
    V_NB = 0.0; V_BOND = 0.0; V_ANGLE = 0.0; V_TORSION = 0.0; V_OOP = 0.0
    CALL NONBONDED_FORCE(NUMATOMS, ATOM, MAXNAB_ATOM, LIST_ATOM, RCUT_ATOM, RCUTSQ_ATOM)
    DO II = 1, NATOMS
       FXYZNB(:,II) = FXYZ(:,II)
    END DO
    CALL BONDED_FORCE()
    DO II = 1, NATOMS
       PT11 = PT11 + (FXYZ(1,II) - FXYZNB(1,II)) * SXYZ(1,II)
       PT22 = PT22 + (FXYZ(2,II) - FXYZNB(2,II)) * SXYZ(2,II)
       PT33 = PT33 + (FXYZ(3,II) - FXYZNB(3,II)) * SXYZ(3,II)
       PT12 = PT12 + (FXYZ(2,II) - FXYZNB(2,II)) * SXYZ(1,II)
       PT13 = PT13 + (FXYZ(3,II) - FXYZNB(3,II)) * SXYZ(1,II)
       PT23 = PT23 + (FXYZ(3,II) - FXYZNB(3,II)) * SXYZ(2,II)
    END DO
  END SUBROUTINE FORCE_NOVS

  SUBROUTINE DUMP_FORCE()
    INTEGER :: AA, KK, NP
    OPEN(20, FILE='ref_dump.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
    WRITE(20) NITEMS, NATOMS
    WRITE(20) FXYZ
    WRITE(20) V_NB, V_BOND, V_ANGLE, V_TORSION, V_OOP
    WRITE(20) PT11, PT22, PT33, PT12, PT13, PT23
    WRITE(20) RXYZ
    ! atom group list: count, then (I, J) pairs
    NP = 0
    DO AA = 1, NUMATOMS
       NP = NP + NNEBS(ATOM(AA))
    END DO
    WRITE(20) NP
    DO AA = 1, NUMATOMS
       DO KK = 1, NNEBS(ATOM(AA))
          WRITE(20) ATOM(AA), LIST_ATOM(KK, AA)
       END DO
    END DO
    NP = 0
    IF (IBRDESCR .eq. 0) THEN
       DO AA = 1, NCOARSE
          NP = NP + NNEBS(BEAD(AA))
       END DO
    END IF
    WRITE(20) NP
    IF (IBRDESCR .eq. 0) THEN
       DO AA = 1, NCOARSE
          DO KK = 1, NNEBS(BEAD(AA))
             WRITE(20) BEAD(AA), LIST_BEAD(KK, AA)
          END DO
       END DO
    END IF
    ! exclusion lists
    WRITE(20) CONNECTIONS
    DO AA = 1, NITEMS
       DO KK = 1, CONNECTIONS(AA)
          WRITE(20) CONNECTED_TO(AA, KK)
       END DO
    END DO
    ! enumerated angles and torsions per owner
    NP = 0
    DO AA = 1, NITEMS
       NP = NP + NIJK(AA)
    END DO
    WRITE(20) NP
    DO AA = 1, NITEMS
       DO KK = 1, NIJK(AA)
          WRITE(20) AA, JANGLEIJK(AA, KK), KANGLEIJK(AA, KK)
       END DO
    END DO
    NP = 0
    IF (NTTYPE > 0) THEN
       DO AA = 1, NATOMS
          NP = NP + NIJKL(AA)
       END DO
    END IF
    WRITE(20) NP
    IF (NTTYPE > 0) THEN
       DO AA = 1, NATOMS
          DO KK = 1, NIJKL(AA)
             WRITE(20) AA, JTORIJKL(AA, KK), KTORIJKL(AA, KK), LTORIJKL(AA, KK)
          END DO
       END DO
    END IF
    CLOSE(20)
  END SUBROUTINE DUMP_FORCE

END PROGRAM REF_HARNESS

! O(N) construction of CONNECTIONS / CONNECTED_TO with the result of
! BUILD_CONNECTIVITY.f90:14-116 (ascending partner lists over atoms), written
! for this harness.  Terms owned by any atom contribute both directions.
SUBROUTINE HARNESS_CONNECTIVITY(N, INDEX_LIST, EXCLUSION)
  USE VAR
  IMPLICIT NONE
  INTEGER, INTENT(IN) :: N, INDEX_LIST(N), EXCLUSION
  INTEGER, PARAMETER :: CAP = 64
  INTEGER, ALLOCATABLE :: TMP(:,:), CNT(:)
  INTEGER :: I, B, A, P, Q, X, M
  ALLOCATE(TMP(CAP, NATOMS), CNT(NATOMS))
  CNT = 0
  DO I = 1, NATOMS
     DO B = 1, NBONDS(I)
        CALL ADDP(I, JBOND(I,B))
     END DO
     DO B = 1, NIJK(I)
        CALL ADDP(I, JANGLEIJK(I,B)); CALL ADDP(I, KANGLEIJK(I,B))
     END DO
     DO B = 1, NOANGLEIJK(I)
        CALL ADDP(I, NOJANGLEIJK(I,B)); CALL ADDP(I, NOKANGLEIJK(I,B))
     END DO
     IF (EXCLUSION .eq. 5) THEN
        IF (NTTYPE > 0) THEN
           DO B = 1, NIJKL(I)
              CALL ADDP(I, JTORIJKL(I,B)); CALL ADDP(I, KTORIJKL(I,B)); CALL ADDP(I, LTORIJKL(I,B))
           END DO
           DO B = 1, FNIJKL(I)
              CALL ADDP(I, FJTORIJKL(I,B)); CALL ADDP(I, FKTORIJKL(I,B)); CALL ADDP(I, FLTORIJKL(I,B))
           END DO
        END IF
        IF (NOTYPE > 0) THEN
           DO B = 1, NOOPIJKL(I)
              CALL ADDP(I, JOOPIJKL(I,B)); CALL ADDP(I, KOOPIJKL(I,B)); CALL ADDP(I, LOOPIJKL(I,B))
           END DO
        END IF
     END IF
  END DO
  DO A = 1, N
     I = INDEX_LIST(A)
     IF (I .gt. NATOMS) EXIT
     ! sort ascending and drop duplicates
     DO P = 2, CNT(I)
        X = TMP(P, I)
        Q = P - 1
        DO WHILE (Q >= 1)
           IF (TMP(Q, I) <= X) EXIT
           TMP(Q + 1, I) = TMP(Q, I)
           Q = Q - 1
        END DO
        TMP(Q + 1, I) = X
     END DO
     M = 0
     DO P = 1, CNT(I)
        IF (P > 1) THEN
           IF (TMP(P, I) == TMP(P - 1, I)) CYCLE
        END IF
        M = M + 1
        CONNECTED_TO(I, M) = TMP(P, I)
     END DO
     CONNECTIONS(I) = M
  END DO
  DEALLOCATE(TMP, CNT)
CONTAINS
  SUBROUTINE ADDP(U, W)
    INTEGER, INTENT(IN) :: U, W
    IF (U <= NATOMS .and. W <= NATOMS) THEN
       CNT(U) = CNT(U) + 1; TMP(CNT(U), U) = W
       CNT(W) = CNT(W) + 1; TMP(CNT(W), W) = U
    END IF
  END SUBROUTINE ADDP
END SUBROUTINE HARNESS_CONNECTIVITY
