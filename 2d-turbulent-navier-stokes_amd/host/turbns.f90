!===============================================================================
! turbns.f90 - driver of the Fortran host (the reference's turbnstwod.f90 /
! turbnsaxis.f90 flow: flow.in, geometry, cold start or restart, RK4 loop,
! monitor, periodic and final restart files, tec_plot.dat).
!
! The RK4 iterations run on the GPU through libtdt.  By default whole
! iterations go through the fused path (tdt_iterate) in chunks that end at
! every monitor (1000 iterations) and restart (20000) boundary; the last
! iteration of the run uses the routine-level entry points in the reference's
! CALL order, so the arrays POSTPROCESS reads are those of the reference's
! last stage.  TDT_HOST_MODE=routines runs every iteration that way.
! TDT_PRECISION=fp32-mixed selects the mixed-precision path.
! Optional bc7.in: PINF AMINF DELTABL, and bc6.in: TFLAME ANCOMB ACOMB RHOPROP,
! the BC type-7 and type-6 inputs the reference never assigns (they stay 0
! without them, as in the reference).
!===============================================================================
PROGRAM TURBNS
    USE, INTRINSIC :: ISO_C_BINDING
    USE TDT_C
    USE CONSTANTS
    USE SOLVER_STATE
    USE GEOMETRY
    USE VISCOUS
    USE SOURCE_TERMS
    USE BOUNDARY
    USE RESIDUAL
    IMPLICIT NONE
    INTEGER :: I, ITER, CHUNK, IOS
    INTEGER(C_INT32_T) :: BAD
    LOGICAL :: ROUTINES, HAVE_BL, HAVE_B6
    CHARACTER(LEN=32) :: ENVV
    DOUBLE PRECISION, ALLOCATABLE :: L2H(:,:), LINFH(:,:)
    DOUBLE PRECISION :: L2(5), LINF(5)

    OPEN(25, FILE='flow.in')
    READ(25, *) FNAMEGRID, FNAMEBC, GAMA, AMOL, LOCAL, NITER, ISTART, CFL, &
                PRANDTL, PRANDTLT, P0, T0, PAMB, AMACHINIT, RTRED, NPSEUDO, IAXIS
    CLOSE(25)
    PINF = 0.0D0; AMINF = 0.0D0; DELTABL = 0.0D0
    INQUIRE(FILE='bc7.in', EXIST=HAVE_BL)
    IF (HAVE_BL) THEN
        OPEN(27, FILE='bc7.in')
        READ(27, *, IOSTAT=IOS) PINF, AMINF, DELTABL
        CLOSE(27)
        IF (IOS .NE. 0) STOP 'bc7.in: expected PINF AMINF DELTABL'
    END IF
    TFLAME = 0.0D0; ANCOMB = 0.0D0; ACOMB = 0.0D0; RHOPROP = 0.0D0
    INQUIRE(FILE='bc6.in', EXIST=HAVE_B6)
    IF (HAVE_B6) THEN
        OPEN(27, FILE='bc6.in')
        READ(27, *, IOSTAT=IOS) TFLAME, ANCOMB, ACOMB, RHOPROP
        CLOSE(27)
        IF (IOS .NE. 0) STOP 'bc6.in: expected TFLAME ANCOMB ACOMB RHOPROP'
    END IF
    CALL GET_ENVIRONMENT_VARIABLE('TDT_HOST_MODE', ENVV)
    ROUTINES = TRIM(ENVV) .EQ. 'routines'
    CALL GET_ENVIRONMENT_VARIABLE('TDT_PRECISION', ENVV)
    IF (TRIM(ENVV) .EQ. 'fp32-mixed') PRECISION_MODE = TDT_FP32_MIXED

    CALL COMPUTE_GEOMETRY
    RR = RUNIV_VAL / AMOL
    NITEROLD = 0
    ALLOCATE(C(NELES, 5))
    IF (ISTART .EQ. 1) THEN
        OPEN(55, FILE='restart.in')
        DO I = 1, NELES
            READ(55, *) C(I,1), C(I,2), C(I,3), C(I,4), C(I,5)
        END DO
        READ(55, *) RTRED
        READ(55, *) NPSEUDO
        READ(55, *) RINIT
        READ(55, *) NITEROLD
        CLOSE(55)
    END IF
    CALL DEVICE_SETUP
    IF (ISTART .EQ. 0) THEN
        CALL TDT_CHECK(TDT_COLD_START(CTX, RINIT), CTX, 'cold start')
    ELSE
        CALL TDT_CHECK(TDT_SET_STATE(CTX, C, RINIT), CTX, 'restart state')
    END IF

    ALLOCATE(L2H(5, 1000), LINFH(5, 1000))
    ITER = 0
    DO WHILE (ITER .LT. NITER)
        IF (ROUTINES .OR. ITER .EQ. NITER - 1) THEN
            CALL ONE_ITERATION
            ITER = ITER + 1
            LINFH(:, 1) = LINF
            CALL AFTER_CHUNK(ITER, 1)
        ELSE
            CHUNK = MIN(NITER - 1 - ITER, 1000 - MOD(ITER, 1000), &
                        20000 - MOD(NITEROLD + ITER, 20000))
            BAD = 0
            I = TDT_ITERATE(CTX, INT(CHUNK, C_INT32_T), L2H, LINFH, BAD)
            IF (BAD .GT. 0) THEN
                WRITE(0, '(A, I0)') 'tdt: non-finite state at iteration ', NITEROLD + ITER + BAD
                STOP 2
            END IF
            CALL TDT_CHECK(I, CTX, 'RK4 iterations')
            ITER = ITER + CHUNK
            CALL AFTER_CHUNK(ITER, CHUNK)
        END IF
    END DO

    NITEROLD = NITEROLD + NITER
    CALL WRITE_RESTART('restart.in', NITEROLD)
    CALL POSTPROCESS
    I = TDT_DESTROY(CTX)

CONTAINS

    ! One iteration with the routine-level entry points, in the reference's
    ! CALL order (turbnstwod.f90:97-168).
    SUBROUTINE ONE_ITERATION
        INTEGER(C_INT32_T) :: RC
        CALL TDT_CHECK(TDT_TIMESTEP(CTX), CTX, 'time step')
        CALL DISSIPATION
        CALL COMPUTE_PRIMITIVE_VARS
        CALL APPLY_BOUNDARY
        CALL COMPUTE_VISCOUS
        CALL COMPUTE_SOURCE_TERMS
        CALL COMPUTE_RESIDUAL
        CALL UPDATE_SOLUTION(ALPHA1_VAL)
        CALL COMPUTE_PRIMITIVE_VARS
        CALL APPLY_BOUNDARY
        CALL COMPUTE_SOURCE_TERMS
        CALL COMPUTE_RESIDUAL
        CALL UPDATE_SOLUTION(ALPHA2_VAL)
        CALL COMPUTE_PRIMITIVE_VARS
        CALL APPLY_BOUNDARY
        CALL COMPUTE_SOURCE_TERMS
        CALL COMPUTE_RESIDUAL
        CALL UPDATE_SOLUTION(ALPHA3_VAL)
        CALL COMPUTE_PRIMITIVE_VARS
        CALL APPLY_BOUNDARY
        CALL COMPUTE_SOURCE_TERMS
        CALL COMPUTE_RESIDUAL
        CALL UPDATE_SOLUTION(ALPHA4_VAL)
        RC = TDT_MONITOR(CTX, L2, LINF)
        IF (RC .EQ. TDT_ERR_NONFINITE) THEN
            WRITE(0, '(A, I0)') 'tdt: non-finite state at iteration ', NITEROLD + ITER + 1
            STOP 2
        END IF
        CALL TDT_CHECK(RC, CTX, 'monitor')
    END SUBROUTINE ONE_ITERATION

    ! Monitor print every 1000 iterations (turbnstwod.f90:147-163) and the
    ! periodic restart file (:170-182) after ITER iterations.
    SUBROUTINE AFTER_CHUNK(ITNOW, NLAST)
        INTEGER, INTENT(IN) :: ITNOW, NLAST
        DOUBLE PRECISION :: CD(5)
        IF (MOD(ITNOW, 1000) .EQ. 0) THEN
            CD(1:4) = MAX(LINFH(1:4, NLAST), CDIFF1_INIT)
            CD(5) = MAX(LINFH(5, NLAST), CDIFF5_INIT)
            WRITE(32, 11) CD
11          FORMAT(5E15.5)
            PRINT 10, ITNOW + NITEROLD, ' ', CD(1), ' ', CD(2), ' ', CD(3), ' ', CD(4), &
                  ' ', CD(5)
10          FORMAT(I10, A5, E12.6, A5, E12.6, A5, E12.6, A5, E12.6, A5, E12.6)
        END IF
        IF (MOD(NITEROLD + ITNOW, 20000) .EQ. 0) THEN
            WRITE(RESTARTTEMPFILE, "(A7, I10, A3)") "restart", NITEROLD + ITNOW, ".in"
            CALL WRITE_RESTART(RESTARTTEMPFILE, ITNOW + NITEROLD)
        END IF
    END SUBROUTINE AFTER_CHUNK

    ! restart file in the reference's layout (turbnstwod.f90:186-195)
    SUBROUTINE WRITE_RESTART(FNAME, NOLDIT)
        CHARACTER(LEN=*), INTENT(IN) :: FNAME
        INTEGER, INTENT(IN) :: NOLDIT
        INTEGER :: II
        CALL TDT_CHECK(TDT_GET_STATE(CTX, C), CTX, 'state download')
        OPEN(31, FILE=FNAME)
        DO II = 1, NELES
            WRITE(31, *) C(II,1), C(II,2), C(II,3), C(II,4), C(II,5)
        END DO
        WRITE(31, *) RTRED
        WRITE(31, *) NPSEUDO
        WRITE(31, *) RINIT
        WRITE(31, *) NOLDIT
        CLOSE(31)
    END SUBROUTINE WRITE_RESTART

END PROGRAM TURBNS

! COMPUTE_PRIMITIVE_VARS (turbnstwod.f90:204-227)
SUBROUTINE COMPUTE_PRIMITIVE_VARS
    USE SOLVER_STATE, ONLY: CTX
    USE TDT_C
    IMPLICIT NONE
    CALL TDT_CHECK(TDT_PRIMITIVES(CTX), CTX, 'COMPUTE_PRIMITIVE_VARS')
END SUBROUTINE COMPUTE_PRIMITIVE_VARS
