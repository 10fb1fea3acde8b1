!===============================================================================
! ref_driver.f90 -- test infrastructure only (never shipped, never measured).
!
! A dump driver linked against the reference solver's own modules (CONSTANTS,
! SOLVER_STATE, GEOMETRY, VISCOUS, SOURCE_TERMS, BOUNDARY, RESIDUAL and the
! external COMPUTE_PRIMITIVE_VARS), built by oracle/build_ref.sh.  It repeats
! the main program's set-up and RK4 loop (turbnstwod.f90:20-168) so that it can
!   * write the per-iteration L2 and L-infinity of C - COLD   -> hist.bin
!   * write the final interior state C1..C5                   -> final.bin
!   * dump every array produced by each reference routine during one extra
!     iteration                                                -> dump.bin
! The loop restated here is checked against the reference executable itself
! (restart.in after the same number of iterations must agree bit for bit).
!
! Inputs in the working directory:
!   flow.in   (reference format, turbnstwod.f90:20-23)
!   driver.in  NRUN IDUMP
!              [PINF AMINF DELTABL]
!              [TFLAME ANCOMB ACOMB RHOPROP]
!              NRUN iterations are run; IDUMP=1 dumps iteration NRUN+1.
!              The optional lines give the BC type-7 and type-6 inputs that
!              the reference never assigns (boundary.f90:138-153, 310-327).
!===============================================================================
PROGRAM REF_DRIVER
    USE CONSTANTS
    USE SOLVER_STATE
    USE GEOMETRY
    USE VISCOUS
    USE SOURCE_TERMS
    USE BOUNDARY
    USE RESIDUAL
    IMPLICIT NONE
    INTEGER :: NRUN, IDUMP, IOS, IOS6, NMAX, IT, ISTG
    DOUBLE PRECISION :: BLP(3), B6P(4), L2(5), LINF(5), STAGEA(4)
    CHARACTER(LEN=2) :: SS

    OPEN(25, FILE='flow.in')
    READ(25, *) FNAMEGRID, FNAMEBC, GAMA, AMOL, LOCAL, NITER, ISTART, CFL, &
                PRANDTL, PRANDTLT, P0, T0, PAMB, AMACHINIT, RTRED, NPSEUDO, IAXIS
    CLOSE(25)

    BLP = 0.0D0
    OPEN(26, FILE='driver.in')
    READ(26, *) NRUN, IDUMP
    READ(26, *, IOSTAT=IOS) BLP(1), BLP(2), BLP(3)
    B6P = 0.0D0
    IOS6 = 1
    IF (IOS .EQ. 0) READ(26, *, IOSTAT=IOS6) B6P(1), B6P(2), B6P(3), B6P(4)
    CLOSE(26)
    IF (IOS .NE. 0) BLP = 0.0D0
    IF (IOS6 .NE. 0) B6P = 0.0D0
    PINF = BLP(1); AMINF = BLP(2); DELTABL = BLP(3)
    TFLAME = B6P(1); ANCOMB = B6P(2); ACOMB = B6P(3); RHOPROP = B6P(4)

    RUNIV = RUNIV_VAL
    GTH = GTH_VAL
    CONVKSC = CONVKSC_VAL

    CALL COMPUTE_GEOMETRY

    RR = RUNIV / AMOL
    CV = RR / (GAMA - 1.0D0)
    CP = CV * GAMA
    RHO0 = P0 / (RR*T0)
    S0 = P0 / (RHO0**GAMA)
    TINF = T0 / (1.0D0 + 0.5D0*(GAMA-1.0D0)*AMINF*AMINF)

    VISU = 0.0D0; VISV = 0.0D0; VIST = 0.0D0; VISR = 0.0D0
    SOURCE5 = 0.0D0
    ITEST = 0
    CFLTURB = CFLTURB_VAL
    NITEROLD = 0; NOLD = 0

    IF (ISTART .EQ. 0) THEN
        DYNRATIO = 1.0D0 + (GAMA-1.0D0)*0.5D0*AMACHINIT**2.0D0
        TINIT = T0 / DYNRATIO
        PINIT = P0 / (DYNRATIO**(GAMA/(GAMA-1.0D0)))
        UINIT = AMACHINIT * DSQRT(GAMA*RR*TINIT)
        VINIT = 0.0D0
        EINIT = CV*TINIT + 0.5D0*(UINIT**2.0D0 + VINIT**2.0D0)
        RHOINIT = PINIT / (RR*TINIT)
        AMUINIT = (0.00000011848D0)*DSQRT(AMOL)*(TINIT**0.6D0)
        RINIT = AMUINIT / RHOINIT
        RTOLERANCE = RINIT * RTRED
        DO I = 1, NELES
            C1(I) = RHOINIT
            C2(I) = RHOINIT*UINIT
            C3(I) = RHOINIT*VINIT
            C4(I) = RHOINIT*EINIT
            C5(I) = RHOINIT*RINIT
        END DO
    ELSE
        OPEN(55, FILE='restart.in')
        DO I = 1, NELES
            READ(55, *) C1(I), C2(I), C3(I), C4(I), C5(I)
        END DO
        READ(55, *) RTRED
        READ(55, *) NPSEUDO
        READ(55, *) RINIT
        READ(55, *) NITEROLD
        CLOSE(55)
        RTOLERANCE = RINIT * RTRED
    END IF

    DO I = 1, NELES
        COLD1(I)=C1(I); COLD2(I)=C2(I); COLD3(I)=C3(I)
        COLD4(I)=C4(I); COLD5(I)=C5(I)
    END DO

    NMAX = NELES
    DO I = 1, NGHOSTS
        NMAX = MAX(NMAX, NGHOST(I))
    END DO

    OPEN(78, FILE='hist.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
    DO IT = 1, NRUN
        ITER = IT
        CALL TIMESTEP
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
        CALL MONITOR_AND_SAVE(IT)
    END DO
    CLOSE(78)

    IF (IDUMP .EQ. 1) THEN
        STAGEA(1) = ALPHA1_VAL; STAGEA(2) = ALPHA2_VAL
        STAGEA(3) = ALPHA3_VAL; STAGEA(4) = ALPHA4_VAL
        OPEN(77, FILE='dump.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
        CALL DUMPA('XCC', XCC, NMAX); CALL DUMPA('YCC', YCC, NMAX)
        CALL DUMPA('VOL', VOL, NMAX); CALL DUMPA('DL', DL, NELES)
        CALL DUMPA('SC1X', SC1X, NELES); CALL DUMPA('SC1Y', SC1Y, NELES)
        CALL DUMPA('SC2X', SC2X, NELES); CALL DUMPA('SC2Y', SC2Y, NELES)
        CALL DUMPA('SC3X', SC3X, NELES); CALL DUMPA('SC3Y', SC3Y, NELES)
        CALL DUMPA('SC4X', SC4X, NELES); CALL DUMPA('SC4Y', SC4Y, NELES)
        CALL DUMPS('RTOL', RTOLERANCE)
        CALL DUMPA('C1', C1, NELES); CALL DUMPA('C2', C2, NELES)
        CALL DUMPA('C3', C3, NELES); CALL DUMPA('C4', C4, NELES)
        CALL DUMPA('C5', C5, NELES)
        ITER = NRUN + 1
        CALL TIMESTEP
        CALL DUMPA('TS_U', U, NELES); CALL DUMPA('TS_V', V, NELES)
        CALL DUMPA('TS_T', T, NELES); CALL DUMPA('TS_P', P, NELES)
        CALL DUMPA('TS_DT', DT, NELES); CALL DUMPS('DTMIN', DTMIN)
        CALL DISSIPATION
        CALL DUMPA('DS_P', P, NMAX); CALL DUMPA('DS_DT', DT, NMAX)
        CALL DUMPA('ANU', ANU, NMAX); CALL DUMPA('ANURT', ANURT, NMAX)
        CALL DUMPA('D2Q1', D2Q1, NMAX); CALL DUMPA('D2Q2', D2Q2, NMAX)
        CALL DUMPA('D2Q3', D2Q3, NMAX); CALL DUMPA('D2Q4', D2Q4, NMAX)
        CALL DUMPA('D2Q5', D2Q5, NMAX)
        CALL DUMPA('D241', D241, NELES); CALL DUMPA('D242', D242, NELES)
        CALL DUMPA('D243', D243, NELES); CALL DUMPA('D244', D244, NELES)
        CALL DUMPA('D245', D245, NELES)
        DO ISTG = 1, 4
            WRITE(SS, '(A1,I1)') 'S', ISTG
            CALL COMPUTE_PRIMITIVE_VARS
            CALL APPLY_BOUNDARY
            CALL DUMPA(SS//'_U', U, NMAX); CALL DUMPA(SS//'_V', V, NMAX)
            CALL DUMPA(SS//'_T', T, NMAX); CALL DUMPA(SS//'_P', P, NMAX)
            CALL DUMPA(SS//'_C1', C1, NMAX); CALL DUMPA(SS//'_C2', C2, NMAX)
            CALL DUMPA(SS//'_C3', C3, NMAX); CALL DUMPA(SS//'_C4', C4, NMAX)
            CALL DUMPA(SS//'_C5', C5, NMAX)
            CALL DUMPA(SS//'_F1', F1, NMAX); CALL DUMPA(SS//'_F2', F2, NMAX)
            CALL DUMPA(SS//'_F3', F3, NMAX); CALL DUMPA(SS//'_F4', F4, NMAX)
            CALL DUMPA(SS//'_F5', F5, NMAX)
            CALL DUMPA(SS//'_G1', G1, NMAX); CALL DUMPA(SS//'_G2', G2, NMAX)
            CALL DUMPA(SS//'_G3', G3, NMAX); CALL DUMPA(SS//'_G4', G4, NMAX)
            CALL DUMPA(SS//'_G5', G5, NMAX)
            IF (ISTG .EQ. 1) THEN
                CALL COMPUTE_VISCOUS
                CALL DUMPA('UN', UN, NODS); CALL DUMPA('VN', VN, NODS)
                CALL DUMPA('TN', TN, NODS); CALL DUMPA('RHON', RHON, NODS)
                CALL DUMPA('RN', RN, NODS)
                CALL DUMPA('VISU', VISU, NELES); CALL DUMPA('VISV', VISV, NELES)
                CALL DUMPA('VIST', VIST, NELES); CALL DUMPA('VISR', VISR, NELES)
            END IF
            CALL COMPUTE_SOURCE_TERMS
            CALL DUMPA(SS//'_SRC3', SOURCE3, NELES)
            CALL DUMPA(SS//'_SRC5', SOURCE5, NELES)
            CALL COMPUTE_RESIDUAL
            CALL DUMPA(SS//'_R1', R1, NELES); CALL DUMPA(SS//'_R2', R2, NELES)
            CALL DUMPA(SS//'_R3', R3, NELES); CALL DUMPA(SS//'_R4', R4, NELES)
            CALL DUMPA(SS//'_R5', R5, NELES)
            CALL UPDATE_SOLUTION(STAGEA(ISTG))
            CALL DUMPA(SS//'_Q1', C1, NELES); CALL DUMPA(SS//'_Q2', C2, NELES)
            CALL DUMPA(SS//'_Q3', C3, NELES); CALL DUMPA(SS//'_Q4', C4, NELES)
            CALL DUMPA(SS//'_Q5', C5, NELES)
        END DO
        CLOSE(77)
        OPEN(78, FILE='hist_dump.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
        CALL MONITOR_AND_SAVE(NRUN + 1)
        CLOSE(78)
    END IF

    OPEN(79, FILE='final.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
    WRITE(79) NELES
    WRITE(79) C1(1:NELES), C2(1:NELES), C3(1:NELES), C4(1:NELES), C5(1:NELES)
    CLOSE(79)

CONTAINS

    ! Time step and primitive variables (turbnstwod.f90:97-114)
    SUBROUTINE TIMESTEP
        DTMIN = 100000.0D0
        DO I = 1, NELES
            RHO = C1(I)
            U(I) = C2(I)/RHO
            V(I) = C3(I)/RHO
            VEL = DSQRT(U(I)*U(I) + V(I)*V(I))
            E = C4(I)/RHO
            T(I) = (E - 0.5D0*(U(I)*U(I)+V(I)*V(I)))/CV
            P(I) = RHO*RR*T(I)
            DT(I) = CFL*DL(I) / (DABS(VEL) + DSQRT(GAMA*RR*T(I)))
            DTMIN = DMIN1(DTMIN, DT(I))
        END DO
        IF (LOCAL .EQ. 0) THEN
            DO I = 1, NELES
                DT(I) = DTMIN
            END DO
        END IF
    END SUBROUTINE TIMESTEP

    ! Per-iteration L2 / L-infinity of C - COLD, the reference monitor print
    ! (turbnstwod.f90:147-163) and COLD <- C (turbnstwod.f90:165-168).
    SUBROUTINE MONITOR_AND_SAVE(ITNOW)
        INTEGER, INTENT(IN) :: ITNOW
        DOUBLE PRECISION :: DD
        INTEGER :: II
        L2 = 0.0D0; LINF = 0.0D0
        DO II = 1, NELES
            DD = C1(II)-COLD1(II); L2(1) = L2(1) + DD*DD; LINF(1) = MAX(LINF(1), DABS(DD))
            DD = C2(II)-COLD2(II); L2(2) = L2(2) + DD*DD; LINF(2) = MAX(LINF(2), DABS(DD))
            DD = C3(II)-COLD3(II); L2(3) = L2(3) + DD*DD; LINF(3) = MAX(LINF(3), DABS(DD))
            DD = C4(II)-COLD4(II); L2(4) = L2(4) + DD*DD; LINF(4) = MAX(LINF(4), DABS(DD))
            DD = C5(II)-COLD5(II); L2(5) = L2(5) + DD*DD; LINF(5) = MAX(LINF(5), DABS(DD))
        END DO
        DO II = 1, 5
            L2(II) = DSQRT(L2(II) / DBLE(NELES))
        END DO
        WRITE(78) ITNOW, L2, LINF
        CDIFF1 = CDIFF1_INIT; CDIFF2 = CDIFF1_INIT; CDIFF3 = CDIFF1_INIT
        CDIFF4 = CDIFF1_INIT; CDIFF5 = CDIFF5_INIT
        IF (MOD(ITNOW,1000) .EQ. 0) THEN
            DO II = 1, NELES
                CDIFF1 = DMAX1(DABS(COLD1(II)-C1(II)), CDIFF1)
                CDIFF2 = DMAX1(DABS(COLD2(II)-C2(II)), CDIFF2)
                CDIFF3 = DMAX1(DABS(COLD3(II)-C3(II)), CDIFF3)
                CDIFF4 = DMAX1(DABS(COLD4(II)-C4(II)), CDIFF4)
                CDIFF5 = DMAX1(DABS(COLD5(II)-C5(II)), CDIFF5)
            END DO
            WRITE(32, 11) CDIFF1, CDIFF2, CDIFF3, CDIFF4, CDIFF5
11          FORMAT(5E15.5)
            PRINT 10, ITNOW+NITEROLD, ' ', CDIFF1, ' ', CDIFF2, ' ', CDIFF3, &
                   ' ', CDIFF4, ' ', CDIFF5
10          FORMAT(I10, A5, E12.6, A5, E12.6, A5, E12.6, A5, E12.6, A5, E12.6)
        END IF
        DO II = 1, NELES
            COLD1(II)=C1(II); COLD2(II)=C2(II); COLD3(II)=C3(II)
            COLD4(II)=C4(II); COLD5(II)=C5(II)
        END DO
    END SUBROUTINE MONITOR_AND_SAVE

    SUBROUTINE DUMPA(NAME, A, NLEN)
        CHARACTER(LEN=*), INTENT(IN) :: NAME
        INTEGER, INTENT(IN) :: NLEN
        DOUBLE PRECISION, INTENT(IN) :: A(NLEN)
        CHARACTER(LEN=16) :: TAG
        TAG = NAME
        WRITE(77) TAG, NLEN, A(1:NLEN)
    END SUBROUTINE DUMPA

    SUBROUTINE DUMPS(NAME, S)
        CHARACTER(LEN=*), INTENT(IN) :: NAME
        DOUBLE PRECISION, INTENT(IN) :: S
        DOUBLE PRECISION :: A(1)
        A(1) = S
        CALL DUMPA(NAME, A, 1)
    END SUBROUTINE DUMPS

END PROGRAM REF_DRIVER
