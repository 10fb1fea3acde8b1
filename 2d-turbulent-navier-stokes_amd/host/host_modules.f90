!===============================================================================
! host_modules.f90 - Fortran host of libtdt.
!
! Same module names and argument-less routines as the reference solver
! (constants.f90, solver_state.f90, geometry.f90, boundary.f90, viscous.f90,
! source_terms.f90, residual.f90), so its CALL sites stay unchanged.  The
! physics runs in libtdt on the GPU: each routine below forwards to the C ABI
! entry point that replaces it.  File formats (flow.in, grid.in, bc.in,
! restart.in, tec_plot.dat) are the reference's.
!===============================================================================
MODULE CONSTANTS
    IMPLICIT NONE
    DOUBLE PRECISION, PARAMETER :: ALPHA1_VAL = 0.25D0, ALPHA2_VAL = 1.0D0/3.0D0
    DOUBLE PRECISION, PARAMETER :: ALPHA3_VAL = 0.5D0, ALPHA4_VAL = 1.0D0
    ! monitor floors: default-real literals, as in constants.f90:44-45
    DOUBLE PRECISION, PARAMETER :: CDIFF1_INIT = 1.0E-10
    DOUBLE PRECISION, PARAMETER :: CDIFF5_INIT = 1.0E-20
    DOUBLE PRECISION, PARAMETER :: MM_TO_M = 0.001D0, PA_TO_BAR = 0.00001D0
    DOUBLE PRECISION, PARAMETER :: RUNIV_VAL = 8314.0D0
END MODULE CONSTANTS

MODULE SOLVER_STATE
    USE, INTRINSIC :: ISO_C_BINDING
    USE TDT_C
    USE CONSTANTS, ONLY: PA_TO_BAR
    IMPLICIT NONE
    CHARACTER*50 :: FNAMEGRID, FNAMEBC, RESTARTTEMPFILE
    INTEGER :: NODS, NELES, NGHOSTS, NMAX, LOCAL, NITER, ISTART, IAXIS, NPSEUDO, NITEROLD
    DOUBLE PRECISION :: GAMA, AMOL, CFL, PRANDTL, PRANDTLT, P0, T0, PAMB, AMACHINIT, RTRED
    DOUBLE PRECISION :: RINIT, PINF, AMINF, DELTABL, RR
    DOUBLE PRECISION :: TFLAME, ANCOMB, ACOMB, RHOPROP
    ! host copies, allocated to the sizes grid.in / bc.in give (no NN limit)
    DOUBLE PRECISION, ALLOCATABLE, TARGET :: X(:), Y(:)
    INTEGER(C_INT32_T), ALLOCATABLE, TARGET :: NOD(:,:), NC(:,:)
    INTEGER(C_INT32_T), ALLOCATABLE, TARGET :: NPARENT(:), NGHOST(:), NTYPE(:), NSIDE(:)
    DOUBLE PRECISION, ALLOCATABLE :: C(:,:)             ! C(I, 1..5) = C1..C5
    DOUBLE PRECISION, ALLOCATABLE :: U(:), V(:), P(:), T(:)
    TYPE(C_PTR) :: CTX = C_NULL_PTR
    INTEGER(C_INT32_T) :: PRECISION_MODE = TDT_FP64

CONTAINS

    ! DISSIPATION (solver_state.f90:70-225)
    SUBROUTINE DISSIPATION
        CALL TDT_CHECK(TDT_DISSIPATION(CTX), CTX, 'DISSIPATION')
    END SUBROUTINE DISSIPATION

    SUBROUTINE FETCH(NAME, A, N)
        CHARACTER(LEN=*), INTENT(IN) :: NAME
        INTEGER, INTENT(IN) :: N
        DOUBLE PRECISION, INTENT(OUT) :: A(N)
        CALL TDT_CHECK(TDT_GET_ARRAY(CTX, TRIM(NAME)//C_NULL_CHAR, A, INT(N, C_INT32_T)), CTX, NAME)
    END SUBROUTINE FETCH

    ! POSTPROCESS (solver_state.f90:230-283): cell-to-node averages of P, U, V, T
    ! over the cells and boundary ghosts around each node, written as tec_plot.dat.
    SUBROUTINE POSTPROCESS
        ! solver_state.f90:230-283: the node averages and derived columns are
        ! computed on the device (TDT_POSTPROCESS, same order and expressions);
        ! the host writes tec_plot.dat in the reference's format
        DOUBLE PRECISION, ALLOCATABLE :: NODAL(:, :)
        INTEGER :: I, IE
        PRINT*, 'Postprocessing.'
        PRINT*, ' '
        ALLOCATE(NODAL(NODS, 6))
        CALL TDT_CHECK(TDT_POSTPROCESS(CTX, NODAL), CTX, 'POSTPROCESS')
        OPEN(UNIT=26, FILE='tec_plot.dat')
        WRITE(26,*)'VARIABLES = "X", "Y", "U", "V", "P", "RHO", "MACH", "T"'
        WRITE(26,*)'ZONE F = FEPOINT, ET = QUADRILATERAL, N = ',NODS,', E = ',NELES
        DO I = 1, NODS
            WRITE(26,*) X(I), Y(I), NODAL(I,1), NODAL(I,2), NODAL(I,3), NODAL(I,4), &
                        NODAL(I,5), NODAL(I,6)
        END DO
        DO IE = 1, NELES
            WRITE(26,*) NOD(IE,1), NOD(IE,2), NOD(IE,3), NOD(IE,4)
        END DO
        CLOSE(26)
    END SUBROUTINE POSTPROCESS

END MODULE SOLVER_STATE

MODULE GEOMETRY
    USE SOLVER_STATE
    USE CONSTANTS, ONLY: MM_TO_M
    IMPLICIT NONE
CONTAINS

    ! COMPUTE_GEOMETRY (geometry.f90:22-151): reads grid.in and bc.in.  The
    ! geometry itself is computed on the GPU by DEVICE_SETUP.
    SUBROUTINE COMPUTE_GEOMETRY
        INTEGER :: I, N
        DOUBLE PRECISION :: XV, YV
        OPEN(21, FILE=FNAMEGRID)
        READ(21, *) NODS, NELES
        ALLOCATE(X(NODS), Y(NODS), NOD(NELES, 4), NC(NELES, 4))
        DO I = 1, NODS
            READ(21, *) N, XV, YV
            X(N) = XV; Y(N) = YV
        END DO
        DO I = 1, NELES
            READ(21, *) N, NOD(I,1), NOD(I,2), NOD(I,3), NOD(I,4), &
                        NC(I,1), NC(I,2), NC(I,3), NC(I,4)
        END DO
        CLOSE(21)
        OPEN(22, FILE=FNAMEBC)
        READ(22, *) NGHOSTS
        ALLOCATE(NPARENT(NGHOSTS), NGHOST(NGHOSTS), NTYPE(NGHOSTS), NSIDE(NGHOSTS))
        DO I = 1, NGHOSTS
            READ(22, *) NPARENT(I), NGHOST(I), NTYPE(I), NSIDE(I)
        END DO
        CLOSE(22)
        NMAX = NELES
        DO I = 1, NGHOSTS
            NMAX = MAX(NMAX, NGHOST(I))
        END DO
    END SUBROUTINE COMPUTE_GEOMETRY

    ! Uploads the mesh and creates the device context (after restart.in, whose
    ! RTRED / NPSEUDO replace flow.in's, turbnstwod.f90:80-81).
    SUBROUTINE DEVICE_SETUP
        TYPE(TDT_MESH_T) :: M
        TYPE(TDT_PARAMS_T) :: PRM
        M%NODS = NODS; M%NELES = NELES; M%NGHOSTS = NGHOSTS
        M%X_MM = C_LOC(X); M%Y_MM = C_LOC(Y)
        M%NOD = C_LOC(NOD); M%NC = C_LOC(NC)
        M%PARENT = C_LOC(NPARENT); M%GHOST = C_LOC(NGHOST)
        M%BTYPE = C_LOC(NTYPE); M%SIDE = C_LOC(NSIDE)
        PRM%GAMA = GAMA; PRM%AMOL = AMOL; PRM%CFL = CFL; PRM%PRANDTL = PRANDTL
        PRM%PRANDTLT = PRANDTLT; PRM%P0 = P0; PRM%T0 = T0; PRM%PAMB = PAMB
        PRM%AMACHINIT = AMACHINIT; PRM%RTRED = RTRED
        PRM%PINF = PINF; PRM%AMINF = AMINF; PRM%DELTABL = DELTABL
        PRM%TFLAME = TFLAME; PRM%ANCOMB = ANCOMB; PRM%ACOMB = ACOMB; PRM%RHOPROP = RHOPROP
        PRM%LOCAL = LOCAL; PRM%NPSEUDO = NPSEUDO; PRM%IAXIS = IAXIS; PRM%RESERVED = 0
        CALL TDT_CHECK(TDT_CREATE(CTX, M, PRM, PRECISION_MODE, TDT_ORDER_AUTO), C_NULL_PTR, &
                       'COMPUTE_GEOMETRY')
        ! the file stores mm; tec_plot.dat prints metres (geometry.f90:35-36)
        X = X * MM_TO_M
        Y = Y * MM_TO_M
    END SUBROUTINE DEVICE_SETUP

END MODULE GEOMETRY

MODULE BOUNDARY
    USE SOLVER_STATE
    IMPLICIT NONE
CONTAINS
    ! APPLY_BOUNDARY (boundary.f90:24-48)
    SUBROUTINE APPLY_BOUNDARY
        CALL TDT_CHECK(TDT_APPLY_BOUNDARY(CTX), CTX, 'APPLY_BOUNDARY')
    END SUBROUTINE APPLY_BOUNDARY
END MODULE BOUNDARY

MODULE VISCOUS
    USE SOLVER_STATE
    IMPLICIT NONE
CONTAINS
    ! COMPUTE_VISCOUS (viscous.f90:29-72)
    SUBROUTINE COMPUTE_VISCOUS
        CALL TDT_CHECK(TDT_VISCOUS(CTX), CTX, 'COMPUTE_VISCOUS')
    END SUBROUTINE COMPUTE_VISCOUS
END MODULE VISCOUS

MODULE SOURCE_TERMS
    USE SOLVER_STATE
    IMPLICIT NONE
CONTAINS
    ! COMPUTE_SOURCE_TERMS (source_terms.f90:24-104)
    SUBROUTINE COMPUTE_SOURCE_TERMS
        CALL TDT_CHECK(TDT_SOURCE_TERMS(CTX), CTX, 'COMPUTE_SOURCE_TERMS')
    END SUBROUTINE COMPUTE_SOURCE_TERMS
END MODULE SOURCE_TERMS

MODULE RESIDUAL
    USE SOLVER_STATE
    IMPLICIT NONE
CONTAINS
    ! COMPUTE_RESIDUAL (residual.f90:23-62)
    SUBROUTINE COMPUTE_RESIDUAL
        CALL TDT_CHECK(TDT_RESIDUAL(CTX), CTX, 'COMPUTE_RESIDUAL')
    END SUBROUTINE COMPUTE_RESIDUAL

    ! UPDATE_SOLUTION (residual.f90:67-83)
    SUBROUTINE UPDATE_SOLUTION(ALPHA_COEFF)
        DOUBLE PRECISION, INTENT(IN) :: ALPHA_COEFF
        CALL TDT_CHECK(TDT_UPDATE(CTX, ALPHA_COEFF), CTX, 'UPDATE_SOLUTION')
    END SUBROUTINE UPDATE_SOLUTION
END MODULE RESIDUAL
