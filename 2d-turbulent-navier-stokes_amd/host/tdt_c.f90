!===============================================================================
! tdt_c.f90 - iso_c_binding interfaces to libtdt (include/tdt.h)
!===============================================================================
MODULE TDT_C
    USE, INTRINSIC :: ISO_C_BINDING
    IMPLICIT NONE

    INTEGER(C_INT32_T), PARAMETER :: TDT_OK = 0, TDT_ERR_NONFINITE = -4
    INTEGER(C_INT32_T), PARAMETER :: TDT_FP64 = 0, TDT_FP32_MIXED = 1
    INTEGER(C_INT32_T), PARAMETER :: TDT_ORDER_AUTO = 2

    TYPE, BIND(C) :: TDT_MESH_T
        INTEGER(C_INT32_T) :: NODS, NELES, NGHOSTS
        TYPE(C_PTR) :: X_MM, Y_MM, NOD, NC, PARENT, GHOST, BTYPE, SIDE
    END TYPE TDT_MESH_T

    TYPE, BIND(C) :: TDT_PARAMS_T
        REAL(C_DOUBLE) :: GAMA, AMOL, CFL, PRANDTL, PRANDTLT, P0, T0, PAMB, AMACHINIT, RTRED
        REAL(C_DOUBLE) :: PINF, AMINF, DELTABL
        REAL(C_DOUBLE) :: TFLAME, ANCOMB, ACOMB, RHOPROP
        INTEGER(C_INT32_T) :: LOCAL, NPSEUDO, IAXIS, RESERVED
    END TYPE TDT_PARAMS_T

    INTERFACE
        INTEGER(C_INT32_T) FUNCTION TDT_CREATE(CTX, MESH, PARAMS, PRECISION, ORDERING) &
                BIND(C, NAME='tdt_create')
            IMPORT :: C_PTR, C_INT32_T, TDT_MESH_T, TDT_PARAMS_T
            TYPE(C_PTR), INTENT(OUT) :: CTX
            TYPE(TDT_MESH_T), INTENT(IN) :: MESH
            TYPE(TDT_PARAMS_T), INTENT(IN) :: PARAMS
            INTEGER(C_INT32_T), VALUE :: PRECISION, ORDERING
        END FUNCTION
        INTEGER(C_INT32_T) FUNCTION TDT_DESTROY(CTX) BIND(C, NAME='tdt_destroy')
            IMPORT :: C_PTR, C_INT32_T
            TYPE(C_PTR), VALUE :: CTX
        END FUNCTION
        TYPE(C_PTR) FUNCTION TDT_LAST_ERROR(CTX) BIND(C, NAME='tdt_last_error')
            IMPORT :: C_PTR
            TYPE(C_PTR), VALUE :: CTX
        END FUNCTION
        INTEGER(C_INT32_T) FUNCTION TDT_COLD_START(CTX, RINIT) BIND(C, NAME='tdt_cold_start')
            IMPORT :: C_PTR, C_INT32_T, C_DOUBLE
            TYPE(C_PTR), VALUE :: CTX
            REAL(C_DOUBLE), INTENT(OUT) :: RINIT
        END FUNCTION
        INTEGER(C_INT32_T) FUNCTION TDT_SET_STATE(CTX, C, RINIT) BIND(C, NAME='tdt_set_state')
            IMPORT :: C_PTR, C_INT32_T, C_DOUBLE
            TYPE(C_PTR), VALUE :: CTX
            REAL(C_DOUBLE), INTENT(IN) :: C(*)
            REAL(C_DOUBLE), VALUE :: RINIT
        END FUNCTION
        INTEGER(C_INT32_T) FUNCTION TDT_GET_STATE(CTX, C) BIND(C, NAME='tdt_get_state')
            IMPORT :: C_PTR, C_INT32_T, C_DOUBLE
            TYPE(C_PTR), VALUE :: CTX
            REAL(C_DOUBLE), INTENT(OUT) :: C(*)
        END FUNCTION
        INTEGER(C_INT32_T) FUNCTION TDT_TIMESTEP(CTX) BIND(C, NAME='tdt_timestep')
            IMPORT :: C_PTR, C_INT32_T
            TYPE(C_PTR), VALUE :: CTX
        END FUNCTION
        INTEGER(C_INT32_T) FUNCTION TDT_DISSIPATION(CTX) BIND(C, NAME='tdt_dissipation')
            IMPORT :: C_PTR, C_INT32_T
            TYPE(C_PTR), VALUE :: CTX
        END FUNCTION
        INTEGER(C_INT32_T) FUNCTION TDT_PRIMITIVES(CTX) BIND(C, NAME='tdt_primitives')
            IMPORT :: C_PTR, C_INT32_T
            TYPE(C_PTR), VALUE :: CTX
        END FUNCTION
        INTEGER(C_INT32_T) FUNCTION TDT_APPLY_BOUNDARY(CTX) BIND(C, NAME='tdt_apply_boundary')
            IMPORT :: C_PTR, C_INT32_T
            TYPE(C_PTR), VALUE :: CTX
        END FUNCTION
        INTEGER(C_INT32_T) FUNCTION TDT_VISCOUS(CTX) BIND(C, NAME='tdt_viscous')
            IMPORT :: C_PTR, C_INT32_T
            TYPE(C_PTR), VALUE :: CTX
        END FUNCTION
        INTEGER(C_INT32_T) FUNCTION TDT_SOURCE_TERMS(CTX) BIND(C, NAME='tdt_source_terms')
            IMPORT :: C_PTR, C_INT32_T
            TYPE(C_PTR), VALUE :: CTX
        END FUNCTION
        INTEGER(C_INT32_T) FUNCTION TDT_RESIDUAL(CTX) BIND(C, NAME='tdt_residual')
            IMPORT :: C_PTR, C_INT32_T
            TYPE(C_PTR), VALUE :: CTX
        END FUNCTION
        INTEGER(C_INT32_T) FUNCTION TDT_UPDATE(CTX, ALPHA) BIND(C, NAME='tdt_update')
            IMPORT :: C_PTR, C_INT32_T, C_DOUBLE
            TYPE(C_PTR), VALUE :: CTX
            REAL(C_DOUBLE), VALUE :: ALPHA
        END FUNCTION
        INTEGER(C_INT32_T) FUNCTION TDT_MONITOR(CTX, L2, LINF) BIND(C, NAME='tdt_monitor')
            IMPORT :: C_PTR, C_INT32_T, C_DOUBLE
            TYPE(C_PTR), VALUE :: CTX
            REAL(C_DOUBLE), INTENT(OUT) :: L2(5), LINF(5)
        END FUNCTION
        INTEGER(C_INT32_T) FUNCTION TDT_ITERATE(CTX, NITER, L2, LINF, FIRST_BAD) &
                BIND(C, NAME='tdt_iterate')
            IMPORT :: C_PTR, C_INT32_T, C_DOUBLE
            TYPE(C_PTR), VALUE :: CTX
            INTEGER(C_INT32_T), VALUE :: NITER
            REAL(C_DOUBLE), INTENT(OUT) :: L2(*), LINF(*)
            INTEGER(C_INT32_T), INTENT(OUT) :: FIRST_BAD
        END FUNCTION
        INTEGER(C_INT32_T) FUNCTION TDT_GET_ARRAY(CTX, NAME, OUT, LEN) BIND(C, NAME='tdt_get_array')
            IMPORT :: C_PTR, C_INT32_T, C_DOUBLE, C_CHAR
            TYPE(C_PTR), VALUE :: CTX
            CHARACTER(KIND=C_CHAR), INTENT(IN) :: NAME(*)
            REAL(C_DOUBLE), INTENT(OUT) :: OUT(*)
            INTEGER(C_INT32_T), VALUE :: LEN
        END FUNCTION
        ! POSTPROCESS node values: NODAL(NODS,6) = U, V, P [bar], RHO, MACH, T
        INTEGER(C_INT32_T) FUNCTION TDT_POSTPROCESS(CTX, NODAL) BIND(C, NAME='tdt_postprocess')
            IMPORT :: C_PTR, C_INT32_T, C_DOUBLE
            TYPE(C_PTR), VALUE :: CTX
            REAL(C_DOUBLE), INTENT(OUT) :: NODAL(*)
        END FUNCTION
    END INTERFACE

CONTAINS

    ! Stops the program with the library's message when a call failed.
    SUBROUTINE TDT_CHECK(RC, CTX, WHAT)
        INTEGER(C_INT32_T), INTENT(IN) :: RC
        TYPE(C_PTR), INTENT(IN) :: CTX
        CHARACTER(LEN=*), INTENT(IN) :: WHAT
        CHARACTER(KIND=C_CHAR), POINTER :: MSG(:)
        TYPE(C_PTR) :: P
        INTEGER :: K
        IF (RC .EQ. TDT_OK) RETURN
        P = TDT_LAST_ERROR(CTX)
        CALL C_F_POINTER(P, MSG, [1024])
        K = 0
        DO WHILE (K .LT. 1024)
            IF (MSG(K+1) .EQ. C_NULL_CHAR) EXIT
            K = K + 1
        END DO
        WRITE(0, '(A, A, A, I0, A)', ADVANCE='NO') 'tdt error in ', WHAT, ' (code ', RC, '): '
        WRITE(0, '(1024A)') MSG(1:K)
        STOP 1
    END SUBROUTINE TDT_CHECK

END MODULE TDT_C
