! tritok_amd_bind.f90 -- ISO_C_BINDING interface of libtritok_amd.so
! (include/tritok_bem.h) for the Tritok Fortran host.
!
! The derived types mirror the C structs field for field; arrays are passed
! as the reference holds them (column-major, 1-based index arrays, INTEGER =
! C int32, REAL(8) = C double, REAL(4) = C float).  The adapter subroutines at
! the end keep the argument lists of the reference routines they replace, so
! a call site changes only its subroutine name (and keeps a device handle).
MODULE tritok_amd_bind
  USE, INTRINSIC :: ISO_C_BINDING
  IMPLICIT NONE

  TYPE, BIND(C) :: tt_mesh_view
    INTEGER(C_INT32_T) :: nnodes, nbnodes, nbelem, nicell, nq
    TYPE(C_PTR) :: x, ibc, idc, ibf, xq, flpm, gbn, lbn, nx, ny, nz
  END TYPE

  TYPE, BIND(C) :: tt_quad
    INTEGER(C_INT32_T) :: kmBr, kmBs, kmDr, kmDs, iiDiv
    TYPE(C_PTR) :: gi, ome, ng1, ng2
  END TYPE

  TYPE, BIND(C) :: tt_case_view
    TYPE(tt_mesh_view) :: mesh
    TYPE(tt_quad) :: quad
    INTEGER(C_INT32_T) :: copy
    TYPE(C_PTR) :: kode, kobc, wkode, wqkode, Tkode, Tqkode
    INTEGER(C_INT32_T) :: nunk(3), nbk(3), Tnunk, Tnb
    TYPE(C_PTR) :: kmu
    INTEGER(C_INT32_T) :: nsmcol, nrhscol, iDT
    REAL(C_DOUBLE) :: ReN, RaN, PrN, gx, gy, gz, tstep
    REAL(C_DOUBLE) :: urBw(3), urDv(3), urDw(3), urDT, eps
    INTEGER(C_INT32_T) :: nnlit, lsqs_maxit, iDLSE
    REAL(C_FLOAT) :: lsqs_eps, dlse_r, dlse_epsth
    INTEGER(C_INT32_T) :: iBw
  END TYPE

  INTEGER(C_INT32_T), PARAMETER :: TT_PREC_FP64 = 0, TT_PREC_FP32 = 1, TT_PREC_BF16 = 2, TT_PREC_FP16 = 3

  INTERFACE
    FUNCTION tt_last_error() BIND(C, NAME="tt_last_error")
      IMPORT :: C_PTR
      TYPE(C_PTR) :: tt_last_error
    END FUNCTION

    ! FMATkmdcSteady / corMin_kmdcSteady (lib_fmat.f:3, lib_int.f:1254)
    FUNCTION tt_macro_integrals(m, q, copy, H, G, B, Abdx, Abdy, Abdz, Htx, Hty, Htz, Dx, Dy, Dz, ont) &
        BIND(C, NAME="tt_macro_integrals")
      IMPORT :: C_INT, C_INT32_T, C_DOUBLE, tt_mesh_view, tt_quad
      TYPE(tt_mesh_view), INTENT(IN) :: m
      TYPE(tt_quad), INTENT(IN) :: q
      INTEGER(C_INT32_T), VALUE :: copy
      REAL(C_DOUBLE) :: H(*), G(*), B(*), Abdx(*), Abdy(*), Abdz(*), Htx(*), Hty(*), Htz(*)
      REAL(C_DOUBLE) :: Dx(*), Dy(*), Dz(*), ont(11)
      INTEGER(C_INT) :: tt_macro_integrals
    END FUNCTION

    ! FMATmodHelm / corMin_modHelm (lib_mHelm.f:2, :1400)
    FUNCTION tt_macro_integrals_mh(m, q, copy, diff, qdiff, beta, H, G, B, Abdx, Abdy, Abdz, Htx, Hty, Htz, &
                                   Dx, Dy, Dz, ont, mumax) BIND(C, NAME="tt_macro_integrals_mh")
      IMPORT :: C_INT, C_INT32_T, C_DOUBLE, tt_mesh_view, tt_quad
      TYPE(tt_mesh_view), INTENT(IN) :: m
      TYPE(tt_quad), INTENT(IN) :: q
      INTEGER(C_INT32_T), VALUE :: copy
      REAL(C_DOUBLE), INTENT(IN) :: diff(*), qdiff(*)
      REAL(C_DOUBLE), VALUE :: beta
      REAL(C_DOUBLE) :: H(*), G(*), B(*), Abdx(*), Abdy(*), Abdz(*), Htx(*), Hty(*), Htz(*)
      REAL(C_DOUBLE) :: Dx(*), Dy(*), Dz(*), ont(11), mumax
      INTEGER(C_INT) :: tt_macro_integrals_mh
    END FUNCTION

    ! FMATsdkm / sdKm_corInt rows zac:kon (lib_sdkm.f:1286, :2377)
    FUNCTION tt_sd_integrals(m, q, row0, row1, H, Htx, Hty, Htz, Dx, Dy, Dz, ont) &
        BIND(C, NAME="tt_sd_integrals")
      IMPORT :: C_INT, C_INT32_T, C_DOUBLE, tt_mesh_view, tt_quad
      TYPE(tt_mesh_view), INTENT(IN) :: m
      TYPE(tt_quad), INTENT(IN) :: q
      INTEGER(C_INT32_T), VALUE :: row0, row1
      REAL(C_DOUBLE) :: H(*), Htx(*), Hty(*), Htz(*), Dx(*), Dy(*), Dz(*), ont(10)
      INTEGER(C_INT) :: tt_sd_integrals
    END FUNCTION

    ! prelsqr2 / slvlsqr2 (lib_Solve_v11.f:4501, :4401)
    FUNCTION tt_lsqr_setup(m, n, nnz, ia, ja, a, q, scale_host, h) BIND(C, NAME="tt_lsqr_setup")
      IMPORT :: C_INT, C_INT32_T, C_DOUBLE, C_PTR
      INTEGER(C_INT32_T), VALUE :: m, n, nnz, scale_host
      INTEGER(C_INT32_T), INTENT(IN) :: ia(*), ja(*)
      REAL(C_DOUBLE) :: a(*), q(*)
      TYPE(C_PTR) :: h
      INTEGER(C_INT) :: tt_lsqr_setup
    END FUNCTION
    FUNCTION tt_lsqr_solve(h, b, x, maxit, eps4, nit, istop) BIND(C, NAME="tt_lsqr_solve")
      IMPORT :: C_INT, C_INT32_T, C_DOUBLE, C_FLOAT, C_PTR
      TYPE(C_PTR), VALUE :: h
      REAL(C_DOUBLE), INTENT(IN) :: b(*)
      REAL(C_DOUBLE) :: x(*)
      INTEGER(C_INT32_T), VALUE :: maxit
      REAL(C_FLOAT), VALUE :: eps4
      INTEGER(C_INT32_T) :: nit, istop
      INTEGER(C_INT) :: tt_lsqr_solve
    END FUNCTION
    ! the same solve on device vectors, allocating nothing (b_dev, x_dev: device addresses)
    FUNCTION tt_lsqr_solve_noalloc(h, b_dev, x_dev, maxit, eps4, nit, istop) BIND(C, NAME="tt_lsqr_solve_noalloc")
      IMPORT :: C_INT, C_INT32_T, C_FLOAT, C_PTR
      TYPE(C_PTR), VALUE :: h, b_dev, x_dev
      INTEGER(C_INT32_T), VALUE :: maxit
      REAL(C_FLOAT), VALUE :: eps4
      INTEGER(C_INT32_T) :: nit, istop
      INTEGER(C_INT) :: tt_lsqr_solve_noalloc
    END FUNCTION
    SUBROUTINE tt_lsqr_free(h) BIND(C, NAME="tt_lsqr_free")
      IMPORT :: C_PTR
      TYPE(C_PTR), VALUE :: h
    END SUBROUTINE

    ! CRSxV (lib_util.f:1351)
    FUNCTION tt_csr_mv(m, n, nnz, ia, ja, a, x, y) BIND(C, NAME="tt_csr_mv")
      IMPORT :: C_INT, C_INT32_T, C_DOUBLE
      INTEGER(C_INT32_T), VALUE :: m, n, nnz
      INTEGER(C_INT32_T), INTENT(IN) :: ia(*), ja(*)
      REAL(C_DOUBLE), INTENT(IN) :: a(*), x(*)
      REAL(C_DOUBLE) :: y(*)
      INTEGER(C_INT) :: tt_csr_mv
    END FUNCTION

    ! dgefa / dgesl (lib_Solve_v11.f:787, :905)
    FUNCTION tt_dense_lu(n, a, ipvt) BIND(C, NAME="tt_dense_lu")
      IMPORT :: C_INT, C_INT32_T, C_DOUBLE
      INTEGER(C_INT32_T), VALUE :: n
      REAL(C_DOUBLE) :: a(*), ipvt(*)
      INTEGER(C_INT) :: tt_dense_lu
    END FUNCTION
    FUNCTION tt_dense_lu_solve(n, a, ipvt, b) BIND(C, NAME="tt_dense_lu_solve")
      IMPORT :: C_INT, C_INT32_T, C_DOUBLE
      INTEGER(C_INT32_T), VALUE :: n
      REAL(C_DOUBLE), INTENT(IN) :: a(*), ipvt(*)
      REAL(C_DOUBLE) :: b(*)
      INTEGER(C_INT) :: tt_dense_lu_solve
    END FUNCTION

    ! sdKm_corInt + sdKm_sysm + FormPREfm, then sdkm_solve x, y, z
    FUNCTION tt_sd_create(m, q, kode, h) BIND(C, NAME="tt_sd_create")
      IMPORT :: C_INT, C_INT32_T, C_PTR, tt_mesh_view, tt_quad
      TYPE(tt_mesh_view), INTENT(IN) :: m
      TYPE(tt_quad), INTENT(IN) :: q
      INTEGER(C_INT32_T), INTENT(IN) :: kode(*)
      TYPE(C_PTR) :: h
      INTEGER(C_INT) :: tt_sd_create
    END FUNCTION
    FUNCTION tt_sd_solve(h, vel, vort, urBw) BIND(C, NAME="tt_sd_solve")
      IMPORT :: C_INT, C_DOUBLE, C_PTR
      TYPE(C_PTR), VALUE :: h
      REAL(C_DOUBLE), INTENT(IN) :: vel(*), urBw(3)
      REAL(C_DOUBLE) :: vort(*)
      INTEGER(C_INT) :: tt_sd_solve
    END FUNCTION
    ! mixed-precision solve of one equation: FP64 iterative refinement
    FUNCTION tt_sd_refine(h, eqn, prec, nrhs, B, X, tol, maxit, iters, relres, ms) BIND(C, NAME="tt_sd_refine")
      IMPORT :: C_INT, C_INT32_T, C_DOUBLE, C_PTR
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT32_T), VALUE :: eqn, prec, nrhs, maxit
      REAL(C_DOUBLE), INTENT(IN) :: B(*)
      REAL(C_DOUBLE) :: X(*), relres, ms(2)
      REAL(C_DOUBLE), VALUE :: tol
      INTEGER(C_INT32_T) :: iters
      INTEGER(C_INT) :: tt_sd_refine
    END FUNCTION
    SUBROUTINE tt_sd_free(h) BIND(C, NAME="tt_sd_free")
      IMPORT :: C_PTR
      TYPE(C_PTR), VALUE :: h
    END SUBROUTINE

    ! main.f90 setup (:237-1084) and the steady Picard loop (:1200-1589)
    FUNCTION tt_case_create(v, vel, vort, qvort, temp, qtemp, prec, refine_tol, h) BIND(C, NAME="tt_case_create")
      IMPORT :: C_INT, C_INT32_T, C_DOUBLE, C_PTR, tt_case_view
      TYPE(tt_case_view), INTENT(IN) :: v
      REAL(C_DOUBLE), INTENT(IN) :: vel(*), vort(*), qvort(*), temp(*), qtemp(*)
      INTEGER(C_INT32_T), VALUE :: prec
      REAL(C_DOUBLE), VALUE :: refine_tol
      TYPE(C_PTR) :: h
      INTEGER(C_INT) :: tt_case_create
    END FUNCTION
    FUNCTION tt_case_picard_step(h, err, nit) BIND(C, NAME="tt_case_picard_step")
      IMPORT :: C_INT, C_INT32_T, C_DOUBLE, C_PTR
      TYPE(C_PTR), VALUE :: h
      REAL(C_DOUBLE) :: err(9)
      INTEGER(C_INT32_T) :: nit(12)
      INTEGER(C_INT) :: tt_case_picard_step
    END FUNCTION
    ! several cases per step (err(9,ncase), nit(12,ncase), status(ncase) per case)
    FUNCTION tt_case_picard_step_batch(hs, ncase, err, nit, status) BIND(C, NAME="tt_case_picard_step_batch")
      IMPORT :: C_INT, C_INT32_T, C_DOUBLE, C_PTR
      TYPE(C_PTR) :: hs(*)
      INTEGER(C_INT32_T), VALUE :: ncase
      REAL(C_DOUBLE) :: err(9, *)
      INTEGER(C_INT32_T) :: nit(12, *), status(*)
      INTEGER(C_INT) :: tt_case_picard_step_batch
    END FUNCTION
    FUNCTION tt_case_fields(h, vel, vort, qvort, temp, qtemp) BIND(C, NAME="tt_case_fields")
      IMPORT :: C_INT, C_DOUBLE, C_PTR
      TYPE(C_PTR), VALUE :: h
      REAL(C_DOUBLE) :: vel(*), vort(*), qvort(*), temp(*), qtemp(*)
      INTEGER(C_INT) :: tt_case_fields
    END FUNCTION
    SUBROUTINE tt_case_free(h) BIND(C, NAME="tt_case_free")
      IMPORT :: C_PTR
      TYPE(C_PTR), VALUE :: h
    END SUBROUTINE
    ! precision of the iteration matrices and the FP64 anchor period
    FUNCTION tt_case_set_precision(h, prec) BIND(C, NAME="tt_case_set_precision")
      IMPORT :: C_INT, C_INT32_T, C_PTR
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT32_T), VALUE :: prec
      INTEGER(C_INT) :: tt_case_set_precision
    END FUNCTION
    FUNCTION tt_case_set_grid(h, grid) BIND(C, NAME="tt_case_set_grid")
      IMPORT :: C_INT, C_INT32_T, C_PTR
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT32_T), VALUE :: grid
      INTEGER(C_INT) :: tt_case_set_grid
    END FUNCTION
    FUNCTION tt_case_set_graphs(h, on) BIND(C, NAME="tt_case_set_graphs")
      IMPORT :: C_INT, C_INT32_T, C_PTR
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT32_T), VALUE :: on
      INTEGER(C_INT) :: tt_case_set_graphs
    END FUNCTION
    FUNCTION tt_case_probe(h, stride) BIND(C, NAME="tt_case_probe")
      IMPORT :: C_INT, C_INT32_T, C_PTR
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT32_T), VALUE :: stride
      INTEGER(C_INT) :: tt_case_probe
    END FUNCTION
    FUNCTION tt_case_probe_stats(h, rows, avg_ms, sampled, launches) BIND(C, NAME="tt_case_probe_stats")
      IMPORT :: C_INT, C_INT32_T, C_INT64_T, C_DOUBLE, C_PTR
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT32_T), VALUE :: rows
      REAL(C_DOUBLE) :: avg_ms
      INTEGER(C_INT64_T) :: sampled, launches
      INTEGER(C_INT) :: tt_case_probe_stats
    END FUNCTION
    FUNCTION tt_case_set_anchor_period(h, period) BIND(C, NAME="tt_case_set_anchor_period")
      IMPORT :: C_INT, C_INT32_T, C_PTR
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT32_T), VALUE :: period
      INTEGER(C_INT) :: tt_case_set_anchor_period
    END FUNCTION

    ! AcaDo (lib_fm.f:2003): aca_CrossAprox / acaj_CrossAprox / aca_RankCrossAprox
    FUNCTION tt_aca_compress(m, n, a, typ, eps, vec, h, rank, norm) BIND(C, NAME="tt_aca_compress")
      IMPORT :: C_INT, C_INT32_T, C_DOUBLE, C_PTR
      INTEGER(C_INT32_T), VALUE :: m, n, typ
      REAL(C_DOUBLE), INTENT(IN) :: a(*), vec(*)
      REAL(C_DOUBLE), VALUE :: eps
      TYPE(C_PTR) :: h
      INTEGER(C_INT32_T) :: rank
      REAL(C_DOUBLE) :: norm
      INTEGER(C_INT) :: tt_aca_compress
    END FUNCTION
    FUNCTION tt_aca_factors(h, a, b) BIND(C, NAME="tt_aca_factors")
      IMPORT :: C_INT, C_DOUBLE, C_PTR
      TYPE(C_PTR), VALUE :: h
      REAL(C_DOUBLE) :: a(*), b(*)
      INTEGER(C_INT) :: tt_aca_factors
    END FUNCTION
    ! aca_abxv (lib_ACA.f:226)
    FUNCTION tt_aca_mv(h, x, y) BIND(C, NAME="tt_aca_mv")
      IMPORT :: C_INT, C_DOUBLE, C_PTR
      TYPE(C_PTR), VALUE :: h
      REAL(C_DOUBLE), INTENT(IN) :: x(*)
      REAL(C_DOUBLE) :: y(*)
      INTEGER(C_INT) :: tt_aca_mv
    END FUNCTION
    SUBROUTINE tt_aca_free(h) BIND(C, NAME="tt_aca_free")
      IMPORT :: C_PTR
      TYPE(C_PTR), VALUE :: h
    END SUBROUTINE

    ! Che_WcomMatrix (lib_WT.f:148), Che_FormCRS (:355), Che_WTAWx (:297)
    FUNCTION tt_wt_compress(m, n, a, kappa, h, nnz) BIND(C, NAME="tt_wt_compress")
      IMPORT :: C_INT, C_INT32_T, C_DOUBLE, C_PTR
      INTEGER(C_INT32_T), VALUE :: m, n
      REAL(C_DOUBLE), INTENT(IN) :: a(*)
      REAL(C_DOUBLE), VALUE :: kappa
      TYPE(C_PTR) :: h
      INTEGER(C_INT32_T) :: nnz
      INTEGER(C_INT) :: tt_wt_compress
    END FUNCTION
    FUNCTION tt_wt_csr(h, zvr, c, v) BIND(C, NAME="tt_wt_csr")
      IMPORT :: C_INT, C_INT32_T, C_DOUBLE, C_PTR
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT32_T) :: zvr(*), c(*)
      REAL(C_DOUBLE) :: v(*)
      INTEGER(C_INT) :: tt_wt_csr
    END FUNCTION
    FUNCTION tt_wt_mv(h, x, y) BIND(C, NAME="tt_wt_mv")
      IMPORT :: C_INT, C_DOUBLE, C_PTR
      TYPE(C_PTR), VALUE :: h
      REAL(C_DOUBLE), INTENT(IN) :: x(*)
      REAL(C_DOUBLE) :: y(*)
      INTEGER(C_INT) :: tt_wt_mv
    END FUNCTION
    SUBROUTINE tt_wt_free(h) BIND(C, NAME="tt_wt_free")
      IMPORT :: C_PTR
      TYPE(C_PTR), VALUE :: h
    END SUBROUTINE

    ! cTree_FillMatrices (lib_ctree.f:2) on HmatrixType parts, cTree_MxV (:99)
    FUNCTION tt_hmat_create_parts(nrow, ncol, a, lda, nparts, rptr, rnodes, cptr, cnodes, admiss, ctype, &
                                  aca_type, aca_eps, wt_kappa, h) BIND(C, NAME="tt_hmat_create_parts")
      IMPORT :: C_INT, C_INT32_T, C_INT64_T, C_DOUBLE, C_PTR
      INTEGER(C_INT32_T), VALUE :: nrow, ncol, nparts, ctype, aca_type
      REAL(C_DOUBLE), INTENT(IN) :: a(*)
      INTEGER(C_INT64_T), VALUE :: lda
      INTEGER(C_INT32_T), INTENT(IN) :: rptr(*), rnodes(*), cptr(*), cnodes(*), admiss(*)
      REAL(C_DOUBLE), VALUE :: aca_eps, wt_kappa
      TYPE(C_PTR) :: h
      INTEGER(C_INT) :: tt_hmat_create_parts
    END FUNCTION
    FUNCTION tt_hmat_mv(h, nrhs, x, y) BIND(C, NAME="tt_hmat_mv")
      IMPORT :: C_INT, C_INT32_T, C_DOUBLE, C_PTR
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT32_T), VALUE :: nrhs
      REAL(C_DOUBLE), INTENT(IN) :: x(*)
      REAL(C_DOUBLE) :: y(*)
      INTEGER(C_INT) :: tt_hmat_mv
    END FUNCTION
    SUBROUTINE tt_hmat_free(h) BIND(C, NAME="tt_hmat_free")
      IMPORT :: C_PTR
      TYPE(C_PTR), VALUE :: h
    END SUBROUTINE

    ! many independent prelsqr2 / slvlsqr2 systems in one launch
    FUNCTION tt_lsqr_batch_setup(nsys, m, n, nnz, ia, ja, a, q, scale_host, h) BIND(C, NAME="tt_lsqr_batch_setup")
      IMPORT :: C_INT, C_INT32_T, C_INT64_T, C_DOUBLE, C_PTR
      INTEGER(C_INT32_T), VALUE :: nsys, scale_host
      INTEGER(C_INT32_T), INTENT(IN) :: m(*), n(*), ia(*), ja(*)
      INTEGER(C_INT64_T), INTENT(IN) :: nnz(*)
      REAL(C_DOUBLE) :: a(*), q(*)
      TYPE(C_PTR) :: h
      INTEGER(C_INT) :: tt_lsqr_batch_setup
    END FUNCTION
    FUNCTION tt_lsqr_batch_solve(h, b, x, maxit, eps4, nit, istop, ms) BIND(C, NAME="tt_lsqr_batch_solve")
      IMPORT :: C_INT, C_INT32_T, C_DOUBLE, C_FLOAT, C_PTR
      TYPE(C_PTR), VALUE :: h
      REAL(C_DOUBLE), INTENT(IN) :: b(*)
      REAL(C_DOUBLE) :: x(*), ms
      INTEGER(C_INT32_T), VALUE :: maxit
      REAL(C_FLOAT), VALUE :: eps4
      INTEGER(C_INT32_T) :: nit(*), istop(*)
      INTEGER(C_INT) :: tt_lsqr_batch_solve
    END FUNCTION
    FUNCTION tt_lsqr_batch_set_precision(h, prec) BIND(C, NAME="tt_lsqr_batch_set_precision")
      IMPORT :: C_INT, C_INT32_T, C_PTR
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT32_T), VALUE :: prec
      INTEGER(C_INT) :: tt_lsqr_batch_set_precision
    END FUNCTION
    SUBROUTINE tt_lsqr_batch_free(h) BIND(C, NAME="tt_lsqr_batch_free")
      IMPORT :: C_PTR
      TYPE(C_PTR), VALUE :: h
    END SUBROUTINE

    ! the same systems on per-cell dense blocks (rows in groups of 51 or 27 per cell)
    FUNCTION tt_block_batch_setup(nsys, m, n, nnz, ia, ja, a, rows_per_cell, h) &
        BIND(C, NAME="tt_block_batch_setup")
      IMPORT :: C_INT, C_INT32_T, C_INT64_T, C_DOUBLE, C_PTR
      INTEGER(C_INT32_T), VALUE :: nsys, rows_per_cell
      INTEGER(C_INT32_T), INTENT(IN) :: m(*), n(*), ia(*), ja(*)
      INTEGER(C_INT64_T), INTENT(IN) :: nnz(*)
      REAL(C_DOUBLE), INTENT(IN) :: a(*)
      TYPE(C_PTR) :: h
      INTEGER(C_INT) :: tt_block_batch_setup
    END FUNCTION
    FUNCTION tt_block_batch_set_precision(h, prec) BIND(C, NAME="tt_block_batch_set_precision")
      IMPORT :: C_INT, C_INT32_T, C_PTR
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT32_T), VALUE :: prec
      INTEGER(C_INT) :: tt_block_batch_set_precision
    END FUNCTION
    FUNCTION tt_block_batch_solve(h, b, x, maxit, eps4, nit, istop, ms) BIND(C, NAME="tt_block_batch_solve")
      IMPORT :: C_INT, C_INT32_T, C_DOUBLE, C_FLOAT, C_PTR
      TYPE(C_PTR), VALUE :: h
      REAL(C_DOUBLE), INTENT(IN) :: b(*)
      REAL(C_DOUBLE) :: x(*), ms
      INTEGER(C_INT32_T), VALUE :: maxit
      REAL(C_FLOAT), VALUE :: eps4
      INTEGER(C_INT32_T) :: nit(*), istop(*)
      INTEGER(C_INT) :: tt_block_batch_solve
    END FUNCTION
    SUBROUTINE tt_block_batch_free(h) BIND(C, NAME="tt_block_batch_free")
      IMPORT :: C_PTR
      TYPE(C_PTR), VALUE :: h
    END SUBROUTINE

    ! batched solves: systems with bit-identical reduced-precision blocks read one copy
    FUNCTION tt_lsqr_set_block_sharing(on) BIND(C, NAME="tt_lsqr_set_block_sharing")
      IMPORT :: C_INT, C_INT32_T
      INTEGER(C_INT32_T), VALUE :: on
      INTEGER(C_INT) :: tt_lsqr_set_block_sharing
    END FUNCTION
    FUNCTION tt_lsqr_block_sharing_stats(launches, systems) BIND(C, NAME="tt_lsqr_block_sharing_stats")
      IMPORT :: C_INT, C_INT64_T
      INTEGER(C_INT64_T) :: launches, systems
      INTEGER(C_INT) :: tt_lsqr_block_sharing_stats
    END FUNCTION

    ! host preprocessing: ReadMeshGid3D + GenMakroMesh, ReadBic, ReadInputFile
    ! (strings NUL-terminated: TRIM(s)//C_NULL_CHAR)
    FUNCTION tt_host_read_geo(path, h) BIND(C, NAME="tt_host_read_geo")
      IMPORT :: C_INT, C_CHAR, C_PTR
      CHARACTER(KIND=C_CHAR), INTENT(IN) :: path(*)
      TYPE(C_PTR) :: h
      INTEGER(C_INT) :: tt_host_read_geo
    END FUNCTION
    FUNCTION tt_host_read_bic(h, text) BIND(C, NAME="tt_host_read_bic")
      IMPORT :: C_INT, C_CHAR, C_PTR
      TYPE(C_PTR), VALUE :: h
      CHARACTER(KIND=C_CHAR), INTENT(IN) :: text(*)
      INTEGER(C_INT) :: tt_host_read_bic
    END FUNCTION
    FUNCTION tt_host_read_inp(h, text) BIND(C, NAME="tt_host_read_inp")
      IMPORT :: C_INT, C_CHAR, C_PTR
      TYPE(C_PTR), VALUE :: h
      CHARACTER(KIND=C_CHAR), INTENT(IN) :: text(*)
      INTEGER(C_INT) :: tt_host_read_inp
    END FUNCTION
    FUNCTION tt_host_case_view(h, v) BIND(C, NAME="tt_host_case_view")
      IMPORT :: C_INT, C_PTR, tt_case_view
      TYPE(C_PTR), VALUE :: h
      TYPE(tt_case_view) :: v
      INTEGER(C_INT) :: tt_host_case_view
    END FUNCTION
    FUNCTION tt_host_fields(h, vel, vort, qvort, temp, qtemp) BIND(C, NAME="tt_host_fields")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: h
      TYPE(C_PTR) :: vel, vort, qvort, temp, qtemp
      INTEGER(C_INT) :: tt_host_fields
    END FUNCTION
    SUBROUTINE tt_host_free(h) BIND(C, NAME="tt_host_free")
      IMPORT :: C_PTR
      TYPE(C_PTR), VALUE :: h
    END SUBROUTINE
    FUNCTION tt_case_run(h, max_iter, sta, ite) BIND(C, NAME="tt_case_run")
      IMPORT :: C_INT, C_INT32_T, C_DOUBLE, C_PTR
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT32_T), VALUE :: max_iter
      REAL(C_DOUBLE) :: sta(*)
      INTEGER(C_INT32_T) :: ite(*)
      INTEGER(C_INT) :: tt_case_run
    END FUNCTION

    ! the reference's files: WrstF (lib_IO.f:604), tri.sta / tri.ite (lib_util.f:1141-1146)
    FUNCTION tt_case_write_restart(h, path, tstep, glnnlit, rtime) BIND(C, NAME="tt_case_write_restart")
      IMPORT :: C_INT, C_INT32_T, C_DOUBLE, C_CHAR, C_PTR
      TYPE(C_PTR), VALUE :: h
      CHARACTER(KIND=C_CHAR), INTENT(IN) :: path(*)
      INTEGER(C_INT32_T), VALUE :: tstep, glnnlit
      REAL(C_DOUBLE), VALUE :: rtime
      INTEGER(C_INT) :: tt_case_write_restart
    END FUNCTION
    FUNCTION tt_write_sta_ite(sta_path, ite_path, id_line, nrows, tstep, tsnnlit, err, nit) &
        BIND(C, NAME="tt_write_sta_ite")
      IMPORT :: C_INT, C_INT32_T, C_DOUBLE, C_CHAR, C_PTR
      CHARACTER(KIND=C_CHAR), INTENT(IN) :: sta_path(*), ite_path(*)
      TYPE(C_PTR), VALUE :: id_line
      INTEGER(C_INT32_T), VALUE :: nrows
      INTEGER(C_INT32_T), INTENT(IN) :: tstep(*), tsnnlit(*), nit(*)
      REAL(C_DOUBLE), INTENT(IN) :: err(*)
      INTEGER(C_INT) :: tt_write_sta_ite
    END FUNCTION

    ! single-domain modified Helmholtz (EQNS imH 3-7): corMin_SDkin_mHfs
    ! (lib_fm.f:1646), fmmHSysM + FormPREfm (:1061), the compressions of
    ! main.f90:609-777 and fmSolveModHelmDC / DCwawt / DCaca / DChmat
    FUNCTION tt_host_mh_codes(h, kode, qkode, u, q, ibcf, iwn, bewn, nunk, nknown) BIND(C, NAME="tt_host_mh_codes")
      IMPORT :: C_INT, C_INT32_T, C_PTR
      TYPE(C_PTR), VALUE :: h
      TYPE(C_PTR) :: kode, qkode, u, q, ibcf, iwn, bewn
      INTEGER(C_INT32_T) :: nunk, nknown
      INTEGER(C_INT) :: tt_host_mh_codes
    END FUNCTION
    FUNCTION tt_mhsd_create(m, q, ibcf, diff, qdiff, beta, h, ont, mumax) BIND(C, NAME="tt_mhsd_create")
      IMPORT :: C_INT, C_INT32_T, C_DOUBLE, C_PTR, tt_mesh_view, tt_quad
      TYPE(tt_mesh_view), INTENT(IN) :: m
      TYPE(tt_quad), INTENT(IN) :: q
      INTEGER(C_INT32_T), INTENT(IN) :: ibcf(*)
      REAL(C_DOUBLE), INTENT(IN) :: diff(*), qdiff(*)
      REAL(C_DOUBLE), VALUE :: beta
      TYPE(C_PTR) :: h
      REAL(C_DOUBLE) :: ont(5), mumax
      INTEGER(C_INT) :: tt_mhsd_create
    END FUNCTION
    FUNCTION tt_mhsd_matrix(h, set, k, dst) BIND(C, NAME="tt_mhsd_matrix")
      IMPORT :: C_INT, C_INT32_T, C_DOUBLE, C_PTR
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT32_T), VALUE :: set, k
      REAL(C_DOUBLE) :: dst(*)
      INTEGER(C_INT) :: tt_mhsd_matrix
    END FUNCTION
    FUNCTION tt_mhsd_system(h, kode, qkode, nunk, nknown) BIND(C, NAME="tt_mhsd_system")
      IMPORT :: C_INT, C_INT32_T, C_PTR
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT32_T), INTENT(IN) :: kode(*), qkode(*)
      INTEGER(C_INT32_T), VALUE :: nunk, nknown
      INTEGER(C_INT) :: tt_mhsd_system
    END FUNCTION
    FUNCTION tt_mhsd_compress(h, mode, aca_type, aca_eps, wt_kappa, ct_type, ct_eta, leaf, ratio) &
        BIND(C, NAME="tt_mhsd_compress")
      IMPORT :: C_INT, C_INT32_T, C_DOUBLE, C_PTR
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT32_T), VALUE :: mode, aca_type, ct_type, leaf
      REAL(C_DOUBLE), VALUE :: aca_eps, wt_kappa, ct_eta
      REAL(C_DOUBLE) :: ratio
      INTEGER(C_INT) :: tt_mhsd_compress
    END FUNCTION
    FUNCTION tt_mhsd_solve(h, u, qu, velocity, rhsv, dgrad, urmh) BIND(C, NAME="tt_mhsd_solve")
      IMPORT :: C_INT, C_DOUBLE, C_PTR
      TYPE(C_PTR), VALUE :: h
      REAL(C_DOUBLE) :: u(*), qu(*)
      REAL(C_DOUBLE), INTENT(IN) :: velocity(*), rhsv(*), dgrad(*)
      REAL(C_DOUBLE), VALUE :: urmh
      INTEGER(C_INT) :: tt_mhsd_solve
    END FUNCTION
    SUBROUTINE tt_mhsd_free(h) BIND(C, NAME="tt_mhsd_free")
      IMPORT :: C_PTR
      TYPE(C_PTR), VALUE :: h
    END SUBROUTINE
  END INTERFACE

CONTAINS

  ! fmSolveModHelmDC (and its wavelet / ACA / block-matrix forms, by the
  ! operator's compression mode) with the arrays of the reference's call
  ! (main.f90:1349-1352); the matrices live behind h (tt_mhsd_create,
  ! tt_mhsd_system, tt_mhsd_compress)
  SUBROUTINE fmSolveModHelmDC_gpu(h, nnodes, nbnpof, u, qu, velocity, rhsv, diffGrad, urmH)
    TYPE(C_PTR) :: h
    INTEGER :: nnodes, nbnpof
    REAL(8) :: u(nnodes), qu(nbnpof), velocity(nnodes, 3), rhsv(nnodes), diffGrad(nnodes, 3), urmH
    IF (tt_mhsd_solve(h, u, qu, velocity, rhsv, diffGrad, urmH) < 0) CALL tt_fail('fmSolveModHelmDC_gpu')
  END SUBROUTINE

  ! prelsqr2 with the reference's arguments; h receives the device handle
  ! that slvlsqr2_gpu takes (a is scaled in place, as prelsqr2 does)
  SUBROUTINE prelsqr2_gpu(m, n, nnz, q, ia, ja, a, h)
    INTEGER :: m, n, nnz, ia(m + 1), ja(nnz)
    REAL(8) :: q(n), a(nnz)
    TYPE(C_PTR) :: h
    IF (tt_lsqr_setup(m, n, nnz, ia, ja, a, q, 1, h) < 0) CALL tt_fail('prelsqr2_gpu')
  END SUBROUTINE

  ! slvlsqr2 with the reference's arguments (ierr = istop when istop > 3)
  SUBROUTINE slvlsqr2_gpu(h, m, n, nnz, maxit, eps4, nit, ierr, q, ia, ja, a, b, x)
    TYPE(C_PTR) :: h
    INTEGER :: m, n, nnz, maxit, nit, ierr, ia(m + 1), ja(nnz)
    REAL(4) :: eps4
    REAL(8) :: q(n), a(nnz), b(m), x(n)
    INTEGER(C_INT32_T) :: istop, nit4
    IF (tt_lsqr_solve(h, b, x, maxit, eps4, nit4, istop) < 0) CALL tt_fail('slvlsqr2_gpu')
    nit = nit4
    ierr = 0
    IF (istop > 3) ierr = istop
  END SUBROUTINE

  SUBROUTINE tt_fail(where)
    CHARACTER(*) :: where
    CHARACTER(KIND=C_CHAR), POINTER :: msg(:)
    INTEGER :: k
    CALL C_F_POINTER(tt_last_error(), msg, [512])
    k = 0
    DO WHILE (k < 512)
      IF (msg(k + 1) == C_NULL_CHAR) EXIT
      k = k + 1
    END DO
    WRITE (*, '(A,": ",512A)') where, msg(1:k)
    STOP 4
  END SUBROUTINE
END MODULE tritok_amd_bind
