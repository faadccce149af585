! tt_fortran_driver.f90 -- a Fortran host calling libtritok_amd.so through
! tritok_amd_bind, the way the reference's call sites would:
!
!   tt_fortran_driver lsqr <in> <out>
!       prelsqr2_gpu + slvlsqr2_gpu (the reference's argument lists,
!       lib_Solve_v11.f:4401, :4501) on the CSR system of the stream file <in>:
!       m, n, nnz, maxit (int32), eps (float32), ia(m+1), ja(nnz) (int32),
!       a(nnz), b(m), x0(n) (float64); writes q(n), the scaled a(nnz), x(n)
!       (float64) and nit, ierr (int32) to <out>.
!   tt_fortran_driver case <geo> <bic> <inp> <dir>
!       the steady Picard loop of main.f90 on the device: mesh, boundary
!       conditions and tri.inp read by the host routines, the case built and
!       run to convergence, then tri.rst (WrstF), tri.sta and tri.ite written
!       into <dir>.
PROGRAM tt_fortran_driver
  USE tritok_amd_bind
  IMPLICIT NONE
  CHARACTER(LEN=1024) :: mode
  CALL GET_COMMAND_ARGUMENT(1, mode)
  SELECT CASE (TRIM(mode))
  CASE ('lsqr')
    CALL run_lsqr()
  CASE ('case')
    CALL run_case()
  CASE DEFAULT
    PRINT *, 'usage: tt_fortran_driver lsqr <in> <out> | case <geo> <bic> <inp> <dir>'
    STOP 2
  END SELECT

CONTAINS

  SUBROUTINE run_lsqr()
    CHARACTER(LEN=1024) :: fin, fout
    INTEGER :: m, n, nnz, maxit, nit, ierr
    REAL(4) :: eps4
    INTEGER, ALLOCATABLE :: ia(:), ja(:)
    REAL(8), ALLOCATABLE :: a(:), b(:), x(:), q(:)
    TYPE(C_PTR) :: h
    CALL GET_COMMAND_ARGUMENT(2, fin)
    CALL GET_COMMAND_ARGUMENT(3, fout)
    OPEN (10, FILE=TRIM(fin), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD')
    READ (10) m, n, nnz, maxit, eps4
    ALLOCATE (ia(m + 1), ja(nnz), a(nnz), b(m), x(n), q(n))
    READ (10) ia, ja, a, b, x
    CLOSE (10)
    CALL prelsqr2_gpu(m, n, nnz, q, ia, ja, a, h)
    CALL slvlsqr2_gpu(h, m, n, nnz, maxit, eps4, nit, ierr, q, ia, ja, a, b, x)
    CALL tt_lsqr_free(h)
    OPEN (11, FILE=TRIM(fout), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
    WRITE (11) q, a, x, nit, ierr
    CLOSE (11)
  END SUBROUTINE

  ! the contents of a text file as one NUL-terminated string
  FUNCTION slurp(path) RESULT(text)
    CHARACTER(*), INTENT(IN) :: path
    CHARACTER(KIND=C_CHAR, LEN=:), ALLOCATABLE :: text
    INTEGER :: sz
    OPEN (12, FILE=TRIM(path), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD')
    INQUIRE (UNIT=12, SIZE=sz)
    ALLOCATE (CHARACTER(KIND=C_CHAR, LEN=sz + 1) :: text)
    READ (12) text(1:sz)
    CLOSE (12)
    text(sz + 1:sz + 1) = C_NULL_CHAR
  END FUNCTION

  SUBROUTINE run_case()
    CHARACTER(LEN=1024) :: geo, bic, inp, dir
    TYPE(C_PTR) :: hh, hc, pvel, pvort, pqvort, ptemp, pqtemp
    TYPE(tt_case_view) :: v
    REAL(8), POINTER :: vel(:), vort(:), qvort(:), temp(:), qtemp(:)
    REAL(8), ALLOCATABLE :: sta(:, :)
    INTEGER(C_INT32_T), ALLOCATABLE :: ite(:, :), ts(:), tsn(:)
    INTEGER :: its, k, nn, nq
    CALL GET_COMMAND_ARGUMENT(2, geo)
    CALL GET_COMMAND_ARGUMENT(3, bic)
    CALL GET_COMMAND_ARGUMENT(4, inp)
    CALL GET_COMMAND_ARGUMENT(5, dir)
    IF (tt_host_read_geo(TRIM(geo)//C_NULL_CHAR, hh) < 0) CALL tt_fail('tt_host_read_geo')
    IF (tt_host_read_bic(hh, slurp(bic)) < 0) CALL tt_fail('tt_host_read_bic')
    IF (tt_host_read_inp(hh, slurp(inp)) < 0) CALL tt_fail('tt_host_read_inp')
    IF (tt_host_case_view(hh, v) < 0) CALL tt_fail('tt_host_case_view')
    IF (tt_host_fields(hh, pvel, pvort, pqvort, ptemp, pqtemp) < 0) CALL tt_fail('tt_host_fields')
    nn = v%mesh%nnodes
    nq = v%mesh%nq
    CALL C_F_POINTER(pvel, vel, [3 * nn])
    CALL C_F_POINTER(pvort, vort, [3 * nn])
    CALL C_F_POINTER(pqvort, qvort, [3 * nq])
    CALL C_F_POINTER(ptemp, temp, [nn])
    CALL C_F_POINTER(pqtemp, qtemp, [nq])
    IF (tt_case_create(v, vel, vort, qvort, temp, qtemp, TT_PREC_FP64, 0.0D0, hc) < 0) CALL tt_fail('tt_case_create')
    ALLOCATE (sta(9, v%nnlit), ite(12, v%nnlit))
    its = tt_case_run(hc, v%nnlit, sta, ite)
    IF (its < 0) CALL tt_fail('tt_case_run')
    IF (tt_case_write_restart(hc, TRIM(dir)//'/tri.rst'//C_NULL_CHAR, 1, its, v%tstep) < 0) &
      CALL tt_fail('tt_case_write_restart')
    ALLOCATE (ts(its), tsn(its))
    ts = 1
    tsn = [(k, k = 1, its)]
    IF (tt_write_sta_ite(TRIM(dir)//'/tri.sta'//C_NULL_CHAR, TRIM(dir)//'/tri.ite'//C_NULL_CHAR, C_NULL_PTR, its, &
                         ts, tsn, sta, ite) < 0) CALL tt_fail('tt_write_sta_ite')
    CALL tt_case_free(hc)
    CALL tt_host_free(hh)
    PRINT '(A,I6,A)', 'converged in ', its, ' Picard iterations'
  END SUBROUTINE
END PROGRAM tt_fortran_driver
