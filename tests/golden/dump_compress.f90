! Runs the reference's compressed-matrix routines on the matrices of
! compress_in.bin and writes what they produce to compress_out.bin (stream
! files, native byte order).  Linked against the objects oracle/Makefile.ref
! builds from the reference's own sources:
!   kind 1  aca_CrossAprox      (src/lib_ACA.f:126-181)
!   kind 2  ACA_AB              (src/lib_ACA.f:372-400, EPS_ACA_NORM :438-474)
!   kind 3  aca_RankCrossAprox  (src/lib_ACA.f:24-125)
!   kind 4  Che_WcomMatrix      (src/lib_WT.f:148-210)
! The Fortran RNG is seeded per case with RANDOM_SEED(PUT=...), and the vector
! the routine draws is drawn once more from the same seed and written out, so
! the random-vector tests can be replayed.
!
! compress_in.bin : ncase, then per case kind, m, n (int32), eps (float64),
!                   seed (int32), a(m,n) (float64, column-major)
! compress_out.bin: per case
!   kinds 1-3: rank (int32), norm (float64), vec(n), a(m,rank), b(rank,n),
!              the residual left in the input matrix (m,n)
!   kind 4   : nnz, nrow, stored (int32), zvr(m+1), c(stored) (int32),
!              v(stored), norm, vec(n) (float64)
program dump_compress
  use inc_types
  implicit none
  integer :: ncase, ic, kind, m, n, seed, nseed, i, stored
  real(8) :: eps, norm, mb, comp, spar
  real(8), allocatable :: a(:,:), vec(:)
  integer, allocatable :: sarr(:)
  type(acaMatrixType) :: aca
  type(crsMatrixType) :: crs

  call random_seed(size=nseed)
  allocate(sarr(nseed))
  open(10, file='compress_in.bin', access='stream', form='unformatted', status='old')
  open(11, file='compress_out.bin', access='stream', form='unformatted', status='replace')
  read(10) ncase
  do ic = 1, ncase
    read(10) kind, m, n, eps, seed
    allocate(a(m,n), vec(n))
    read(10) a
    do i = 1, nseed
      sarr(i) = seed + 7919*i
    end do
    call random_seed(put=sarr)
    do i = 1, n
      call random_number(vec(i))
    end do
    call random_seed(put=sarr)
    norm = -1.0d0
    select case (kind)
    case (1)
      call aca_CrossAprox(a, m, n, aca, eps, norm)
    case (2)
      call ACA_AB(aca, a, m, n, eps)
      vec = 0.0d0
    case (3)
      call aca_RankCrossAprox(a, m, n, aca, eps, norm)
    case (4)
      call Che_WcomMatrix(a, m, n, crs, eps, mb, comp, spar, norm)
    end select
    if (kind <= 3) then
      write(11) aca%rank, norm, vec(1:n), aca%a(1:m,1:aca%rank), aca%b(1:aca%rank,1:n), a
      deallocate(aca%a, aca%b)
    else
      stored = crs%zvr(m+1) - 1
      write(11) crs%nnz, crs%nrow, stored, crs%zvr(1:m+1), crs%c(1:stored), crs%v(1:stored), norm, vec(1:n)
      deallocate(crs%zvr, crs%c, crs%v)
    end if
    deallocate(a, vec)
  end do
  close(10)
  close(11)
end program dump_compress
