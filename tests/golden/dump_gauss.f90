! Writes the reference's Gauss-Legendre table, as its own GaussPosWeights
! (src/lib_int.f:2673-3173) fills TYPE(GaussType), to a stream file:
! ng1(9), ng2(9) as int32, then gi(220), ome(220) as float64.
! Linked against the objects oracle/Makefile.ref builds from the reference.
program dump_gauss
  use inc_types
  implicit none
  type(gausstype) :: g
  call GaussPosWeights(g)
  open(10, file='gauss_ref.bin', access='stream', form='unformatted', status='replace')
  write(10) g%ng1(1:9), g%ng2(1:9), g%gi(1:220), g%ome(1:220)
  close(10)
end program dump_gauss
