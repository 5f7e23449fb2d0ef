!-------------------------------------------------------------------------
! Driver in the style of the reference's mg_test (fmg_test.F90:8-77): fills
! the source of the drop-in module `mg`, runs mg_cycle and writes the solution.
! Usage: mg_drop_in_test N PRECISION IN.bin OUT.bin
!   IN.bin: N^3 float64 source, x fastest;  OUT.bin: N^3 float64 solution.
!-------------------------------------------------------------------------
program mg_drop_in_test
  use mg
  use sg_capi, only : SG_STOP_TRUE_RESIDUAL
  implicit none
  integer :: n, prec, u_in, u_out
  character(len=512) :: arg, fin, fout
  real(kind=8),pointer,dimension(:,:,:,:) :: f, u
  real(kind=8),allocatable :: dense(:,:,:)
  call get_command_argument(1, arg); read(arg, *) n
  call get_command_argument(2, arg); read(arg, *) prec
  call get_command_argument(3, fin)
  call get_command_argument(4, fout)
  allocate(dense(0:n-1, 0:n-1, 0:n-1))
  open(newunit=u_in, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u_in) dense
  close(u_in)
  call mg_set_domain(n, 1.d0 / n)
  call mg_set_solver(1.d-10, 30, SG_STOP_TRUE_RESIDUAL, prec)
  call mg_init(0)
  f => mg_get_arrp(0, IRHO)
  if (lbound(f, 1) /= -1 .or. ubound(f, 1) /= n) then
     print *, '*** mg_get_arrp: unexpected bounds', lbound(f), ubound(f)
     stop 1
  end if
  f(0:n-1, 0:n-1, 0:n-1, 0) = dense
  call mg_fill0(0, IU)
  call mg_cycle(IU, IRHO)
  u => mg_get_arrp(0, IU)
  dense = u(0:n-1, 0:n-1, 0:n-1, 0)
  open(newunit=u_out, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u_out) dense
  close(u_out)
  print '(a,i0,a,es12.4)', 'mg_drop_in_test: iterations ', mg_last_iterations(), ' residual ', mg_last_residual()
  call mg_finalize
end program mg_drop_in_test
