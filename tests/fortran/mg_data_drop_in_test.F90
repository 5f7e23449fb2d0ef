!-------------------------------------------------------------------------
! Driver of the drop-in modules `mg_data` / `mg` through the call sequences
! of the reference's own callers, with the FMG-side arrays held by the test
! harness's fmg_data and grid modules:
!  1. mg_test (fmg_test.F90:38-55): mg_init(fmglev); mg_dataPrepare((/IRHO, IU/),
!     (/IRHO, IU/)); mg_fill0(0, IU); mg_cycle(IU, IRHO); mg_dataRestore((/IU/),
!     (/IU/)); the solution read back from the FMG blocks' IU;
!  2. vmg_slvsml (vmg.F90:407-428): mg_dataPrepare((/ISRC, IPSI/), (/IRHS, IRES/));
!     mg_cycle(IPSI, ISRC); mg_dataRestore((/IRES/), (/IPSI/)), starting from the
!     initial guess the FMG blocks' IRES hold.
! Usage: mg_data_drop_in_test NB FMGLEV PRECISION IN.bin GUESS.bin OUT1.bin OUT2.bin
!   NB base blocks per side; the MG grid has n = NB*8/2^FMGLEV cells per side;
!   IN.bin, GUESS.bin: n^3 float64 (x fastest); OUT*.bin: n^3 float64.
!-------------------------------------------------------------------------
program mg_data_drop_in_test
  use grid, only : grid_host_alloc, grid_host_base, CellWidth, GidBase
  use fmg_data, only : fmg_data_host_alloc, fmg_get_arrp, AMR_LevelMin
  use mg_data
  use mg
  use sg_capi, only : SG_STOP_TRUE_RESIDUAL
  implicit none
  integer :: nb, fmglev, prec, n, bs, i, j, k, u_in
  character(len=512) :: arg, fin, fguess, fout1, fout2
  real(kind=8),allocatable :: src(:,:,:), guess(:,:,:), out(:,:,:)
  real(kind=8),pointer,dimension(:,:,:,:) :: uf
  call get_command_argument(1, arg); read(arg, *) nb
  call get_command_argument(2, arg); read(arg, *) fmglev
  call get_command_argument(3, arg); read(arg, *) prec
  call get_command_argument(4, fin)
  call get_command_argument(5, fguess)
  call get_command_argument(6, fout1)
  call get_command_argument(7, fout2)
  bs = 8 / 2**fmglev
  n = nb * bs
  allocate(src(0:n-1, 0:n-1, 0:n-1), guess(0:n-1, 0:n-1, 0:n-1), out(0:n-1, 0:n-1, 0:n-1))
  open(newunit=u_in, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u_in) src
  close(u_in)
  open(newunit=u_in, file=trim(fguess), access='stream', form='unformatted', status='old')
  read(u_in) guess
  close(u_in)
  ! host side: base blocks of the grid module and their FMG arrays (junk in the ghost layers)
  call grid_host_alloc(nb**3, 1, 1)
  call grid_host_base(nb, nb, nb)
  CellWidth(:, 0) = 1.d0 / (nb * 8)
  call fmg_data_host_alloc(nb**3, fmglev)
  do k = 0, nb - 1
     do j = 0, nb - 1
        do i = 0, nb - 1
           uf => fmg_get_arrp(AMR_LevelMin, fmglev, GidBase(i, j, k), IRHO)
           uf = 7.d0
           uf(0:bs-1, 0:bs-1, 0:bs-1, 0) = src(bs*i:bs*i+bs-1, bs*j:bs*j+bs-1, bs*k:bs*k+bs-1)
           uf => fmg_get_arrp(AMR_LevelMin, fmglev, GidBase(i, j, k), IU)
           uf = -3.d0
           uf => fmg_get_arrp(AMR_LevelMin, fmglev, GidBase(i, j, k), IRHS)
           uf = 5.d0
           uf(0:bs-1, 0:bs-1, 0:bs-1, 0) = src(bs*i:bs*i+bs-1, bs*j:bs*j+bs-1, bs*k:bs*k+bs-1)
           uf => fmg_get_arrp(AMR_LevelMin, fmglev, GidBase(i, j, k), IRES)
           uf = 11.d0
           uf(0:bs-1, 0:bs-1, 0:bs-1, 0) = guess(bs*i:bs*i+bs-1, bs*j:bs*j+bs-1, bs*k:bs*k+bs-1)
        end do
     end do
  end do
  call mg_set_solver(1.d-10, 30, SG_STOP_TRUE_RESIDUAL, prec)
  ! 1. mg_test
  call mg_init(fmglev)
  call mg_dataPrepare((/IRHO, IU/), (/IRHO, IU/))
  call mg_fill0(0, IU)
  call mg_cycle(IU, IRHO)
  call mg_dataRestore((/IU/), (/IU/))
  call gather(IU, fout1)
  print '(a,i0,a,es12.4)', 'mg_test: iterations ', mg_last_iterations(), ' residual ', mg_last_residual()
  ! 2. vmg_slvsml with jrhs = IRHS, ju = IRES
  call mg_dataPrepare((/ISRC, IPSI/), (/IRHS, IRES/))
  call mg_cycle(IPSI, ISRC)
  call mg_dataRestore((/IRES/), (/IPSI/))
  call gather(IRES, fout2)
  print '(a,i0,a,es12.4)', 'vmg_slvsml: iterations ', mg_last_iterations(), ' residual ', mg_last_residual()
  ! the ghost layers of the FMG blocks are not touched by the exchange
  uf => fmg_get_arrp(AMR_LevelMin, fmglev, GidBase(0, 0, 0), IRES)
  if (uf(-1, 0, 0, 0) /= 11.d0 .or. uf(bs, bs, bs, 0) /= 11.d0) then
     print *, '*** ghost layer changed'
     stop 1
  end if
  call mg_finalize
contains
  subroutine gather(code, path)
    integer,intent(IN) :: code
    character(len=*),intent(IN) :: path
    integer :: ii, jj, kk, u_out
    do kk = 0, nb - 1
       do jj = 0, nb - 1
          do ii = 0, nb - 1
             uf => fmg_get_arrp(AMR_LevelMin, fmglev, GidBase(ii, jj, kk), code)
             out(bs*ii:bs*ii+bs-1, bs*jj:bs*jj+bs-1, bs*kk:bs*kk+bs-1) = uf(0:bs-1, 0:bs-1, 0:bs-1, 0)
          end do
       end do
    end do
    open(newunit=u_out, file=trim(path), access='stream', form='unformatted', status='replace')
    write(u_out) out
    close(u_out)
  end subroutine gather
end program mg_data_drop_in_test
