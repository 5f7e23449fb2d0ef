!-------------------------------------------------------------------------
! Driver of the drop-in module `fmg`: builds a hierarchy in the harness grid
! module the way the reference's grid holds it, puts rho in the U1 slot and
! the initial psi in U (with junk in the hydro ghost layers, which the solve
! must not read), calls fmg_poisson, and writes psi and g of every block.
! Usage: fmg_drop_in_test IN.bin OUT.bin TOL MAXIT STOP_MODE PRECISION
!   IN.bin (stream): int32 nblk, nlev, bc; float64 h0, pi4g, x0(3);
!     int32 level(nblk), ijk(3,nblk), parent(nblk), child(8,nblk),
!     nbr_same(6,nblk) (same-level neighbours, -1 otherwise);
!     float64 rho(512,nblk), psi0(512,nblk)      (blocks 0-based, x fastest)
!   OUT.bin: float64 psi(512,nblk), g(512,3,nblk)
! Grid ids are assigned in reverse block order, so the module's gid <-> block
! mapping is exercised.
!-------------------------------------------------------------------------
program fmg_drop_in_test
  use grid
  use fmg
  implicit none
  integer,parameter :: MRHO = 0, MPSI = 5, MGX = 6, NM = 9
  integer :: nblk, nlev, bc, u_in, u_out, b, gid, l, n, o, d, lr, maxit, stop_mode, prec
  real(kind=8) :: h0, pi4g, x0(3), tol
  integer,allocatable :: level(:), ijk(:,:), parent(:), child(:,:), nbr(:,:)
  real(kind=8),allocatable :: rho(:,:), psi(:,:)
  real(kind=8),pointer :: u(:,:,:)
  character(len=512) :: fin, fout, arg
  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  call get_command_argument(3, arg); read(arg, *) tol
  call get_command_argument(4, arg); read(arg, *) maxit
  call get_command_argument(5, arg); read(arg, *) stop_mode
  call get_command_argument(6, arg); read(arg, *) prec
  open(newunit=u_in, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u_in) nblk, nlev, bc
  read(u_in) h0, pi4g, x0
  allocate(level(0:nblk-1), ijk(3, 0:nblk-1), parent(0:nblk-1), child(8, 0:nblk-1), nbr(6, 0:nblk-1))
  allocate(rho(512, 0:nblk-1), psi(512, 0:nblk-1))
  read(u_in) level, ijk, parent, child, nbr, rho, psi
  close(u_in)
  call grid_host_alloc(nblk, nlev, NM)
  LevelMax = nlev - 1
  do l = 0, nlev - 1
     CellWidth(:, l) = h0 / 2.d0**l
     GidListMax(l) = Gidmin - 1
  end do
  do b = 0, nblk - 1
     gid = nblk - 1 - b
     l = level(b)
     GidListMax(l) = GidListMax(l) + 1
     GidList(GidListMax(l), l) = gid
     Igrid(gid) = ijk(1, b); Jgrid(gid) = ijk(2, b); Kgrid(gid) = ijk(3, b)
     if (parent(b) >= 0) ParentGid(gid, 0) = nblk - 1 - parent(b)
     do o = 0, 7
        if (child(o+1, b) >= 0) ChildGid(iand(o, 1), iand(ishft(o, -1), 1), ishft(o, -2), gid, 0) = nblk - 1 - child(o+1, b)
     end do
     do d = 0, 2
        do lr = 0, 1
           if (nbr(2*d+lr+1, b) >= 0) NeighborGid(lr, d, gid, 0) = nblk - 1 - nbr(2*d+lr+1, b)
        end do
     end do
     BlockMem(gid)%u = 1.d300
     BlockMem1(gid)%u = -1.d300
     u => get_U1comp(MRHO, gid)
     u(0:7, 0:7, 0:7) = reshape(rho(:, b), [8, 8, 8])
     u => get_Ucomp(MPSI, gid)
     u(0:7, 0:7, 0:7) = reshape(psi(:, b), [8, 8, 8])
  end do
  ! GidList order inside a level: ascending block number (the ids run backwards)
  call fmg_set_problem(pi4g, bc, x0, MRHO, MPSI, MGX, 1)
  call fmg_set_solver(tol, maxit, stop_mode, prec)
  call fmg_poisson
  open(newunit=u_out, file=trim(fout), access='stream', form='unformatted', status='replace')
  do b = 0, nblk - 1
     u => get_Ucomp(MPSI, nblk - 1 - b)
     write(u_out) u(0:7, 0:7, 0:7)
  end do
  do b = 0, nblk - 1
     do d = 0, 2
        u => get_Ucomp(MGX + d, nblk - 1 - b)
        write(u_out) u(0:7, 0:7, 0:7)
     end do
  end do
  close(u_out)
  print '(a,i0,a,es12.4)', 'fmg_drop_in_test: iterations ', fmg_last_iterations(), ' residual ', fmg_last_residual()
  call fmg_finalize
end program fmg_drop_in_test
