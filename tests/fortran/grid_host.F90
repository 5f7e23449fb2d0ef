!-------------------------------------------------------------------------
! Test harness: a host-side module `grid` with the names, index ranges and
! accessor semantics that the reference's grid module gives its solver
! (grid.F90:5-194): GidList / GidListMax per level, Igrid / Jgrid / Kgrid,
! ParentGid, ChildGid, NeighborGid (same-level neighbours, Undefi
! elsewhere), CellWidth, BlockMem(gid)%u(-2:9,-2:9,-2:9,0:NM-1) and the
! component pointers get_Ucomp / get_U1comp (lower bounds -2, as slice3
! returns them), and the base-block tables GidBase / RankBase (grid.F90:54).
! Array sizes are set at run time by grid_host_alloc / grid_host_base.  It
! exists so that the drop-in module `fmg` can be compiled and driven as the
! reference's time step would drive it; it is not part of the library.
!-------------------------------------------------------------------------
module grid
  implicit none
  integer,parameter :: Undefi = -1
  integer,parameter :: Left = 0, Right = 1
  integer,parameter :: Imin = 0, Jmin = 0, Kmin = 0, Imax = 7, Jmax = 7, Kmax = 7, Gidmin = 0, Lmin = 0
  integer,parameter :: Ngh = 2
  type t_blockMem
     real(kind=8),dimension(:,:,:,:),pointer :: u => null()
  end type t_blockMem
  type(t_blockMem),save,allocatable :: BlockMem(:), BlockMem1(:)
  integer,save,allocatable :: Igrid(:), Jgrid(:), Kgrid(:)
  integer,save,allocatable :: GidList(:,:), GidListMax(:)
  integer,save,allocatable :: ParentGid(:,:), ChildGid(:,:,:,:,:), NeighborGid(:,:,:,:)
  real(kind=8),save,allocatable :: CellWidth(:,:)
  integer,save,allocatable :: GidBase(:,:,:), RankBase(:,:,:)
  integer,save :: LevelMax = Undefi
contains
  subroutine grid_host_alloc(ngid, nl, nm)
    integer,intent(IN) :: ngid, nl, nm
    integer :: g
    allocate(BlockMem(0:ngid-1), BlockMem1(0:ngid-1))
    do g = 0, ngid - 1
       allocate(BlockMem(g)%u(Imin-Ngh:Imax+Ngh, Jmin-Ngh:Jmax+Ngh, Kmin-Ngh:Kmax+Ngh, 0:nm-1))
       allocate(BlockMem1(g)%u(Imin-Ngh:Imax+Ngh, Jmin-Ngh:Jmax+Ngh, Kmin-Ngh:Kmax+Ngh, 0:nm-1))
    end do
    allocate(Igrid(0:ngid-1), Jgrid(0:ngid-1), Kgrid(0:ngid-1))
    allocate(GidList(0:ngid-1, 0:nl-1), GidListMax(0:nl-1))
    allocate(ParentGid(0:ngid-1, 0:0), ChildGid(0:1, 0:1, 0:1, 0:ngid-1, 0:0), NeighborGid(0:1, 0:2, 0:ngid-1, 0:0))
    allocate(CellWidth(0:2, 0:nl-1))
    Igrid = Undefi; Jgrid = Undefi; Kgrid = Undefi
    GidList = Undefi; GidListMax = Undefi
    ParentGid = Undefi; ChildGid = Undefi; NeighborGid = Undefi
  end subroutine grid_host_alloc

  ! base blocks nx x ny x nz, numbered x fastest, all on rank 0
  subroutine grid_host_base(nx, ny, nz)
    integer,intent(IN) :: nx, ny, nz
    integer :: i, j, k
    if (allocated(GidBase)) deallocate(GidBase, RankBase)
    allocate(GidBase(0:nx-1, 0:ny-1, 0:nz-1), RankBase(0:nx-1, 0:ny-1, 0:nz-1))
    do k = 0, nz - 1
       do j = 0, ny - 1
          do i = 0, nx - 1
             GidBase(i, j, k) = (k * ny + j) * nx + i
          end do
       end do
    end do
    RankBase = 0
  end subroutine grid_host_base

  function get_Ucomp(ncomp, gid) result(ptr)
    integer,intent(IN) :: ncomp, gid
    real(kind=8),dimension(:,:,:),pointer :: ptr
    ptr(Imin-Ngh:, Jmin-Ngh:, Kmin-Ngh:) => BlockMem(gid)%u(:, :, :, ncomp)
  end function get_Ucomp

  function get_U1comp(ncomp, gid) result(ptr)
    integer,intent(IN) :: ncomp, gid
    real(kind=8),dimension(:,:,:),pointer :: ptr
    ptr(Imin-Ngh:, Jmin-Ngh:, Kmin-Ngh:) => BlockMem1(gid)%u(:, :, :, ncomp)
  end function get_U1comp
end module grid
