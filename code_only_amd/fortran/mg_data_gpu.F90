!-------------------------------------------------------------------------
! Drop-in for the reference's module `mg_data` (mg_data.F90:6-571) as the
! uniform solver's data exchange uses it: the level-0 arrays of every code in
! the ghosted layout u(-1:n,-1:n,-1:n,0:0) (mg_get_arrp, mg_alloc_arr), the FMG
! level they belong to (FMG_Level, mg_data_init / mg_data_finalize), and the
! gather of the base blocks' FMG arrays into them and the scatter back
! (mg_dataPrepare / mg_dataRestore, mg_data.F90:179-330), so that
! vmg_slvsml (vmg.F90:407-428) and mg_test (fmg_test.F90:38-55) link against
! it unchanged.  Coarser MG levels live on the GPU inside libsg's context
! (module mg), so only level MG_LevelMin is held here.
!
! One rank: the reference's mpi_gather(v) / mpi_scatter(v) to PRIMARY_RANK
! become local copies; every base block must be on rank 0 (RankBase == 0).
! The domain comes from the host's module grid (GidBase extent x 8 cells,
! CellWidth at level Lmin) unless mg_set_domain was called first.
!-------------------------------------------------------------------------
module mg_data
  use fmg_data, only : fmg_get_arrp, AMR_LevelMin, IDBG, IU, IRES, IRHS, IRHO, IFLX, ISRC, IPSI, IRUF, &
       ICODEMIN, ICODEMAX
  implicit none
  private
  public :: IDBG, IU, IRES, IRHS, IRHO, IFLX, ISRC, IPSI, IRUF
  integer,save,public :: FMG_Level = -1       ! mg_data.F90:52-54
  integer,save,public :: MG_LevelMin = 0, MG_LevelMax = -1
  integer,save,public :: Nmg = 0              ! cells per side at MG_LevelMin
  real(kind=8),save,public :: Hmg = 0.d0      ! cell width at MG_LevelMin
  integer,save :: NBase = 0                   ! cells per side of the base grid at FMG level 0
  real(kind=8),save :: H0 = 0.d0
  integer,save :: Bs = 0                      ! block size at FMG_Level (fmg_get_gridsize)
  type t_mgarr
     real(kind=8),dimension(:,:,:,:),pointer :: u => null()
  end type t_mgarr
  type(t_mgarr),save :: Arr(ICODEMIN:ICODEMAX)
  public :: mg_data_init, mg_data_finalize, mg_dataPrepare, mg_dataRestore, mg_get_arrp, mg_alloc_arr, mg_set_domain, &
       mg_get_h
contains
  !-------------------------------------------------------------------------
  ! base grid: ncells per side (power of two) at FMG level 0, cell width width
  !-------------------------------------------------------------------------
  subroutine mg_set_domain(ncells, width)
    integer,intent(IN) :: ncells
    real(kind=8),intent(IN) :: width
    NBase = ncells
    H0 = width
  end subroutine mg_set_domain
  !-------------------------------------------------------------------------
  ! mg_data.F90:45-90
  !-------------------------------------------------------------------------
  subroutine mg_data_init(fmglev)
    use grid, only : GidBase, CellWidth, Lmin
    integer,intent(IN) :: fmglev
    integer :: l
    if (fmglev == FMG_Level) return
    if (FMG_Level >= 0) call mg_data_finalize
    if (NBase <= 0) then
       if (.not. allocated(GidBase)) then
          print *, '*** mg_data_init: no base grid (module grid) and no mg_set_domain'
          stop 1
       end if
       if (size(GidBase, 1) /= size(GidBase, 2) .or. size(GidBase, 1) /= size(GidBase, 3)) then
          print *, '*** mg_data_init: the base grid must be a cube of blocks', shape(GidBase)
          stop 1
       end if
       NBase = size(GidBase, 1) * 8
       H0 = CellWidth(lbound(CellWidth, 1), Lmin)
    end if
    FMG_Level = fmglev
    Bs = 8 / 2**fmglev
    Nmg = NBase / 2**fmglev
    Hmg = H0 * 2**fmglev
    MG_LevelMin = 0
    MG_LevelMax = 0
    l = Nmg
    do while (l > 1)
       l = l / 2
       MG_LevelMax = MG_LevelMax + 1
    end do
  end subroutine mg_data_init

  subroutine mg_data_finalize
    integer :: c
    do c = ICODEMIN, ICODEMAX
       if (associated(Arr(c)%u)) deallocate(Arr(c)%u)
       nullify(Arr(c)%u)
    end do
    FMG_Level = -1
  end subroutine mg_data_finalize
  !-------------------------------------------------------------------------
  ! mg_data.F90:432-470 (level MG_LevelMin only)
  !-------------------------------------------------------------------------
  subroutine mg_alloc_arr(mglev, icode)
    integer,intent(IN) :: mglev, icode
    call check_code(mglev, icode)
    if (associated(Arr(icode)%u)) return
    allocate(Arr(icode)%u(-1:Nmg, -1:Nmg, -1:Nmg, 0:0))
    Arr(icode)%u = 0.d0
  end subroutine mg_alloc_arr

  function mg_get_arrp(mglev, icode) result(arrp)
    integer,intent(IN) :: mglev, icode
    real(kind=8),pointer,dimension(:,:,:,:) :: arrp
    call mg_alloc_arr(mglev, icode)
    arrp => Arr(icode)%u
  end function mg_get_arrp

  function mg_get_h(mglev) result(h)
    integer,intent(IN) :: mglev
    real(kind=8) :: h
    h = Hmg * 2**mglev
  end function mg_get_h

  subroutine check_code(mglev, icode)
    integer,intent(IN) :: mglev, icode
    if (FMG_Level < 0) then
       print *, '*** mg_data: call mg_init (mg_data_init) first'
       stop 1
    end if
    if (mglev /= MG_LevelMin .or. icode < ICODEMIN .or. icode > ICODEMAX) then
       print *, '*** This code is not supported', mglev, icode
       stop 1
    end if
  end subroutine check_code
  !-------------------------------------------------------------------------
  ! mg_data.F90:179-253: iumg(m) at MG_LevelMin <- interiors of iufmg(m) of
  ! every base block at FMG_Level
  !-------------------------------------------------------------------------
  subroutine mg_dataPrepare(iumg, iufmg)
    use grid, only : GidBase, RankBase
    integer,dimension(:),intent(IN) :: iufmg, iumg
    integer :: i, j, k, m
    real(kind=8),pointer,dimension(:,:,:,:) :: u, uf
    if (size(iufmg, 1) /= size(iumg, 1)) then
       print *, '*** mg_dataPrepare: error. iufmg and iumg are not consistent.', size(iufmg, 1), size(iumg, 1)
       stop
    end if
    call check_ranks(RankBase)
    do m = lbound(iumg, 1), ubound(iumg, 1)
       u => mg_get_arrp(MG_LevelMin, iumg(m))
       do k = lbound(GidBase, 3), ubound(GidBase, 3)
          do j = lbound(GidBase, 2), ubound(GidBase, 2)
             do i = lbound(GidBase, 1), ubound(GidBase, 1)
                uf => fmg_get_arrp(AMR_LevelMin, FMG_Level, GidBase(i, j, k), iufmg(m))
                u(Bs*i:Bs*(i+1)-1, Bs*j:Bs*(j+1)-1, Bs*k:Bs*(k+1)-1, :) = uf(0:Bs-1, 0:Bs-1, 0:Bs-1, :)
             end do
          end do
       end do
    end do
  end subroutine mg_dataPrepare
  !-------------------------------------------------------------------------
  ! mg_data.F90:257-330: interiors of iufmg(m) of every base block <- iumg(m)
  !-------------------------------------------------------------------------
  subroutine mg_dataRestore(iufmg, iumg)
    use grid, only : GidBase, RankBase
    integer,dimension(:),intent(IN) :: iufmg, iumg
    integer :: i, j, k, m
    real(kind=8),pointer,dimension(:,:,:,:) :: u, uf
    if (size(iufmg, 1) /= size(iumg, 1)) then
       print *, '*** mg_dataRestore: error. iufmg and iumg are not consistent.', size(iufmg, 1), size(iumg, 1)
       stop
    end if
    call check_ranks(RankBase)
    do m = lbound(iumg, 1), ubound(iumg, 1)
       u => mg_get_arrp(MG_LevelMin, iumg(m))
       do k = lbound(GidBase, 3), ubound(GidBase, 3)
          do j = lbound(GidBase, 2), ubound(GidBase, 2)
             do i = lbound(GidBase, 1), ubound(GidBase, 1)
                uf => fmg_get_arrp(AMR_LevelMin, FMG_Level, GidBase(i, j, k), iufmg(m))
                uf(0:Bs-1, 0:Bs-1, 0:Bs-1, :) = u(Bs*i:Bs*(i+1)-1, Bs*j:Bs*(j+1)-1, Bs*k:Bs*(k+1)-1, :)
             end do
          end do
       end do
    end do
  end subroutine mg_dataRestore

  subroutine check_ranks(rank_base)
    integer,dimension(:,:,:),intent(IN) :: rank_base
    if (any(rank_base /= 0)) then
       print *, '*** mg_data: every base block must be on rank 0 (one rank, one GPU)'
       stop 1
    end if
    if (Bs * size(rank_base, 1) /= Nmg .or. Bs * size(rank_base, 2) /= Nmg .or. Bs * size(rank_base, 3) /= Nmg) then
       print *, '*** mg_data: base blocks do not tile the MG grid', shape(rank_base), Bs, Nmg
       stop 1
    end if
  end subroutine check_ranks
end module mg_data
