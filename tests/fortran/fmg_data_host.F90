!-------------------------------------------------------------------------
! Test harness: a host-side module `fmg_data` with the names the reference's
! fmg_data gives the uniform solver's data exchange (fmg_data.F90:21,89-91,
! 156-188, 796-804): the array codes, AMR_LevelMin, fmg_get_arrp(amr_level,
! fmg_level, grid_id, icode) returning a block's ghosted array
! u(-1:n,-1:n,-1:n,0:0), and fmg_getNcomp.  Storage: base-level blocks only,
! one FMG level, set up by fmg_data_host_alloc.  It exists so that the drop-in
! modules `mg_data` / `mg` can be compiled and driven as the reference's
! mg_test (fmg_test.F90:8-77) and vmg_slvsml (vmg.F90:387-432) drive them; it
! is not part of the library.
!-------------------------------------------------------------------------
module fmg_data
  implicit none
  integer,parameter :: IDBG=-1, IU=0, IRES=1, IRHS=2, IRHO=3, IFLX=4, ISRC=5, IPSI=6, IRUF=7, IETA=8, IDOD=9, &
       IDHE=10, IDAD=11
  integer,parameter :: ICODEMIN = -1, ICODEMAX = 11
  integer,save :: AMR_LevelMin = 0
  integer,save,private :: HostFmgLevel = -1, HostN = 0
  type t_hostarr
     real(kind=8),dimension(:,:,:,:),pointer :: u => null()
  end type t_hostarr
  type(t_hostarr),save,allocatable,private :: Store(:,:)   ! (gid, icode)
contains
  subroutine fmg_data_host_alloc(ngid, fmglev)
    integer,intent(IN) :: ngid, fmglev
    integer :: g, c
    HostFmgLevel = fmglev
    HostN = 8 / 2**fmglev
    allocate(Store(0:ngid-1, ICODEMIN:ICODEMAX))
    do c = ICODEMIN, ICODEMAX
       do g = 0, ngid - 1
          allocate(Store(g, c)%u(-1:HostN, -1:HostN, -1:HostN, 0:0))
          Store(g, c)%u = 0.d0
       end do
    end do
  end subroutine fmg_data_host_alloc

  function fmg_get_arrp(amr_level, fmg_level, grid_id, icode) result(arrp)
    integer,intent(IN) :: amr_level, fmg_level, grid_id, icode
    real(kind=8),pointer,dimension(:,:,:,:) :: arrp
    if (amr_level /= AMR_LevelMin .or. fmg_level /= HostFmgLevel) then
       print *, '*** fmg_get_arrp (test harness): only the base level at FMG level', HostFmgLevel
       stop 1
    end if
    arrp => Store(grid_id, icode)%u
  end function fmg_get_arrp

  function fmg_getNcomp(icode) result(ncomp)
    integer,intent(IN) :: icode
    integer :: ncomp
    ncomp = 1
    if (icode < ICODEMIN .or. icode > ICODEMAX) ncomp = 0
  end function fmg_getNcomp
end module fmg_data
