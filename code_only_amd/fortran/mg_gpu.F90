!-------------------------------------------------------------------------
! Drop-in for the reference's uniform multigrid module `mg` (mg.F90:8-14):
! same module name and public entry points (mg_init, mg_cycle, mg_lin,
! mg_fill0, mg_boundary_fill0, mg_finalize), same array codes (IU, IRHO,
! IPSI, ISRC, ... as fmg_data.F90:89) and the same ghosted storage
! u(-1:n,-1:n,-1:n,0:0) per code (mg_data.F90:398-417).  The solve itself runs
! on the GPU through libsg's C ABI (sg_capi).  Only the level-0 arrays are held
! on the host: coarser levels live on the device for the whole run.
!
! Differences from the reference, by design:
!  * the domain size and cell width come from mg_set_domain instead of the
!    config.h macros NGI_BASE*NI (mg_data.F90:58-75);
!  * the stopping test is selectable (mg_set_solver); the default is the
!    reference's Resmaxg/rhomax < 1e-2 with at most 10 iterations
!    (mg.F90:50-52,91).
! mg_nonlin / mg_nonlin_vcycle belong to the nonlinear PDEs (ohmic, ambipolar
! diffusion), which this library does not provide: they stop with a message.
!-------------------------------------------------------------------------
module mg
  use iso_c_binding
  use sg_capi
  implicit none
  private
  integer,parameter,public :: IDBG=-1, IU=0, IRES=1, IRHS=2, IRHO=3, IFLX=4, ISRC=5, IPSI=6, IRUF=7
  integer,parameter :: ICODEMIN = -1, ICODEMAX = 7
  integer,save :: NBase = 0, N = 0
  real(kind=8),save :: H0 = 0.d0, H = 0.d0
  integer,save :: FMG_Level = -1
  type(c_ptr),save :: Ctx = c_null_ptr
  real(kind=8),save,pointer :: Arr(:,:,:,:,:) => null()   ! (-1:N,-1:N,-1:N,0:0,ICODEMIN:ICODEMAX)
  type(sg_params),save :: Params
  logical,save :: ParamsSet = .false.
  type(sg_stats),save :: LastStats
  real(kind=8),save,public :: Resmaxg = 0.d0   ! mg.F90:12
  public :: mg_lin, mg_cycle, mg_boundary_fill0, mg_fill0, mg_nonlin, mg_nonlin_vcycle, mg_init, mg_finalize
  public :: mg_set_domain, mg_set_solver, mg_get_arrp, mg_last_iterations, mg_last_residual
contains
  !-------------------------------------------------------------------------
  ! base grid: nbase cells per side (power of two), cell width h0 at fmglev 0
  !-------------------------------------------------------------------------
  subroutine mg_set_domain(ncells, width)
    integer,intent(IN) :: ncells
    real(kind=8),intent(IN) :: width
    NBase = ncells
    H0 = width
  end subroutine mg_set_domain
  !-------------------------------------------------------------------------
  ! stopping test and arithmetic of mg_cycle
  !-------------------------------------------------------------------------
  subroutine mg_set_solver(tol, maxit, stop_mode, precision)
    real(kind=8),intent(IN) :: tol
    integer,intent(IN) :: maxit, stop_mode, precision
    if (.not. ParamsSet) call sg_default_params(Params)
    ParamsSet = .true.
    Params%tol = tol
    Params%maxit = maxit
    Params%stop_mode = stop_mode
    Params%precision = precision
  end subroutine mg_set_solver
  !-------------------------------------------------------------------------
  ! mg.F90:525-528 / mg_data.F90:45-90
  !-------------------------------------------------------------------------
  subroutine mg_init(fmglev)
    integer,intent(IN) :: fmglev
    if (fmglev == FMG_Level) return
    if (FMG_Level >= 0) call mg_finalize
    if (NBase <= 0) then
       print *, '*** mg_init: call mg_set_domain first'
       stop 1
    end if
    FMG_Level = fmglev
    N = NBase / 2**fmglev
    H = H0 * 2**fmglev
    if (.not. ParamsSet) then
       call sg_default_params(Params)
       ParamsSet = .true.
    end if
    call sg_check(sg_uniform_create(Ctx, int(N, c_int), real(H, c_double), SG_PREC_FP64 + SG_PREC_MIXED), 'mg_init')
    allocate(Arr(-1:N, -1:N, -1:N, 0:0, ICODEMIN:ICODEMAX))
    Arr = 0.d0
  end subroutine mg_init
  !-------------------------------------------------------------------------
  ! mg.F90:532-536
  !-------------------------------------------------------------------------
  subroutine mg_finalize
    integer(c_int) :: rc
    if (c_associated(Ctx)) rc = sg_uniform_destroy(Ctx)
    Ctx = c_null_ptr
    if (associated(Arr)) deallocate(Arr)
    nullify(Arr)
    FMG_Level = -1
  end subroutine mg_finalize
  !-------------------------------------------------------------------------
  ! level-0 array of a code (mg_data.F90:472-504); coarser levels are on the GPU
  !-------------------------------------------------------------------------
  function mg_get_arrp(mglev, icode) result(arrp)
    integer,intent(IN) :: mglev, icode
    real(kind=8),pointer,dimension(:,:,:,:) :: arrp
    if (mglev /= 0 .or. icode < ICODEMIN .or. icode > ICODEMAX .or. .not. associated(Arr)) then
       print *, '*** This code is not supported', mglev, icode
       stop 1
    end if
    ! keep the reference's bounds u(-1:n,-1:n,-1:n,0:0): a plain section pointer would start at 1
    arrp(-1:, -1:, -1:, 0:) => Arr(:,:,:,:,icode)
  end function mg_get_arrp
  !-------------------------------------------------------------------------
  ! mg.F90:654-660
  !-------------------------------------------------------------------------
  subroutine mg_fill0(mglev, icode)
    integer,intent(IN) :: mglev, icode
    real(kind=8),pointer,dimension(:,:,:,:) :: u
    u => mg_get_arrp(mglev, icode)
    u = 0.d0
  end subroutine mg_fill0
  !-------------------------------------------------------------------------
  ! mg.F90:735-750; the periodic domain has no physical boundary, and every
  ! solve wraps the ghosts itself (fmg_boundary_phys_periodic.F90:71-104)
  !-------------------------------------------------------------------------
  subroutine mg_boundary_fill0(mglev, icode)
    integer,intent(IN) :: mglev, icode
    real(kind=8),pointer,dimension(:,:,:,:) :: u
    u => mg_get_arrp(mglev, icode)
  end subroutine mg_boundary_fill0
  !-------------------------------------------------------------------------
  ! mg.F90:42-94: ju (initial guess in, solution out) from the source jrho
  !-------------------------------------------------------------------------
  subroutine mg_cycle(ju, jrho)
    integer,intent(IN) :: ju, jrho
    real(kind=8),pointer,dimension(:,:,:,:) :: u, f
    integer(c_int) :: rc
    u => mg_get_arrp(0, ju)
    f => mg_get_arrp(0, jrho)
    rc = sg_uniform_cycle_host(Ctx, u, f, Params, LastStats)
    if (rc /= SG_ERR_NOT_CONVERGED) call sg_check(rc, 'mg_cycle')
    Resmaxg = LastStats%resmaxg
  end subroutine mg_cycle
  !-------------------------------------------------------------------------
  ! mg.F90:106-218: one FMG solve of IU from IRHO with zero initial guess
  !-------------------------------------------------------------------------
  subroutine mg_lin
    type(sg_params) :: p
    real(kind=8),pointer,dimension(:,:,:,:) :: u, f
    integer(c_int) :: rc
    p = Params
    p%maxit = 1
    p%tol = 0.d0
    p%stop_mode = SG_STOP_TRUE_RESIDUAL
    u => mg_get_arrp(0, IU)
    f => mg_get_arrp(0, IRHO)
    u = 0.d0
    rc = sg_uniform_cycle_host(Ctx, u, f, p, LastStats)
    if (rc /= SG_ERR_NOT_CONVERGED) call sg_check(rc, 'mg_lin')
    Resmaxg = LastStats%resmaxg
  end subroutine mg_lin

  integer function mg_last_iterations()
    mg_last_iterations = LastStats%iterations
  end function mg_last_iterations

  real(kind=8) function mg_last_residual()
    mg_last_residual = LastStats%rel_linf
  end function mg_last_residual

  subroutine mg_nonlin(ju, jrho)
    integer,intent(IN) :: ju, jrho
    print *, '*** mg_nonlin: nonlinear PDEs are not provided by this library', ju, jrho
    stop 1
  end subroutine mg_nonlin

  subroutine mg_nonlin_vcycle(ju, jrho)
    integer,intent(IN) :: ju, jrho
    print *, '*** mg_nonlin_vcycle: nonlinear PDEs are not provided by this library', ju, jrho
    stop 1
  end subroutine mg_nonlin_vcycle
end module mg
