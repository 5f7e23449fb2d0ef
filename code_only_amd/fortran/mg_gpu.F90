!-------------------------------------------------------------------------
! Drop-in for the reference's uniform multigrid module `mg` (mg.F90:8-14):
! same module name and public entry points (mg_init, mg_cycle, mg_lin,
! mg_fill0, mg_boundary_fill0, mg_finalize), working on module mg_data's
! arrays (mg_data_gpu.F90: same codes as fmg_data.F90:89, same ghosted storage
! u(-1:n,-1:n,-1:n,0:0) per code, filled by mg_dataPrepare).  The solve itself
! runs on the GPU through libsg's C ABI (sg_capi).  Only the level-0 arrays are
! held on the host: coarser levels live on the device for the whole run.
!
! Differences from the reference, by design:
!  * the domain size and cell width come from the host's module grid
!    (GidBase, CellWidth) or from mg_set_domain, instead of the config.h
!    macros NGI_BASE*NI (mg_data.F90:58-75);
!  * the stopping test is selectable (mg_set_solver); the default is the
!    reference's Resmaxg/rhomax < 1e-2 with at most 10 iterations
!    (mg.F90:50-52,91).
! mg_nonlin / mg_nonlin_vcycle belong to the nonlinear PDEs (ohmic, ambipolar
! diffusion), which this library does not provide: they stop with a message.
!-------------------------------------------------------------------------
module mg
  use iso_c_binding
  use sg_capi
  use mg_data
  implicit none
  private
  type(c_ptr),save :: Ctx = c_null_ptr
  integer,save :: CtxLevel = -1
  type(sg_params),save :: Params
  logical,save :: ParamsSet = .false.
  type(sg_stats),save :: LastStats
  real(kind=8),save,public :: Resmaxg = 0.d0   ! mg.F90:12
  public :: IDBG, IU, IRES, IRHS, IRHO, IFLX, ISRC, IPSI, IRUF
  public :: mg_lin, mg_cycle, mg_boundary_fill0, mg_fill0, mg_nonlin, mg_nonlin_vcycle, mg_init, mg_finalize
  public :: mg_set_domain, mg_set_solver, mg_get_arrp, mg_last_iterations, mg_last_residual
contains
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
  ! mg.F90:525-528: mg_data_init, then the device context of that grid
  !-------------------------------------------------------------------------
  subroutine mg_init(fmglev)
    integer,intent(IN) :: fmglev
    if (fmglev == CtxLevel .and. c_associated(Ctx)) return
    call release_ctx
    call mg_data_init(fmglev)
    if (.not. ParamsSet) then
       call sg_default_params(Params)
       ParamsSet = .true.
    end if
    call sg_check(sg_uniform_create(Ctx, int(Nmg, c_int), real(Hmg, c_double), SG_PREC_FP64 + SG_PREC_MIXED), 'mg_init')
    CtxLevel = fmglev
  end subroutine mg_init

  subroutine release_ctx
    integer(c_int) :: rc
    if (c_associated(Ctx)) rc = sg_uniform_destroy(Ctx)
    Ctx = c_null_ptr
    CtxLevel = -1
  end subroutine release_ctx
  !-------------------------------------------------------------------------
  ! mg.F90:532-536
  !-------------------------------------------------------------------------
  subroutine mg_finalize
    call release_ctx
    call mg_data_finalize
  end subroutine mg_finalize
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
    call need_ctx('mg_cycle')
    u => mg_get_arrp(MG_LevelMin, ju)
    f => mg_get_arrp(MG_LevelMin, jrho)
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
    call need_ctx('mg_lin')
    p = Params
    p%maxit = 1
    p%tol = 0.d0
    p%stop_mode = SG_STOP_TRUE_RESIDUAL
    u => mg_get_arrp(MG_LevelMin, IU)
    f => mg_get_arrp(MG_LevelMin, IRHO)
    u = 0.d0
    rc = sg_uniform_cycle_host(Ctx, u, f, p, LastStats)
    if (rc /= SG_ERR_NOT_CONVERGED) call sg_check(rc, 'mg_lin')
    Resmaxg = LastStats%resmaxg
  end subroutine mg_lin

  subroutine need_ctx(where)
    character(len=*),intent(IN) :: where
    if (.not. c_associated(Ctx)) then
       print *, '*** ', where, ': call mg_init first'
       stop 1
    end if
  end subroutine need_ctx

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
