!-------------------------------------------------------------------------
! Drop-in for the reference's module `fmg` (fmg.F90:13-74): same module name
! and the same argument-less entry point fmg_poisson, as multi_timestep.F90
! calls it.  It reads the hydro grid through module `grid` where the
! reference's staging reads it (fmg_data_if.F90): GidList / GidListMax per
! level, Igrid / Jgrid / Kgrid, ParentGid, ChildGid, NeighborGid, CellWidth,
! and the block components through get_Ucomp / get_U1comp.  The solve runs on
! the GPU through libsg's AMR entry points (sg_amr_*).  The potential goes to
! component MPSI and g = -grad psi to MGX..MGZ (fmg.F90:214-215,
! fmg_poisson.F90:4-49).
!
! What fmg_poisson does, in the reference's order:
!  * link lists (fmg_make_linklist, fmg_data_if.F90:298-411): blocks numbered
!    level by level in GidList order; a face without a same-level neighbour
!    takes the parent's neighbour (NeighborParentLevel) or none (domain
!    boundary).  The tables and the device context are kept while the
!    hierarchy is unchanged;
!  * physical variables (fmg_make_physvar, fmg_data_if.F90:13-84): 4 pi G rho
!    from the chosen time slice of component MRHO, the initial guess from
!    component MPSI; the leaf mean is removed for periodic boundaries
!    (ZERO_AVERAGE_DENSITY) inside the library;
!  * fmg_cycle (fmg.F90:134-233): by default its own stopping test,
!    Resh2maxg / h2absmax < 1e-3 with at most 10 iterations and the stall test;
!  * fmg_restore_u and fmg_psi2g (fmg_data_if.F90:269-292,
!    fmg_poisson.F90:4-49).
!
! Differences from the reference, by design:
!  * 4 pi G, the boundary kind, the domain origin and the component indices
!    (config.h MRHO / MPSI / MGX) come from fmg_set_problem instead of
!    modelParameter, the boundary module and the preprocessor;
!  * one MPI rank: link arrays are read at rank 0.
!-------------------------------------------------------------------------
module fmg
  use iso_c_binding
  use sg_capi
  use grid, only : LevelMax, GidList, GidListMax, Gidmin, Igrid, Jgrid, Kgrid, ParentGid, ChildGid, &
       NeighborGid, CellWidth, get_Ucomp, get_U1comp, Left, Right, Undefi
  implicit none
  private
  integer,save :: Mrho = 0, Mpsi = 5, Mgx = 6        ! config.h:63-69 (isothermal MHD layout)
  integer,save :: SourceSlice = 1                     ! 1: U1 (SINGLE_STEP predictor), 0: U
  real(kind=8),save :: Pi4G = 1.d0
  integer(c_int),save :: Bc = SG_BC_PERIODIC
  real(c_double),save :: X0(3) = 0.d0
  logical,save :: X0Set = .false.
  type(sg_params),save :: Params
  logical,save :: ParamsSet = .false.
  type(sg_stats),save :: LastStats
  type(c_ptr),save :: Ctx = c_null_ptr
  ! block tables of the current context (0-based block numbers)
  integer(c_int),save,allocatable :: TLevel(:), TIjk(:), TParent(:), TChild(:), TNbr(:)
  integer,save,allocatable :: BlockGid(:)
  integer,save :: Nblk = 0, Nlev = 0
  public :: fmg_poisson, fmg_set_problem, fmg_set_solver, fmg_finalize, fmg_last_iterations, fmg_last_residual
contains
  !-------------------------------------------------------------------------
  ! pi4g = 4 pi G; bc = SG_BC_PERIODIC or SG_BC_SPHERE; x0 = domain lower
  ! corner (used by the sphere boundary's multipole geometry); components of
  ! rho, psi and gx; source_slice = 1 for U1 (predictor), 0 for U.
  !-------------------------------------------------------------------------
  subroutine fmg_set_problem(pi4g_in, bc_in, x0_in, mrho_in, mpsi_in, mgx_in, source_slice)
    real(kind=8),intent(IN) :: pi4g_in
    integer,intent(IN) :: bc_in
    real(kind=8),intent(IN) :: x0_in(3)
    integer,intent(IN) :: mrho_in, mpsi_in, mgx_in, source_slice
    Pi4G = pi4g_in
    if (bc_in /= Bc) call fmg_finalize
    Bc = int(bc_in, c_int)
    X0 = x0_in
    X0Set = .true.
    Mrho = mrho_in
    Mpsi = mpsi_in
    Mgx = mgx_in
    SourceSlice = source_slice
  end subroutine fmg_set_problem
  !-------------------------------------------------------------------------
  ! stopping test and arithmetic of fmg_cycle (default: fmg.F90:146-179)
  !-------------------------------------------------------------------------
  subroutine fmg_set_solver(tol, maxit, stop_mode, precision)
    real(kind=8),intent(IN) :: tol
    integer,intent(IN) :: maxit, stop_mode, precision
    call default_params
    Params%tol = tol
    Params%maxit = maxit
    Params%stop_mode = stop_mode
    Params%precision = precision
  end subroutine fmg_set_solver

  subroutine default_params
    if (ParamsSet) return
    call sg_default_params(Params)
    Params%tol = 1.d-3          ! fmg.F90:146 errormax
    Params%maxit = 10           ! fmg.F90:179 nfmgmax
    Params%stop_mode = SG_STOP_REFERENCE
    Params%precision = SG_PREC_FP64
    ParamsSet = .true.
  end subroutine default_params

  integer function fmg_last_iterations()
    fmg_last_iterations = LastStats%iterations
  end function fmg_last_iterations

  real(kind=8) function fmg_last_residual()
    fmg_last_residual = LastStats%rel_linf
  end function fmg_last_residual
  !-------------------------------------------------------------------------
  ! fmg.F90:63-74
  !-------------------------------------------------------------------------
  subroutine fmg_poisson
    real(kind=8),allocatable :: psi(:), src(:), g(:)
    real(kind=8),dimension(:,:,:),pointer :: u
    integer :: b, gid, m
    integer(c_int) :: rc
    call default_params
    call make_linklist
    allocate(psi(0:512*Nblk-1), src(0:512*Nblk-1), g(0:3*512*Nblk-1))
    do b = 0, Nblk - 1
       gid = BlockGid(b)
       if (SourceSlice == 1) then
          u => get_U1comp(Mrho, gid)
       else
          u => get_Ucomp(Mrho, gid)
       end if
       src(512*b:512*b+511) = reshape(u(0:7, 0:7, 0:7), [512]) * Pi4G
       u => get_Ucomp(Mpsi, gid)
       psi(512*b:512*b+511) = reshape(u(0:7, 0:7, 0:7), [512])
    end do
    rc = sg_amr_solve_host(Ctx, psi, src, g, Params, LastStats)
    ! fmg_cycle ends its loop after nfmgmax iterations or on the stall test without an error
    ! (fmg.F90:179-196), so non-convergence is reported in LastStats, not fatal
    if (rc /= SG_ERR_NOT_CONVERGED) call sg_check(rc, 'fmg_poisson')
    do b = 0, Nblk - 1
       gid = BlockGid(b)
       u => get_Ucomp(Mpsi, gid)
       u(0:7, 0:7, 0:7) = reshape(psi(512*b:512*b+511), [8, 8, 8])
       do m = 0, 2
          u => get_Ucomp(Mgx + m, gid)
          u(0:7, 0:7, 0:7) = reshape(g(512*(3*b+m):512*(3*b+m)+511), [8, 8, 8])
       end do
    end do
    deallocate(psi, src, g)
  end subroutine fmg_poisson
  !-------------------------------------------------------------------------
  ! fmg_make_linklist (fmg_data_if.F90:298-411) on one rank; the device
  ! context is rebuilt only when the tables change.
  !-------------------------------------------------------------------------
  subroutine make_linklist
    integer :: nb, nl, lev, n, b, gid, pgid, cgid, ngid, lri, lrj, lrk, lr, d, base_dims(3)
    integer,allocatable :: rev(:), lv(:), ijk(:), par(:), chi(:), nbr(:), gids(:)
    logical :: same
    real(c_double) :: x0c(3), h0
    nl = LevelMax + 1
    nb = 0
    do lev = 0, LevelMax
       nb = nb + (GidListMax(lev) - Gidmin + 1)
    end do
    allocate(lv(0:nb-1), ijk(0:3*nb-1), par(0:nb-1), chi(0:8*nb-1), nbr(0:6*nb-1), gids(0:nb-1))
    allocate(rev(lbound(Igrid, 1):ubound(Igrid, 1)))
    rev = -1
    b = 0
    do lev = 0, LevelMax
       do n = Gidmin, GidListMax(lev)
          gid = GidList(n, lev)
          gids(b) = gid
          rev(gid) = b
          lv(b) = lev
          ijk(3*b:3*b+2) = [Igrid(gid), Jgrid(gid), Kgrid(gid)]
          b = b + 1
       end do
    end do
    base_dims = 0
    do b = 0, nb - 1
       gid = gids(b)
       par(b) = -1
       if (lv(b) > 0) par(b) = rev(ParentGid(gid, 0))
       do lrk = Left, Right
          do lrj = Left, Right
             do lri = Left, Right
                cgid = ChildGid(lri, lrj, lrk, gid, 0)
                chi(8*b + lri + 2*lrj + 4*lrk) = -1
                if (cgid /= Undefi) chi(8*b + lri + 2*lrj + 4*lrk) = rev(cgid)
             end do
          end do
       end do
       do d = 0, 2
          do lr = Left, Right
             ngid = NeighborGid(lr, d, gid, 0)
             if (ngid /= Undefi) then
                nbr(6*b + 2*d + lr) = rev(ngid)
             else if (lv(b) > 0) then
                ! NeighborParentLevel: the parent's neighbour across the same face
                pgid = ParentGid(gid, 0)
                ngid = NeighborGid(lr, d, pgid, 0)
                nbr(6*b + 2*d + lr) = -1
                if (ngid /= Undefi) nbr(6*b + 2*d + lr) = rev(ngid)
             else
                nbr(6*b + 2*d + lr) = -1
             end if
          end do
       end do
       if (lv(b) == 0) base_dims = max(base_dims, ijk(3*b:3*b+2) + 1)
    end do
    same = c_associated(Ctx) .and. nb == Nblk .and. nl == Nlev
    if (same) same = all(TLevel == lv) .and. all(TIjk == ijk) .and. all(TParent == par) .and. &
         all(TChild == chi) .and. all(TNbr == nbr)
    if (.not. same) then
       call fmg_finalize
       Nblk = nb
       Nlev = nl
       allocate(TLevel(0:nb-1), TIjk(0:3*nb-1), TParent(0:nb-1), TChild(0:8*nb-1), TNbr(0:6*nb-1), BlockGid(0:nb-1))
       TLevel = lv; TIjk = ijk; TParent = par; TChild = chi; TNbr = nbr
       h0 = CellWidth(0, 0)
       x0c = X0
       if (.not. X0Set .and. Bc == SG_BC_SPHERE) x0c = -0.5d0 * 8 * base_dims * h0   ! centred domain
       call sg_check(sg_amr_create(Ctx, int(nb, c_int), int(nl, c_int), TLevel, TIjk, TParent, TChild, TNbr, &
            int(base_dims, c_int), h0, x0c, Bc), 'fmg_poisson (link list)')
    end if
    BlockGid = gids
    deallocate(rev, lv, ijk, par, chi, nbr, gids)
  end subroutine make_linklist
  !-------------------------------------------------------------------------
  ! fmg.F90:246-258 (release of the FMG data)
  !-------------------------------------------------------------------------
  subroutine fmg_finalize
    integer(c_int) :: rc
    if (c_associated(Ctx)) rc = sg_amr_destroy(Ctx)
    Ctx = c_null_ptr
    if (allocated(TLevel)) deallocate(TLevel, TIjk, TParent, TChild, TNbr, BlockGid)
    Nblk = 0
    Nlev = 0
  end subroutine fmg_finalize
end module fmg
