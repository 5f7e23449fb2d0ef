!-------------------------------------------------------------------------
! ISO_C_BINDING interface to libsg (include/sg.h): the C ABI that the
! Fortran drop-in modules call.  Types and constants mirror sg.h one to one.
!-------------------------------------------------------------------------
module sg_capi
  use iso_c_binding
  implicit none
  integer(c_int),parameter :: SG_OK = 0, SG_ERR_ARG = 1, SG_ERR_HIP = 2, &
       SG_ERR_NOT_CONVERGED = 3, SG_ERR_NONFINITE = 4
  integer(c_int),parameter :: SG_STOP_REFERENCE = 0, SG_STOP_TRUE_RESIDUAL = 1
  integer(c_int),parameter :: SG_PREC_FP64 = 1, SG_PREC_MIXED = 2
  integer(c_int),parameter :: SG_BC_PERIODIC = 0, SG_BC_SPHERE = 1
  integer(c_int),parameter :: SG_COARSE_EXACT = 0, SG_COARSE_VCYCLE = 1
  integer,parameter :: SG_HIST = 64

  type, bind(C) :: sg_params
     real(c_double) :: tol
     integer(c_int) :: maxit
     integer(c_int) :: stop_mode
     integer(c_int) :: precision
     integer(c_int) :: ncycle
     integer(c_int) :: npre
     integer(c_int) :: npost
     integer(c_int) :: use_graph
     integer(c_int) :: fused
     integer(c_int) :: coarse
  end type sg_params

  type, bind(C) :: sg_stats
     integer(c_int) :: iterations
     integer(c_int) :: status
     real(c_double) :: rel_linf
     real(c_double) :: rel_l2
     real(c_double) :: resmaxg
     real(c_double) :: ms
     real(c_double) :: hist(SG_HIST)
     real(c_double) :: rel_h2
  end type sg_stats

  ! hydro block memory on the device (sg.h: sg_blockmem, sg_timeslice)
  integer,parameter :: SG_AMR_MAX_LEVELS = 16
  type, bind(C) :: sg_blockmem
     integer(c_int) :: ng
     integer(c_int) :: ncomp
     integer(c_long_long) :: stride
     integer(c_int) :: mrho, mvx, mvy, mvz, mp, mpsi, mgx, mgy, mgz
  end type sg_blockmem

  type, bind(C) :: sg_timeslice
     integer(c_int) :: source
     real(c_double) :: step_max
     real(c_double) :: step(SG_AMR_MAX_LEVELS)
     real(c_double) :: dstep(SG_AMR_MAX_LEVELS)
  end type sg_timeslice

  interface
     function sg_last_error() bind(C, name='sg_last_error') result(p)
       import :: c_ptr
       type(c_ptr) :: p
     end function sg_last_error

     subroutine sg_default_params(p) bind(C, name='sg_default_params')
       import :: sg_params
       type(sg_params),intent(OUT) :: p
     end subroutine sg_default_params

     function sg_uniform_create(ctx, n, h, precisions) bind(C, name='sg_uniform_create') result(rc)
       import :: c_ptr, c_int, c_double
       type(c_ptr),intent(OUT) :: ctx
       integer(c_int),value :: n, precisions
       real(c_double),value :: h
       integer(c_int) :: rc
     end function sg_uniform_create

     function sg_uniform_destroy(ctx) bind(C, name='sg_uniform_destroy') result(rc)
       import :: c_ptr, c_int
       type(c_ptr),value :: ctx
       integer(c_int) :: rc
     end function sg_uniform_destroy

     function sg_uniform_cycle_host(ctx, psi_gh, src_gh, p, st) bind(C, name='sg_uniform_cycle_host') result(rc)
       import :: c_ptr, c_int, c_double, sg_params, sg_stats
       type(c_ptr),value :: ctx
       real(c_double),intent(INOUT) :: psi_gh(*)
       real(c_double),intent(IN) :: src_gh(*)
       type(sg_params),intent(IN) :: p
       type(sg_stats),intent(OUT) :: st
       integer(c_int) :: rc
     end function sg_uniform_cycle_host

     function sg_amr_create(ctx, nblk, nlev, level, ijk, parent, child, nbr, base_dims, h0, x0, bc) &
          bind(C, name='sg_amr_create') result(rc)
       import :: c_ptr, c_int, c_double
       type(c_ptr),intent(OUT) :: ctx
       integer(c_int),value :: nblk, nlev, bc
       integer(c_int),intent(IN) :: level(*), ijk(*), parent(*), child(*), nbr(*), base_dims(3)
       real(c_double),value :: h0
       real(c_double),intent(IN) :: x0(3)
       integer(c_int) :: rc
     end function sg_amr_create

     function sg_amr_destroy(ctx) bind(C, name='sg_amr_destroy') result(rc)
       import :: c_ptr, c_int
       type(c_ptr),value :: ctx
       integer(c_int) :: rc
     end function sg_amr_destroy

     function sg_amr_set_zero_average(ctx, on) bind(C, name='sg_amr_set_zero_average') result(rc)
       import :: c_ptr, c_int
       type(c_ptr),value :: ctx
       integer(c_int),value :: on
       integer(c_int) :: rc
     end function sg_amr_set_zero_average

     function sg_amr_solve_host(ctx, psi, src, g, p, st) bind(C, name='sg_amr_solve_host') result(rc)
       import :: c_ptr, c_int, c_double, sg_params, sg_stats
       type(c_ptr),value :: ctx
       real(c_double),intent(INOUT) :: psi(*)
       real(c_double),intent(IN) :: src(*)
       real(c_double),intent(OUT) :: g(*)
       type(sg_params),intent(IN) :: p
       type(sg_stats),intent(OUT) :: st
       integer(c_int) :: rc
     end function sg_amr_solve_host

     function sg_amr_set_blockmem(ctx, lay, pool_slots, slot_u, slot_u1, slot_u2) &
          bind(C, name='sg_amr_set_blockmem') result(rc)
       import :: c_ptr, c_int, c_long_long, sg_blockmem
       type(c_ptr),value :: ctx
       type(sg_blockmem),intent(IN) :: lay
       integer(c_long_long),value :: pool_slots
       integer(c_int),intent(IN) :: slot_u(*)
       type(c_ptr),value :: slot_u1, slot_u2   ! c_loc of the tables, or c_null_ptr
       integer(c_int) :: rc
     end function sg_amr_set_blockmem

     ! pool: the device address of the grids' arrays (from the GPU-resident hydro code)
     function sg_amr_solve_blockmem_dev(ctx, pool, ts, pi4g, p, st) &
          bind(C, name='sg_amr_solve_blockmem_dev') result(rc)
       import :: c_ptr, c_int, c_double, sg_timeslice, sg_params, sg_stats
       type(c_ptr),value :: ctx, pool
       type(sg_timeslice),intent(IN) :: ts
       real(c_double),value :: pi4g
       type(sg_params),intent(IN) :: p
       type(sg_stats),intent(OUT) :: st
       integer(c_int) :: rc
     end function sg_amr_solve_blockmem_dev

     function sg_amr_source_g_dev(ctx, pool, dtime, gamma) bind(C, name='sg_amr_source_g_dev') result(rc)
       import :: c_ptr, c_int, c_double
       type(c_ptr),value :: ctx, pool
       real(c_double),intent(IN) :: dtime(*)
       real(c_double),value :: gamma
       integer(c_int) :: rc
     end function sg_amr_source_g_dev
  end interface

contains

  ! message of the last failure, as a Fortran string
  function sg_error_message() result(msg)
    character(len=1024) :: msg
    character(kind=c_char),pointer :: cs(:)
    type(c_ptr) :: p
    integer :: i
    msg = ''
    p = sg_last_error()
    if (.not. c_associated(p)) return
    call c_f_pointer(p, cs, [1024])
    do i = 1, 1024
       if (cs(i) == c_null_char) exit
       msg(i:i) = cs(i)
    end do
  end function sg_error_message

  ! the reference prints a message and stops on errors (e.g. mg_data.F90:187-190)
  subroutine sg_check(rc, where)
    integer(c_int),intent(IN) :: rc
    character(len=*),intent(IN) :: where
    if (rc == SG_OK) return
    print *, '*** ', where, ': ', trim(sg_error_message())
    stop 1
  end subroutine sg_check
end module sg_capi
