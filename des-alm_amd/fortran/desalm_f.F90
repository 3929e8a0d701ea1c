! desalm_f.F90 -- ISO_C_BINDING interface to libdesalm.so (include/desalm.h)
! and the one-time context set-up shared by the drop-in mg_driver.F90 and
! solver_daf.F90 of this directory.
!
! Compiled inside the DES-ALM build (it uses that code's modules global,
! global_param, global_app, global_mpi).  Arrays cross the boundary as
! real(c_double); the module arrays are real(kind=rdf), FP32 as the
! reference ships (src/precision.F90:11) or FP64, and are converted here.
module desalm_f
  use, intrinsic :: iso_c_binding
  implicit none
  private

  integer(c_int32_t), parameter, public :: DESALM_PREC_FP64 = 0

  ! struct desalm_params of include/desalm.h (C layout via bind(C))
  type, bind(C), public :: desalm_params_t
     integer(c_int32_t) :: ng
     integer(c_int32_t) :: img, jmg, kmg
     integer(c_int32_t) :: icrs, jcrs, kcrs
     integer(c_int32_t) :: iter(6)
     integer(c_int32_t) :: itm
     integer(c_int32_t) :: it_min
     real(c_double) :: er_min, eo_min
     real(c_double) :: ren, beta, delti
     integer(c_int32_t) :: unsteady, des, itk, precision
     real(c_double) :: cfl1, vnn1, cfl2, vnn2
     real(c_double) :: pdiss_coef
     real(c_double) :: ep(3, 2)          ! C ep[a][d] is ep(d+1, a+1) here
     integer(c_int32_t) :: btype(6), bdir(6)
     integer(c_int32_t) :: ifix, jfix, kfix
     integer(c_int32_t) :: reserved(8)
  end type desalm_params_t

  public :: desalm_init, desalm_set_metrics, desalm_set_coordinates, desalm_set_hooks, &
            desalm_set_wall_distance, desalm_set_state, &
            desalm_get_state, desalm_mg_driver, desalm_solver_daf, desalm_finalize
  public :: desalm_f_ctx, desalm_f_setup, desalm_f_check

  interface
     integer(c_int) function desalm_init(p, device, ctx) bind(C, name = 'desalm_init')
       import :: c_int, c_ptr, desalm_params_t
       type(desalm_params_t), intent(in) :: p
       integer(c_int), value :: device
       type(c_ptr), intent(out) :: ctx
     end function desalm_init

     integer(c_int) function desalm_set_metrics(ctx, csi, eta, zet, aj) bind(C, name = 'desalm_set_metrics')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: csi(*), eta(*), zet(*), aj(*)
     end function desalm_set_metrics

     integer(c_int) function desalm_set_coordinates(ctx, x, y, z) bind(C, name = 'desalm_set_coordinates')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: x(*), y(*), z(*)
     end function desalm_set_coordinates

     ! conservation (1) and the j-wall wall functions (2) of mg_driver
     integer(c_int) function desalm_set_hooks(ctx, hooks) bind(C, name = 'desalm_set_hooks')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: hooks
     end function desalm_set_hooks

     integer(c_int) function desalm_set_wall_distance(ctx, wd, dtev) bind(C, name = 'desalm_set_wall_distance')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: wd(*), dtev(*)
     end function desalm_set_wall_distance

     integer(c_int) function desalm_set_state(ctx, q, qn, qnm1, xnut) bind(C, name = 'desalm_set_state')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: q(*), qn(*), qnm1(*), xnut(*)
     end function desalm_set_state

     integer(c_int) function desalm_get_state(ctx, q, qn, qnm1, xnut) bind(C, name = 'desalm_get_state')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: q(*), qn(*), qnm1(*), xnut(*)
     end function desalm_get_state

     integer(c_int) function desalm_mg_driver(ctx, itc, resid) bind(C, name = 'desalm_mg_driver')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), intent(out) :: itc
       real(c_double), intent(out) :: resid(*)
     end function desalm_mg_driver

     integer(c_int) function desalm_solver_daf(ctx, n, recalc, calc_pk, q, qn, qnm1, csi, eta, zet, aj, &
                                               xnut, pk, rh) bind(C, name = 'desalm_solver_daf')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: n, recalc, calc_pk
       real(c_double), intent(inout) :: q(*)
       real(c_double), intent(in) :: qn(*), qnm1(*), csi(*), eta(*), zet(*), aj(*), xnut(*)
       real(c_double), intent(inout) :: pk(*), rh(*)
     end function desalm_solver_daf

     type(c_ptr) function desalm_last_error(ctx) bind(C, name = 'desalm_last_error')
       import :: c_ptr
       type(c_ptr), value :: ctx
     end function desalm_last_error

     integer(c_int) function desalm_finalize(ctx) bind(C, name = 'desalm_finalize')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function desalm_finalize

     integer(c_size_t) function c_strlen(s) bind(C, name = 'strlen')
       import :: c_size_t, c_ptr
       type(c_ptr), value :: s
     end function c_strlen
  end interface

  ! the rank's context: one zone, one device, created on first use
  type(c_ptr), save :: desalm_f_ctx = c_null_ptr

contains

  ! Stop the run on a library error, in the reference's fatal style
  ! (message, then mpi_abort, src/init.F90:710).
  subroutine desalm_f_check(rc, what)
    use global_mpi
    integer(c_int), intent(in) :: rc
    character(len = *), intent(in) :: what
    character(kind = c_char), pointer :: msg(:)
    type(c_ptr) :: p
    integer :: n, ierr2
    if (rc == 0) return
    p = desalm_last_error(desalm_f_ctx)
    if (c_associated(p)) then
       n = int(c_strlen(p))
       call c_f_pointer(p, msg, [n])
       write (*, '(a,a,i0,a,*(a))') what, ' failed, code ', rc, ': ', msg(1:n)
    else
       write (*, '(a,a,i0)') what, ' failed, code ', rc
    end if
    call mpi_abort(mpi_comm_world, 1, ierr2)
  end subroutine desalm_f_check

  ! Build the context from the module state that init leaves behind
  ! (init_input_file src/init.F90:176-228, init_daf_bc :439-511, metrics of
  ! mg_metrics, wall distance of the grid file) on the first call.
  subroutine desalm_f_setup()
    use global
    use global_param
    use global_app
    use global_mpi
    type(desalm_params_t) :: p
    integer :: d, nall, n1, lv
    integer(c_int) :: rc
    if (c_associated(desalm_f_ctx)) return
    if (nproc /= 1 .or. nzone /= 1) then
       write (*, '(a)') 'libdesalm drives one zone on one rank (nproc = 1, nzone = 1)'
       call mpi_abort(mpi_comm_world, 1, ierr)
    end if
    p%ng = ng
    p%img = img(1,1); p%jmg = jmg(1,1); p%kmg = kmg(1,1)
    p%icrs = icrs(1); p%jcrs = jcrs(1); p%kcrs = kcrs(1)
    p%iter = 0
    do lv = 1, min(ng, 6)
       p%iter(lv) = iter(lv)
    end do
    p%itm = itm(1)
    p%it_min = it_min
    p%er_min = er_min; p%eo_min = eo_min
    p%ren = ren; p%beta = beta; p%delti = delti
    p%unsteady = merge(1, 0, unsteady)
    p%des = merge(1, 0, des)
    p%itk = itk
    p%precision = DESALM_PREC_FP64
    p%cfl1 = cfl1(1); p%vnn1 = vnn1(1); p%cfl2 = cfl2(1); p%vnn2 = vnn2(1)
    p%pdiss_coef = pdiss_coef
    do d = 1, 3
       p%ep(d, 1) = ep(1, d, 1)
       p%ep(d, 2) = ep(2, d, 1)
    end do
    p%btype = btype(:, 1)
    p%bdir = bdir(:, 1)
    p%ifix = ifix; p%jfix = jfix; p%kfix = kfix
    p%reserved = 0
    rc = desalm_init(p, 0_c_int, desalm_f_ctx)
    call desalm_f_check(rc, 'desalm_init')
    nall = le_idx_b(ng)
    n1 = le_idx_b(1)
    rc = desalm_set_metrics(desalm_f_ctx, real(csi(:, 1:nall), c_double), real(eta(:, 1:nall), c_double), &
                            real(zet(:, 1:nall), c_double), real(aj(1:nall), c_double))
    call desalm_f_check(rc, 'desalm_set_metrics')
    ! coordinates for the inflow faces' pressure extrapolation and the case hooks
    rc = desalm_set_coordinates(desalm_f_ctx, real(x(1:n1), c_double), real(y(1:n1), c_double), &
                                real(z(1:n1), c_double))
    call desalm_f_check(rc, 'desalm_set_coordinates')
    rc = desalm_set_wall_distance(desalm_f_ctx, real(wd(1:n1), c_double), real(dtev(1:n1), c_double))
    call desalm_f_check(rc, 'desalm_set_wall_distance')
  end subroutine desalm_f_setup

end module desalm_f
