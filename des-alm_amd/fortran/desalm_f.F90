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
     integer(c_int32_t) :: dtau_formula  ! 0: rhs_daf_dtau, the reference's live formula
     integer(c_int32_t) :: reserved(6)
     real(c_double) :: body_force(3)     ! 0: the reference has no body force
  end type desalm_params_t

  ! struct desalm_turbine and struct desalm_polar of include/desalm.h
  type, bind(C), public :: desalm_turbine_t
     integer(c_int32_t) :: nblades, nelems, nangles
     integer(c_int32_t) :: itur1, itur2
     real(c_double) :: tur_rad, hub_rad, tsr
     real(c_double) :: angle
     real(c_double) :: dtheta_interp
     type(c_ptr) :: node_indexes, node_indexes_pos, disk_dat, polar_before, polar_after
     integer(c_int32_t) :: n_disk
     type(c_ptr) :: disk_nodes
     integer(c_int32_t) :: n_ker
     type(c_ptr) :: ker_nodes, ker_pos, ker_dat
  end type desalm_turbine_t

  type, bind(C), public :: desalm_polar_t
     integer(c_int32_t) :: n
     type(c_ptr) :: angle, cl, cd
  end type desalm_polar_t

  integer(c_int), parameter, public :: DESALM_HOOK_CONSERVATION = 1, DESALM_HOOK_WALL_FUNCTIONS = 2

  public :: desalm_init, desalm_set_metrics, desalm_set_coordinates, desalm_set_hooks, &
            desalm_set_wall_distance, desalm_set_state, desalm_set_turbines, desalm_turbine_angle, &
            desalm_get_state, desalm_mg_driver, desalm_solver_daf, desalm_finalize
  public :: desalm_f_ctx, desalm_f_setup, desalm_f_check, desalm_f_angles

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

     integer(c_int) function desalm_set_turbines(ctx, n_tur, turbines, n_polar, polars, length_scale, &
                                                 vel_infty) bind(C, name = 'desalm_set_turbines')
       import :: c_int, c_int32_t, c_ptr, c_double, desalm_turbine_t, desalm_polar_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: n_tur, n_polar
       type(desalm_turbine_t), intent(in) :: turbines(*)
       type(desalm_polar_t), intent(in) :: polars(*)
       real(c_double), value :: length_scale, vel_infty
     end function desalm_set_turbines

     ! blade angle of rotor t (0-based): get when get is not null, then set
     integer(c_int) function desalm_turbine_angle(ctx, t, get, set) bind(C, name = 'desalm_turbine_angle')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: t
       type(c_ptr), value :: get
       real(c_double), intent(in) :: set
     end function desalm_turbine_angle

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
  ! mg_metrics, wall distance of the grid file, the actuator-line tables of
  ! init_lines :1156-1710) on the first call.  The reference's mg_driver runs
  ! its two case hooks, conservation (src/mg_driver.F90:64-75) and the
  ! j-wall wall functions (:218-224), on every call of a one-zone, one-rank
  ! run, so both are switched on; its solver_daf adds the actuator-line
  ! forcing when turbines is set (src/solver_daf.F90:378-381).  Blanking
  ! boxes (src/solver_daf.F90:301-315, src/rhs_convec_quick.F90:163) are not
  ! supported: a deck with nzblanking /= 0 stops here.
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
    p%dtau_formula = 0
    p%reserved = 0
    p%body_force = 0.0_c_double
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
    if (nzblanking(1) /= 0) then
       write (*, '(a)') 'libdesalm does not support blanking boxes (nzblanking /= 0 in the input deck)'
       call mpi_abort(mpi_comm_world, 1, ierr)
    end if
    rc = desalm_set_hooks(desalm_f_ctx, ior(DESALM_HOOK_CONSERVATION, DESALM_HOOK_WALL_FUNCTIONS))
    call desalm_f_check(rc, 'desalm_set_hooks')
    call desalm_f_turbines()
  end subroutine desalm_f_setup

  ! The rotors of module global_turb as init_lines leaves them, and the
  ! airfoil polars that rhs_lines reads per element and pass from the files
  ! '<airfoil name>.txt' (angle of attack, Cl, Cd per line up to the angle
  ! -999, src/rhs_lines.F90:251-330), read once here.
  subroutine desalm_f_turbines()
    use global
    use global_mpi
    use global_turb
    type :: rotor_tables_t
       integer(c_int32_t), allocatable :: ni(:,:,:), pb(:,:), pa(:,:), dn(:,:), kn(:,:)
       real(c_double), allocatable :: np(:,:,:), dd(:,:,:), kp(:,:), kd(:,:)
    end type rotor_tables_t
    type :: polar_tables_t
       real(c_double), allocatable :: a(:), cl(:), cd(:)
    end type polar_tables_t
    type(rotor_tables_t), allocatable, target :: tb(:)
    type(polar_tables_t), allocatable, target :: pl(:)
    type(desalm_turbine_t), allocatable :: tt(:)
    type(desalm_polar_t), allocatable :: pp(:)
    character(len = 10), allocatable :: names(:)
    integer :: t, j, k, c, jm, km, ne, na, nd, nk, npol, ii, jj, kk
    integer(c_int) :: rc
    if (.not. turbines) return
    jm = jmg(1,1); km = kmg(1,1)
    allocate (tb(n_tur), tt(n_tur), names(2 * jm * km * n_tur))
    npol = 0
    do t = 1, n_tur
       ne = nelems(t); na = nangles(t)
       allocate (tb(t)%ni(ne, na, 3), tb(t)%np(ne, na, 27), tb(t)%dd(jm, km, 8), tb(t)%pb(jm, km), &
                 tb(t)%pa(jm, km))
       tb(t)%ni = int(node_indexes(t, 1:ne, 1:na, 1:3), c_int32_t)
       tb(t)%np = real(node_indexes_pos(t, 1:ne, 1:na, 1:27), c_double)
       tb(t)%dd = real(disk_dat(1:jm, 1:km, 1:8, t), c_double)
       do k = 1, km
          do j = 1, jm
             tb(t)%pb(j, k) = polar_index(airfoil_before(j, k, t))
             tb(t)%pa(j, k) = polar_index(airfoil_after(j, k, t))
          end do
       end do
       nd = int(disk_node_coordinates(1, 1, t))
       allocate (tb(t)%dn(3, max(nd, 1)))
       tb(t)%dn = 1
       do c = 1, nd
          tb(t)%dn(1:3, c) = int(disk_node_coordinates(c + 1, 1:3, t), c_int32_t)
       end do
       nk = int(disk_node_coordinates_ker(1, 1, t))
       allocate (tb(t)%kn(3, max(nk, 1)), tb(t)%kp(3, max(nk, 1)), tb(t)%kd(3, max(nk, 1)))
       tb(t)%kn = 1; tb(t)%kp = 0.0_c_double; tb(t)%kd = 0.0_c_double
       do c = 1, nk
          ii = int(disk_node_coordinates_ker(c + 1, 1, t))
          jj = int(disk_node_coordinates_ker(c + 1, 2, t))
          kk = int(disk_node_coordinates_ker(c + 1, 3, t))
          tb(t)%kn(1:3, c) = int([ii, jj, kk], c_int32_t)
          tb(t)%kp(1:3, c) = real(disk_node_coordinates_pos_ker(c + 1, 1:3, t), c_double)
          tb(t)%kd(1:3, c) = real([disk_dat_ker(ii, jj, kk, 1, t), disk_dat_ker(ii, jj, kk, 2, t), &
                                   disk_dat_ker(ii, jj, kk, 4, t)], c_double)
       end do
       tt(t)%nblades = nblades(t); tt(t)%nelems = ne; tt(t)%nangles = na
       tt(t)%itur1 = itur1(t); tt(t)%itur2 = itur2(t)
       tt(t)%tur_rad = tur_rad(t); tt(t)%hub_rad = hub_rad(t); tt(t)%tsr = TSR(t)
       tt(t)%angle = angle(t); tt(t)%dtheta_interp = dtheta_interp(t)
       tt(t)%node_indexes = c_loc(tb(t)%ni); tt(t)%node_indexes_pos = c_loc(tb(t)%np)
       tt(t)%disk_dat = c_loc(tb(t)%dd)
       tt(t)%polar_before = c_loc(tb(t)%pb); tt(t)%polar_after = c_loc(tb(t)%pa)
       tt(t)%n_disk = nd; tt(t)%disk_nodes = c_loc(tb(t)%dn)
       tt(t)%n_ker = nk; tt(t)%ker_nodes = c_loc(tb(t)%kn)
       tt(t)%ker_pos = c_loc(tb(t)%kp); tt(t)%ker_dat = c_loc(tb(t)%kd)
    end do
    if (npol == 0) then
       write (*, '(a)') 'libdesalm: the rotor tables name no airfoil (airfoil_before / airfoil_after are blank)'
       call mpi_abort(mpi_comm_world, 1, ierr)
    end if
    ! nodes without an airfoil name are never an element's node: any polar
    do t = 1, n_tur
       where (tb(t)%pb < 0) tb(t)%pb = 0
       where (tb(t)%pa < 0) tb(t)%pa = 0
    end do
    allocate (pl(npol), pp(npol))
    do c = 1, npol
       call read_polar(names(c), pl(c))
       pp(c)%n = size(pl(c)%a)
       pp(c)%angle = c_loc(pl(c)%a); pp(c)%cl = c_loc(pl(c)%cl); pp(c)%cd = c_loc(pl(c)%cd)
    end do
    rc = desalm_set_turbines(desalm_f_ctx, int(n_tur, c_int32_t), tt, int(npol, c_int32_t), pp, &
                             real(length_scale, c_double), real(vel_infty, c_double))
    call desalm_f_check(rc, 'desalm_set_turbines')
    deallocate (tb, tt, pl, pp, names)

  contains

    ! 0-based index of an airfoil name in names(1:npol), appended when new;
    ! -1 for a blank name
    integer(c_int32_t) function polar_index(nm)
      character(len = *), intent(in) :: nm
      integer :: p
      polar_index = -1
      if (len_trim(nm) == 0) return
      do p = 1, npol
         if (names(p) == nm) then
            polar_index = p - 1
            return
         end if
      end do
      npol = npol + 1
      names(npol) = nm
      polar_index = npol - 1
    end function polar_index

    ! One airfoil file, read as rhs_lines reads it (list-directed, values
    ! of kind rdf, up to the line whose angle is -999)
    subroutine read_polar(nm, pt)
      character(len = *), intent(in) :: nm
      type(polar_tables_t), intent(out) :: pt
      real(kind = rdf) :: a, cl, cd
      integer :: u, ios, n, pass
      open (newunit = u, file = trim(adjustl(nm)) // '.txt', status = 'old', action = 'read', iostat = ios)
      if (ios /= 0) then
         write (*, '(a,a,a)') 'libdesalm: cannot open the airfoil file ', trim(adjustl(nm)), '.txt'
         call mpi_abort(mpi_comm_world, 1, ierr)
      end if
      do pass = 1, 2
         rewind (u)
         n = 0
         do
            read (u, *, iostat = ios) a, cl, cd
            if (ios /= 0) exit
            if (a == -999) exit
            n = n + 1
            if (pass == 2) then
               pt%a(n) = real(a, c_double); pt%cl(n) = real(cl, c_double); pt%cd(n) = real(cd, c_double)
            end if
         end do
         if (pass == 1) allocate (pt%a(max(n, 1)), pt%cl(max(n, 1)), pt%cd(max(n, 1)))
      end do
      close (u)
      if (n == 0) then
         write (*, '(a,a,a)') 'libdesalm: no polar points in ', trim(adjustl(nm)), '.txt'
         call mpi_abort(mpi_comm_world, 1, ierr)
      end if
    end subroutine read_polar

  end subroutine desalm_f_turbines

  ! The blade angles of module global_turb, which the physical time loop
  ! advances before every mg_driver call (src/main.F90:42-49), into the
  ! library's rotors.
  subroutine desalm_f_angles()
    use global_turb
    integer :: t
    integer(c_int) :: rc
    if (.not. turbines) return
    do t = 1, n_tur
       rc = desalm_turbine_angle(desalm_f_ctx, int(t - 1, c_int), c_null_ptr, real(angle(t), c_double))
       call desalm_f_check(rc, 'desalm_turbine_angle')
    end do
  end subroutine desalm_f_angles

end module desalm_f
