! desalm_ref_harness.F90 -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
!
! Drives the reference's own Fortran (mg_driver, solver_daf, des_eddy,
! mg_metrics, mg_inject, compiled from /root/reference/src by build_ref.sh)
! on one rank and one zone, so that the C oracle and the HIP path can be
! compared against the real thing.  It replaces what init.F90 does from its
! input files: the module state below is filled with the same formulas
! (index maps src/init.F90:820-980, init_mg :1025-1124, init_daf_bc
! :439-511, level sizes :254-260) from a small binary input file.
!
! usage: desalm_ref_harness <input.bin> <output.bin>
!
! input  (stream, native endian):
!   int32 hdr(40): mode, unused, img, jmg, kmg, icrs, jcrs, kcrs, iter(1:3),
!                  itm, it_min, unsteady, des, itk, btype(1:6), bdir(1:6),
!                  ifix, jfix, kfix, nsteps, daf_level, recalc, calc_pk, nd, nt,
!                  rotate (1: before every mg_driver call of mode 1 the blade
!                  angles advance as the physical time loop does,
!                  src/main.F90:42-49, dtheta of src/init.F90:1447), (pad)
!   real64 rhd(20): er_min, eo_min, ren, beta, delti, cfl1, vnn1, cfl2, vnn2,
!                   pdiss, ep(1,1:3), ep(2,1:3), (pad)
!   real64 x(n1), y(n1), z(n1), wd(n1), q(5,n1), qn(5,n1), qnm1(5,n1), xnut(n1)
!   then, when hdr(36) = nd > 0, an overset donor table of the fine level
!   (global_osg.F90:18-28, zone 1 to zone 1 on rank 0):
!   int32 hosts(3,nd), real64 coef(3,nd), int32 nodes(3,nd)
!   then, when hdr(37) = nt > 0, the actuator-line tables of module
!   global_turb (as init_lines builds them, src/init.F90:1156-1710):
!   int32 th(8,nt): nblades, nelems, nangles, itur1, itur2, n_disk, n_ker, 0
!   real64 length_scale, vel_infty, (tur_rad, hub_rad, tsr, angle,
!          dtheta_interp) per turbine
!   per turbine: int32 node_indexes(nelems,nangles,3),
!   real64 node_indexes_pos(nelems,nangles,27), real64 disk_dat(jmg,kmg,8),
!   int32 polar_before(jmg,kmg), polar_after(jmg,kmg) (airfoil files
!   'pNNN.txt', NNN = index + 1, in the working directory),
!   int32 disk_nodes(3,n_disk), ker_nodes(3,n_ker), real64 ker_pos(3,n_ker),
!   ker_dat(3,n_ker) (disk_dat_ker components 1, 2, 4)
! modes: 1 = nsteps calls of mg_driver, 2 = one solver_daf on the fine level,
!        3 = one des_eddy, 4 = metrics only, 5 = one mg_bintp_mom,
!        6 = one mg_bintp_tur
! output (stream): int32 nall, n1; real64 q(5,n1), qn(5,n1), qnm1(5,n1),
!   xnut(n1), rh(4,n1), csi(3,nall), eta(3,nall), zet(3,nall), aj(nall),
!   dtev(nall)
program desalm_ref_harness
  use global
  use global_param
  use global_app
  use global_mpi
  use global_osg
  use global_turb
  implicit none

  integer :: hdr(40)
  real (kind = rdf) :: rhd(20)
  character (len=512) :: fin, fout
  integer :: mode, nsteps, n, i, j, k, l, m, ib, nall, n1, istep, nd, nt, t, c
  integer, allocatable :: ibuf(:,:), th(:,:), i3(:,:,:), i2(:,:)
  real (kind = rdf), allocatable :: r2(:,:), r3(:,:,:)
  real (kind = rdf) :: rt(2)
  integer :: la, lb, u
  integer (kind = 8) :: c0, c1, crate
  real (kind = rdf), allocatable :: buf(:)

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)

  call mpi_init(ierr)
  myid = 0
  nproc = 1
  myzone = 1
  myleft = mpi_proc_null; myright = mpi_proc_null
  myback = mpi_proc_null; myfront = mpi_proc_null
  mydown = mpi_proc_null; myup = mpi_proc_null

  open (newunit = u, file = trim(fin), access = 'stream', form = 'unformatted', status = 'old')
  read (u) hdr
  read (u) rhd

  mode = hdr(1)
  img(1,1) = hdr(3); jmg(1,1) = hdr(4); kmg(1,1) = hdr(5)
  icrs(1) = hdr(6); jcrs(1) = hdr(7); kcrs(1) = hdr(8)
  iter(:) = 0
  do n = 1, min(ng, 3)
     iter(n) = hdr(8+n)
  end do
  itm(:) = hdr(12)
  it_min = hdr(13)
  unsteady = hdr(14) /= 0
  des = hdr(15) /= 0
  itk = hdr(16)
  nsteps = hdr(32)
  ns = 1
  turbulence = .true.
  turbines = .false.
  nlinc = .false.
  rfgt = .false.

  er_min = rhd(1); eo_min = rhd(2); ren = rhd(3); beta = rhd(4); delti = rhd(5)
  allocate (ep(2,3,nzone), cfl1(nzone), vnn1(nzone), cfl2(nzone), vnn2(nzone), &
            btype(6,nzone), bdir(6,nzone), nzblanking(nzone), blktype(6,nba_max,nzone), &
            blanking(3,2*nba_max,nzone))
  cfl1(1) = rhd(6); vnn1(1) = rhd(7); cfl2(1) = rhd(8); vnn2(1) = rhd(9)
  pdiss_coef = rhd(10)
  do k = 1, 3
     ep(1,k,1) = rhd(10+k)
     ep(2,k,1) = rhd(13+k)
  end do
  do ib = 1, 6
     btype(ib,1) = hdr(16+ib)
     bdir(ib,1) = hdr(22+ib)
  end do
  ifix = hdr(29); jfix = hdr(30); kfix = hdr(31); nzfix = 1
  e_source = zero
  if (unsteady) e_source = one
  cvin = pt1 / ren
  nzblanking = 0; blanking = 0; blktype = 0
  nblk = 0

  ! level sizes (init.F90:254-260)
  do n = 2, ng
     img(n,1) = (img(n-1,1) + icrs(1)) / 2**icrs(1)
     jmg(n,1) = (jmg(n-1,1) + jcrs(1)) / 2**jcrs(1)
     kmg(n,1) = (kmg(n-1,1) + kcrs(1)) / 2**kcrs(1)
  end do

  call setup_index_maps()
  call setup_mg_arrays()

  ! init_daf_bc (init.F90:439-511)
  pfix_proc = 0
  local_ifix = ifix; local_jfix = jfix; local_kfix = kfix
  allocate (sa(5,6), sb(5,6))
  sa = four_third
  sb = -one_third
  do ib = 1, 6
     if (btype(ib,1) == 1) then
        sa(2:4,ib) = zero
        sb(2:4,ib) = zero
     end if
     if (btype(ib,1) == 2) then
        if (bdir(ib,1) == 1) then
           sa(2,ib) = zero; sb(2,ib) = zero
        else if (bdir(ib,1) == 2) then
           sa(3,ib) = zero; sb(3,ib) = zero
        else if (bdir(ib,1) == 3) then
           sa(4,ib) = zero; sb(4,ib) = zero
        end if
     end if
  end do

  ! overset donor table: donors and fringe nodes both on rank 0, zone 1
  nd = hdr(36)
  allocate (donor(0:0), donee(0:0))
  donor = nd; donee = nd
  allocate (hosts(4, max(nd,1), 0:0), coef(3, max(nd,1), 0:0), nodes(4, max(nd,1), 0:0))
  hosts = 1; coef = zero; nodes = 1

  n1 = img(1,1) * jmg(1,1) * kmg(1,1)
  nall = le_idx_b(ng)
  allocate (buf(5*n1))
  read (u) x(1:n1)
  read (u) y(1:n1)
  read (u) z(1:n1)
  read (u) wd(1:n1)
  read (u) buf(1:5*n1)
  q(1:5,1:n1) = reshape(buf(1:5*n1), (/5, n1/))
  read (u) buf(1:5*n1)
  qn(1:5,1:n1) = reshape(buf(1:5*n1), (/5, n1/))
  read (u) buf(1:5*n1)
  qnm1(1:5,1:n1) = reshape(buf(1:5*n1), (/5, n1/))
  read (u) xnut(1:n1)
  if (nd > 0) then
     allocate (ibuf(3, nd))
     read (u) ibuf
     hosts(1:3, 1:nd, 0) = ibuf
     read (u) coef(1:3, 1:nd, 0)
     read (u) ibuf
     nodes(1:3, 1:nd, 0) = ibuf
     deallocate (ibuf)
  end if
  nt = hdr(37)
  turbines = nt > 0
  if (nt > 0) call read_turbines()
  close (u)

  ! grid on all levels and metrics, as init does (init.F90:88-110)
  call mg_inject(x)
  call mg_inject(y)
  call mg_inject(z)
  call mg_metrics
  call setup_rat()

  n = 1
  la = le_idx_a(1)
  lb = le_idx_b(1)
  select case (mode)
  case (1)
     ! wall time of the mg_driver calls alone, for the CPU baseline
     call system_clock(count = c0, count_rate = crate)
     do istep = 1, nsteps
        if (hdr(38) /= 0 .and. turbines) then
           do t = 1, n_tur
              if (angle(t) > 2 * acos(-1.0_rdf)) angle(t) = angle(t) - 2 * acos(-1.0_rdf)
              angle(t) = angle(t) + TSR(t) * delti * length_scale / tur_rad(t)
           end do
        end if
        call mg_driver
     end do
     call system_clock(count = c1)
     open (newunit = u, file = 'timing', form = 'formatted')
     write (u, '(es24.16)') real(c1 - c0, kind = 8) / real(crate, kind = 8)
     close (u)
  case (2)
     call solver_daf (me, 1, hdr(34), hdr(35), &
          le_ia(n),le_ib(n),le_ja(n),le_jb(n),le_ka(n),le_kb(n), &
          igp(n), jgp(n), kgp(n), dc(n), de(n), dz(n), &
          q(1:4,la:lb), qn(1:4,la:lb), qnm1(1:4,la:lb), &
          csi(1:3,la:lb), eta(1:3,la:lb), zet(1:3,la:lb), aj(la:lb), &
          xnut(la:lb), pk(1:4,la:lb), rh(1:4,la:lb), &
          gi_ia(1),gi_ja(1),gi_ka(1),gi_jb(1),gi_kb(1))
  case (3)
     call des_eddy (le_ia(n),le_ib(n),le_ja(n),le_jb(n),le_ka(n),le_kb(n), &
          igp(n), jgp(n), kgp(n), dc(n), de(n), dz(n), dtev(la:lb), &
          q(1:me,la:lb), qn(me,la:lb), qnm1(me,la:lb), &
          csi(1:3,la:lb), eta(1:3,la:lb), zet(1:3,la:lb), aj(la:lb), wd(la:lb), xnut(la:lb))
  case (5)
     ns = 1
     call harness_bintp(0)
  case (6)
     ns = 1
     call harness_bintp(1)
  case default
  end select

  open (newunit = u, file = trim(fout), access = 'stream', form = 'unformatted', status = 'replace')
  write (u) nall, n1
  write (u) ((q(m,l), m = 1, 5), l = 1, n1)
  write (u) ((qn(m,l), m = 1, 5), l = 1, n1)
  write (u) ((qnm1(m,l), m = 1, 5), l = 1, n1)
  write (u) xnut(1:n1)
  write (u) ((rh(m,l), m = 1, 4), l = 1, n1)
  write (u) ((csi(m,l), m = 1, 3), l = 1, nall)
  write (u) ((eta(m,l), m = 1, 3), l = 1, nall)
  write (u) ((zet(m,l), m = 1, 3), l = 1, nall)
  write (u) aj(1:nall)
  write (u) dtev(1:nall)
  close (u)

  call mpi_finalize(ierr)

contains

  ! the actuator-line tables of module global_turb, with the shapes init
  ! allocates (src/init.F90:56-86, 204-267)
  subroutine read_turbines()
    integer :: mne, mna, mnb, mdk, mkr, jm, km, im
    character (len = 10) :: nm
    n_tur = nt
    allocate (th(8, nt))
    read (u) th
    read (u) rt
    length_scale = rt(1); vel_infty = rt(2)
    allocate (itur1(nt), itur2(nt), jtur1(nt), jtur2(nt), ktur1(nt), ktur2(nt), tur_rad(nt), nelems(nt), &
              nblades(nt), nangles(nt), hub_rad(nt), TSR(nt), angle(nt), dtheta(nt), dtheta_interp(nt), &
              C_T(nt), C_P(nt))
    allocate (r2(5, nt))
    read (u) r2
    do t = 1, nt
       nblades(t) = th(1,t); nelems(t) = th(2,t); nangles(t) = th(3,t)
       itur1(t) = th(4,t); itur2(t) = th(5,t)
       tur_rad(t) = r2(1,t); hub_rad(t) = r2(2,t); TSR(t) = r2(3,t); angle(t) = r2(4,t)
       dtheta_interp(t) = r2(5,t)
    end do
    deallocate (r2)
    mnb = maxval(nblades); mne = maxval(nelems); mna = maxval(nangles)
    mdk = maxval(th(6,:)); mkr = maxval(th(7,:))
    im = img(1,1); jm = jmg(1,1); km = kmg(1,1)
    allocate (node_indexes(nt, mne, mna, 3), node_indexes_pos(nt, mne, mna, 27), &
              forces_lines(nt, mnb, mne, 6), disk_dat(jm, km, 8, nt), disk_dat_ker(im, jm, km, 4, nt), &
              airfoil_before(jm, km, nt), airfoil_after(jm, km, nt), &
              disk_node_coordinates(mdk + 1, 3, nt), disk_node_coordinates_ker(mkr + 1, 3, nt), &
              disk_node_coordinates_pos_ker(mkr + 1, 3, nt))
    node_indexes = 0; node_indexes_pos = zero; forces_lines = zero; disk_dat = zero; disk_dat_ker = zero
    disk_node_coordinates = 0; disk_node_coordinates_ker = 0; disk_node_coordinates_pos_ker = zero
    do t = 1, nt
       allocate (i3(nelems(t), nangles(t), 3), r3(nelems(t), nangles(t), 27))
       read (u) i3
       read (u) r3
       node_indexes(t, 1:nelems(t), 1:nangles(t), :) = i3
       node_indexes_pos(t, 1:nelems(t), 1:nangles(t), :) = r3
       deallocate (i3, r3)
       allocate (r3(jm, km, 8))
       read (u) r3
       disk_dat(:, :, :, t) = r3
       deallocate (r3)
       allocate (i2(jm, km))
       read (u) i2
       do k = 1, km
          do j = 1, jm
             write (nm, '(a,i3.3)') 'p', i2(j,k) + 1
             airfoil_before(j, k, t) = nm
          end do
       end do
       read (u) i2
       do k = 1, km
          do j = 1, jm
             write (nm, '(a,i3.3)') 'p', i2(j,k) + 1
             airfoil_after(j, k, t) = nm
          end do
       end do
       deallocate (i2)
       allocate (i2(3, th(6,t)))
       read (u) i2
       disk_node_coordinates(1, 1, t) = th(6,t)
       do c = 1, th(6,t)
          disk_node_coordinates(c + 1, 1:3, t) = i2(1:3, c)
       end do
       deallocate (i2)
       allocate (i2(3, th(7,t)), r2(3, th(7,t)), r3(3, th(7,t), 1))
       read (u) i2
       read (u) r2
       read (u) r3
       disk_node_coordinates_ker(1, 1, t) = th(7,t)
       do c = 1, th(7,t)
          disk_node_coordinates_ker(c + 1, 1:3, t) = i2(1:3, c)
          disk_node_coordinates_pos_ker(c + 1, 1:3, t) = r2(1:3, c)
          disk_dat_ker(i2(1,c), i2(2,c), i2(3,c), 1, t) = r3(1, c, 1)
          disk_dat_ker(i2(1,c), i2(2,c), i2(3,c), 2, t) = r3(2, c, 1)
          disk_dat_ker(i2(1,c), i2(2,c), i2(3,c), 4, t) = r3(3, c, 1)
       end do
       deallocate (i2, r2, r3)
    end do
  end subroutine read_turbines

  ! index maps of init_decomp_3d for one rank (init.F90:820-980)
  subroutine setup_index_maps()
    integer :: ii, jj, kk
    allocate (igp(ng), jgp(ng), kgp(ng))
    igp = 0; jgp = 0; kgp = 0
    allocate (gi_ia(ng), gi_ib(ng), gi_ja(ng), gi_jb(ng), gi_ka(ng), gi_kb(ng))
    allocate (li_ia(ng), li_ib(ng), li_ja(ng), li_jb(ng), li_ka(ng), li_kb(ng))
    allocate (le_ia(ng), le_ib(ng), le_ja(ng), le_jb(ng), le_ka(ng), le_kb(ng))
    allocate (ge_ia(ng), ge_ib(ng), ge_ja(ng), ge_jb(ng), ge_ka(ng), ge_kb(ng))
    allocate (li_ix(ng), li_jx(ng), li_kx(ng), le_ix(ng), le_jx(ng), le_kx(ng))
    allocate (gi_ix(ng), gi_jx(ng), gi_kx(ng), ge_ix(ng), ge_jx(ng), ge_kx(ng))
    allocate (li_idx_a(ng), li_idx_b(ng), li_idx_mx(ng), le_idx_a(ng), le_idx_b(ng), le_idx_mx(ng))
    allocate (gi_idx_a(ng), gi_idx_b(ng), gi_idx_mx(ng), ge_idx_mx(ng))
    do n = 1, ng
       gi_ia(n) = 1; gi_ib(n) = img(n,1)
       gi_ja(n) = 1; gi_jb(n) = jmg(n,1)
       gi_ka(n) = 1; gi_kb(n) = kmg(n,1)
    end do
    li_ia(:) = 1 ; li_ib(:) = gi_ib(:) - gi_ia(:) + 1
    li_ja(:) = 1 ; li_jb(:) = gi_jb(:) - gi_ja(:) + 1
    li_ka(:) = 1 ; li_kb(:) = gi_kb(:) - gi_ka(:) + 1
    le_ia(:) = 1 - igp(:) ; le_ib(:) = li_ib(:) + igp(:)
    le_ja(:) = 1 - jgp(:) ; le_jb(:) = li_jb(:) + jgp(:)
    le_ka(:) = 1 - kgp(:) ; le_kb(:) = li_kb(:) + kgp(:)
    ge_ia(:) = gi_ia(:) - igp(:); ge_ja(:) = gi_ja(:) - jgp(:); ge_ka(:) = gi_ka(:) - kgp(:)
    ge_ib(:) = gi_ib(:) + igp(:); ge_jb(:) = gi_jb(:) + jgp(:); ge_kb(:) = gi_kb(:) + kgp(:)
    li_ix(:) = li_ib(:) - li_ia(:) + 1; li_jx(:) = li_jb(:) - li_ja(:) + 1; li_kx(:) = li_kb(:) - li_ka(:) + 1
    le_ix(:) = le_ib(:) - le_ia(:) + 1; le_jx(:) = le_jb(:) - le_ja(:) + 1; le_kx(:) = le_kb(:) - le_ka(:) + 1
    gi_ix(:) = gi_ib(:) - gi_ia(:) + 1; gi_jx(:) = gi_jb(:) - gi_ja(:) + 1; gi_kx(:) = gi_kb(:) - gi_ka(:) + 1
    ge_ix(:) = ge_ib(:) - ge_ia(:) + 1; ge_jx(:) = ge_jb(:) - ge_ja(:) + 1; ge_kx(:) = ge_kb(:) - ge_ka(:) + 1
    allocate (li_idx(li_ia(1):li_ib(1),li_ja(1):li_jb(1),li_ka(1):li_kb(1),1:ng), &
              le_idx(le_ia(1):le_ib(1),le_ja(1):le_jb(1),le_ka(1):le_kb(1),1:ng), &
              gi_idx(minval(gi_ia(:)):maxval(gi_ib(:)), minval(gi_ja(:)):maxval(gi_jb(:)), &
                     minval(gi_ka(:)):maxval(gi_kb(:)), 1:ng), &
              gi_2_le_idx(minval(gi_ia(:)):maxval(gi_ib(:)), minval(gi_ja(:)):maxval(gi_jb(:)), &
                          minval(gi_ka(:)):maxval(gi_kb(:)), 1:ng))
    li_idx = 0; le_idx = 0; gi_idx = 0; gi_2_le_idx = 0
    li_idx_a(1) = 1; li_idx_b(1) = li_ix(1) * li_jx(1) * li_kx(1); li_idx_mx(1) = 0
    le_idx_a(1) = 1; le_idx_b(1) = le_ix(1) * le_jx(1) * le_kx(1); le_idx_mx(1) = 0
    gi_idx_a(1) = 1; gi_idx_b(1) = gi_ix(1) * gi_jx(1) * gi_kx(1); gi_idx_mx(1) = 0
    do n = 2, ng
       li_idx_a(n) = li_idx_b(n-1) + 1
       li_idx_b(n) = li_idx_a(n) + li_ix(n) * li_jx(n) * li_kx(n) - 1
       li_idx_mx(n) = li_idx_mx(n-1) + li_ix(n-1) * li_jx(n-1) * li_kx(n-1)
       le_idx_a(n) = le_idx_b(n-1) + 1
       le_idx_b(n) = le_idx_a(n) + le_ix(n) * le_jx(n) * le_kx(n) - 1
       le_idx_mx(n) = le_idx_mx(n-1) + le_ix(n-1) * le_jx(n-1) * le_kx(n-1)
       gi_idx_a(n) = gi_idx_b(n-1) + 1
       gi_idx_b(n) = gi_idx_a(n) + gi_ix(n) * gi_jx(n) * gi_kx(n) - 1
       gi_idx_mx(n) = gi_idx_mx(n-1) + gi_ix(n-1) * gi_jx(n-1) * gi_kx(n-1)
    end do
    l = 0
    do n = 1, ng
       do k = li_ka(n), li_kb(n); do j = li_ja(n), li_jb(n); do i = li_ia(n), li_ib(n)
          l = l + 1; li_idx(i,j,k,n) = l
       end do; end do; end do
    end do
    l = 0
    do n = 1, ng
       do k = le_ka(n), le_kb(n); do j = le_ja(n), le_jb(n); do i = le_ia(n), le_ib(n)
          l = l + 1; le_idx(i,j,k,n) = l
       end do; end do; end do
    end do
    l = 0
    do n = 1, ng
       do k = gi_ka(n), gi_kb(n); do j = gi_ja(n), gi_jb(n); do i = gi_ia(n), gi_ib(n)
          l = l + 1; gi_idx(i,j,k,n) = l
       end do; end do; end do
    end do
    do n = 1, ng
       do k = gi_ka(n), gi_kb(n); do j = gi_ja(n), gi_jb(n); do i = gi_ia(n), gi_ib(n)
          kk = k - gi_ka(n) + 1; jj = j - gi_ja(n) + 1; ii = i - gi_ia(n) + 1
          gi_2_le_idx(i,j,k,n) = le_idx(ii,jj,kk,n)
       end do; end do; end do
    end do
  end subroutine setup_index_maps

  ! allocations of init_mg (init.F90:1025-1124)
  ! The grid ratio of the pressure extrapolation on inflow faces (btype 4),
  ! as init_outlet_bc fills module global_app's rat after init_grid
  ! (src/init.F90:2750-2891): on face b the spacing of the first two node
  ! pairs off the face in direction bdir(b).  Its k faces test btype == 4
  ! .and. btype == 5 (:2855, :2874), which never holds, so they stay as
  ! allocated; here they are zero.
  subroutine setup_rat()
    integer :: ijkmax, b, i, j, k, l0, l1, l2, d
    real (kind = rdf) :: c0, c1, c2
    ijkmax = max(img(1,1), jmg(1,1), kmg(1,1))
    allocate (rat(ijkmax, ijkmax, 6))
    rat = zero
    do b = 1, 4
       if (btype(b,1) /= 4 .and. btype(b,1) /= 5) cycle
       d = bdir(b,1)
       do k = 1, kmg(1,1)
          do j = 1, merge(jmg(1,1), img(1,1), b <= 2)
             if (b == 1) then
                l0 = le_idx(1,j,k,1); l1 = le_idx(2,j,k,1); l2 = le_idx(3,j,k,1)
             else if (b == 2) then
                l0 = le_idx(img(1,1),j,k,1); l1 = le_idx(img(1,1)-1,j,k,1); l2 = le_idx(img(1,1)-2,j,k,1)
             else if (b == 3) then
                l0 = le_idx(j,1,k,1); l1 = le_idx(j,2,k,1); l2 = le_idx(j,3,k,1)
             else
                l0 = le_idx(j,jmg(1,1),k,1); l1 = le_idx(j,jmg(1,1)-1,k,1); l2 = le_idx(j,jmg(1,1)-2,k,1)
             end if
             if (d == 1) then
                c0 = x(l0); c1 = x(l1); c2 = x(l2)
             else if (d == 2) then
                c0 = y(l0); c1 = y(l1); c2 = y(l2)
             else
                c0 = z(l0); c1 = z(l1); c2 = z(l2)
             end if
             if (mod(b, 2) == 1) then
                rat(j,k,b) = (c1 - c0) / (c2 - c1)
             else
                rat(j,k,b) = (c0 - c1) / (c1 - c2)
             end if
          end do
       end do
    end do
  end subroutine setup_rat

  subroutine setup_mg_arrays()
    integer :: la0, lb0, pad
    la0 = le_idx_a(1)
    lb0 = le_idx_b(ng)
    ! rh gets padding on both sides: mg_nlevel's 27-point restriction reads
    ! up to 1+le_ix+le_ix*le_jx nodes outside the fine level at boundary
    ! coarse nodes (src/mg_nlevel.F90:108-143); those results land only in
    ! boundary pk, which mg_driver zeroes (src/mg_driver.F90:134).  Every
    ! access of rh through an assumed-shape dummy (mg_exchng3_2d) is guarded
    ! by a non-null neighbour, so the shifted lower bound is never seen.
    pad = 2 + le_ix(1) + le_ix(1) * le_jx(1)
    allocate (x(la0:lb0), y(la0:lb0), z(la0:lb0), aj(la0:lb0), csi(1:3,la0:lb0), &
              eta(1:3,la0:lb0), zet(1:3,la0:lb0), q(1:me,la0:lb0), qn(1:me,la0:lb0), &
              qnm1(1:me,la0:lb0), qold_mg(1:me,la0:lb0), rh(1:4,la0-pad:lb0+pad), pk(1:4,la0:lb0))
    allocate (xc(la0:lb0), xe(la0:lb0), xz(la0:lb0), yc(la0:lb0), ye(la0:lb0), yz(la0:lb0), &
              zc(la0:lb0), ze(la0:lb0), zz(la0:lb0))
    allocate (wd(la0:lb0), dtev(la0:lb0), xnut(la0:lb0))
    xc = zero; xe = zero; xz = zero; yc = zero; ye = zero; yz = zero; zc = zero; ze = zero; zz = zero
    x = zero; y = zero; z = zero; csi = zero; eta = zero; zet = zero; aj = zero
    q = zero; qn = zero; qnm1 = zero; qold_mg = zero; rh = zero; pk = zero
    wd = zero; dtev = zero; xnut = zero
    allocate (convec(ng), dc(ng), de(ng), dz(ng))
    convec(1) = 'nc_quick'
    convec(2:ng) = 'nc_first'
    do n = 1, ng
       dc(n) = one / two**real(((n - 1) * icrs(1)), kind = rdf)
       de(n) = one / two**real(((n - 1) * jcrs(1)), kind = rdf)
       dz(n) = one / two**real(((n - 1) * kcrs(1)), kind = rdf)
    end do
  end subroutine setup_mg_arrays

end program desalm_ref_harness
