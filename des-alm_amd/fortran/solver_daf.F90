! solver_daf.F90 -- drop-in replacement of the reference's `subroutine
! solver_daf` (src/solver_daf.F90:2-16, dummies :80-89): same name, same
! argument list, same intents.  One DAF smoother pass on level
! decide_grid_level runs on the GPU through libdesalm.so, copying the
! level's arrays in and out on every call (the parity entry; mg_driver.F90
! of this directory is the device-resident production path).
!
! me, il..ku, igp..kgp, dc, de, dz and gi_* are accepted for signature
! compatibility; the library holds the level sizes and spacings itself and
! this routine checks that the array bounds agree with them.
subroutine solver_daf (me, decide_grid_level, decide_recalc_rh, &
     decide_calc_pk, il, iu, jl, ju, kl, ku, &
     igp, jgp, kgp, &
     dc, de, dz, &
     q, qn, qnm1, csi, eta, zet, aj, xnut, pk, rh, &
     gi_ia, gi_ja, gi_ka, gi_jb, gi_kb)
  use, intrinsic :: iso_c_binding
  use global, only: rdf, img, jmg, kmg
  use global_mpi, only: mpi_comm_world, ierr
  use desalm_f
  implicit none

  integer :: me
  integer :: decide_grid_level, decide_recalc_rh, decide_calc_pk
  integer :: il, iu, jl, ju, kl, ku
  integer :: igp, jgp, kgp
  real (kind = rdf) :: dc, de, dz
  real (kind = rdf), dimension(1:4,il:iu,jl:ju,kl:ku), intent(inout) :: q
  real (kind = rdf), dimension(1:4,il:iu,jl:ju,kl:ku), intent(in)    :: qn, qnm1
  real (kind = rdf), dimension(1:4,il:iu,jl:ju,kl:ku), intent(out)   :: rh
  real (kind = rdf), dimension(1:4,il:iu,jl:ju,kl:ku), intent(inout) :: pk
  real (kind = rdf), dimension(1:3,il:iu,jl:ju,kl:ku), intent(in) :: csi, eta, zet
  real (kind = rdf), dimension(il:iu,jl:ju,kl:ku)    , intent(in) :: aj
  real (kind = rdf), dimension(il:iu,jl:ju,kl:ku), intent(in) :: xnut
  integer :: gi_ia, gi_ja, gi_ka, gi_jb, gi_kb

  integer :: n
  integer(c_int) :: rc
  real(c_double), allocatable :: qb(:,:,:,:), pkb(:,:,:,:), rhb(:,:,:,:)

  call desalm_f_setup()
  n = decide_grid_level
  if (iu - il + 1 /= img(n,1) .or. ju - jl + 1 /= jmg(n,1) .or. ku - kl + 1 /= kmg(n,1)) then
     write (*, '(a,i0)') 'solver_daf: array bounds do not match the size of level ', n
     call mpi_abort(mpi_comm_world, 1, ierr)
  end if

  allocate (qb(1:4,il:iu,jl:ju,kl:ku), pkb(1:4,il:iu,jl:ju,kl:ku), rhb(1:4,il:iu,jl:ju,kl:ku))
  qb = real(q, c_double)
  pkb = real(pk, c_double)
  rhb = 0.0_c_double
  call desalm_f_angles()
  rc = desalm_solver_daf(desalm_f_ctx, int(n, c_int), int(decide_recalc_rh, c_int), &
                         int(decide_calc_pk, c_int), qb, real(qn, c_double), real(qnm1, c_double), &
                         real(csi, c_double), real(eta, c_double), real(zet, c_double), &
                         real(aj, c_double), real(xnut, c_double), pkb, rhb)
  call desalm_f_check(rc, 'desalm_solver_daf')
  q = real(qb, rdf)
  pk = real(pkb, rdf)
  rh = real(rhb, rdf)
  deallocate (qb, pkb, rhb)
end subroutine solver_daf
