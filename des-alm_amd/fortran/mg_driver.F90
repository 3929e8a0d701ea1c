! mg_driver.F90 -- drop-in replacement of the reference's `subroutine
! mg_driver` (src/mg_driver.F90:2-323): same name, no arguments, same module
! state in and out.  The pseudo-time inner iteration (qn/qnm1 shift, FAS
! V-cycles of DAF smoothing, SA/DES update, fine-grid BCs, convergence test
! of mg_ck_conver :335-405) runs on the GPU through libdesalm.so.
!
! In:  q, qn, qnm1, xnut of the fine level (module global), the blade
!      angles of module global_turb, everything desalm_f_setup reads on the
!      first call.
! Out: q, qn, qnm1, xnut of the fine level; itc, itmax, nstop; one line per
!      pseudo-iteration appended to 'conver' on the root rank, in the format
!      of mg_ck_conver (:363).
! Coarse-level entries of the module arrays are not written back: every
! V-cycle rebuilds them from the fine level (mg_driver.F90:55-58, mg_nlevel).
subroutine mg_driver
  use, intrinsic :: iso_c_binding
  use global
  use global_param
  use global_mpi
  use desalm_f
  implicit none

  integer :: n1, ipi, u
  integer(c_int) :: rc, itc_c
  real(c_double), allocatable :: qb(:,:), qnb(:,:), qnm1b(:,:), xb(:), resid(:,:)

  call desalm_f_setup()

  nstop = ng
  itc = 0
  itmax = itm(1)

  n1 = le_idx_b(1)
  allocate (qb(5, n1), qnb(5, n1), qnm1b(5, n1), xb(n1), resid(5, max(1, itm(1))))
  qb = real(q(1:5, 1:n1), c_double)
  qnb = real(qn(1:5, 1:n1), c_double)
  qnm1b = real(qnm1(1:5, 1:n1), c_double)
  xb = real(xnut(1:n1), c_double)
  rc = desalm_set_state(desalm_f_ctx, qb, qnb, qnm1b, xb)
  call desalm_f_check(rc, 'desalm_set_state')
  call desalm_f_angles()

  rc = desalm_mg_driver(desalm_f_ctx, itc_c, resid)
  call desalm_f_check(rc, 'desalm_mg_driver')

  rc = desalm_get_state(desalm_f_ctx, qb, qnb, qnm1b, xb)
  call desalm_f_check(rc, 'desalm_get_state')
  q(1:5, 1:n1) = real(qb, rdf)
  qn(1:5, 1:n1) = real(qnb, rdf)
  qnm1(1:5, 1:n1) = real(qnm1b, rdf)
  xnut(1:n1) = real(xb, rdf)
  itc = itc_c

  if (myid == root) then
     open (newunit = u, file = 'conver', position = 'append', form = 'formatted')
     do ipi = 1, itc
        write (unit = u, fmt = '(i5.5,1x,6(g13.6,1x))') ipi, real(resid(:, ipi), rdf)
     end do
     close (unit = u)
  end if

  deallocate (qb, qnb, qnm1b, xb, resid)
end subroutine mg_driver
