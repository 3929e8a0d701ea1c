! hooks_off.F90 -- TEST INFRASTRUCTURE (oracle/_ref build only).
!
! Two case-specific hooks that the reference's mg_driver calls
! unconditionally are switched off for the synthetic configurations by
! linking these empty bodies instead of src/wf_mpi.F90 and
! src/conservation.F90:
!  * wf_eta, the wall-function override "FOR DARGAHI's CASE ONLY"
!    (src/mg_driver.F90:218-224), active on every one-rank zone-1 run;
!  * conservation, the outlet mass-flux rescale (src/mg_driver.F90:64-75),
!    which divides by a zero outlet flux on closed domains
!    (src/conservation.F90:146).
module wf_mpi
  implicit none
  private
  public :: wf_eta
contains
  subroutine wf_eta (jb)
    integer, intent(in) :: jb
  end subroutine wf_eta
end module wf_mpi

subroutine conservation ()
end subroutine conservation
