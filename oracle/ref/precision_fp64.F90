! precision_fp64.F90 -- TEST INFRASTRUCTURE (oracle/_ref build only).
!
! FP64 variant of the reference's `precision` module: the same public names
! (the reference's other modules `use` them), with the working real kind
! raised from 6 to 15 decimal digits.  The reference ships FP32
! (src/precision.F90:11); its FP64 behaviour is what the FP64 HIP path is
! compared with.  Default REAL entities (LAPACK sgtsv, sa/sb) are promoted by
! building with -fdefault-real-8.
module precision
  implicit none
  integer, parameter :: rdf = selected_real_kind(p=15)
  integer, parameter :: idf = selected_int_kind(r=9)

  real (kind = rdf), parameter :: zero = 0.0_rdf, one = 1.0_rdf, two = 2.0_rdf, three = 3.0_rdf
  real (kind = rdf), parameter :: four = 4.0_rdf, five = 5.0_rdf, six = 6.0_rdf, seven = 7.0_rdf
  real (kind = rdf), parameter :: eight = 8.0_rdf, nine = 9.0_rdf, ten = 10.0_rdf
  real (kind = rdf), parameter :: sixteen = 16.0_rdf
  real (kind = rdf), parameter :: one_fourth = one / four, pt1 = one / ten, pt25 = one / four
  real (kind = rdf), parameter :: pt5 = one / two, one_half = one / two, one_third = one / three
  real (kind = rdf), parameter :: one_pt_five = three / two, onept5 = three / two
  real (kind = rdf), parameter :: four_third = four / three
  real (kind = rdf), parameter :: pi = 3.141592653589793_rdf
end module precision
