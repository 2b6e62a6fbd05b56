! bind(C) driver around the reference's self-contained Fortran modules (mCommon, GaussIntegration,
! Triangle, singInt), compiled from the reference sources into oracle/_ref/ by oracle/Makefile.
! Test infrastructure: it lets tests compare the oracle's tables and singular integrals with the
! reference's own compiled routines. No reference source is copied here.
module andr_ref_bind
  use iso_c_binding
  use mCommon
  use GaussIntegration
  use Triangle
  use singInt
  implicit none
contains
  ! Triangle_IntegrationInit(n) then copy tipw (src/mod_trikotnik.f90:282-781)
  integer(c_int) function ref_tri_rule(itri, l1, l2, l3, w) bind(C, name="ref_tri_rule")
    integer(c_int), value :: itri
    real(c_double) :: l1(126), l2(126), l3(126), w(126)
    integer :: i
    call Triangle_IntegrationInit(int(itri))
    do i = 1, tipw%n
      l1(i) = tipw%l1(i); l2(i) = tipw%l2(i); l3(i) = tipw%l3(i); w(i) = tipw%w(i)
    end do
    ref_tri_rule = tipw%n
  end function

  ! GaussPosWeights (src/mod_gauss.f90:33-391): 156 packed nodes/weights, 8 rules
  subroutine ref_gauss(gi, ome, ng1, ng2) bind(C, name="ref_gauss")
    real(c_double) :: gi(156), ome(156)
    integer(c_int) :: ng1(8), ng2(8)
    integer :: i
    call GaussPosWeights()
    do i = 1, 156
      gi(i) = gaus%gi(i); ome(i) = gaus%ome(i)
    end do
    do i = 1, 8
      ng1(i) = gaus%ng1(i); ng2(i) = gaus%ng2(i)
    end do
  end subroutine

  ! Triangle_Area (src/mod_trikotnik.f90:786-802)
  real(c_double) function ref_tri_area(x) bind(C, name="ref_tri_area")
    real(c_double) :: x(9)
    real(rk) :: a
    call Triangle_Area(x(1), x(2), x(3), x(4), x(5), x(6), x(7), x(8), x(9), a)
    ref_tri_area = a
  end function

  ! si_tri_lap (src/mod_singInt.f90:127-167)
  real(c_double) function ref_si_tri_lap(x, area, isrc) bind(C, name="ref_si_tri_lap")
    real(c_double) :: x(9)
    real(c_double), value :: area
    integer(c_int), value :: isrc
    real(rk) :: g
    call si_tri_lap(x(1), x(2), x(3), x(4), x(5), x(6), x(7), x(8), x(9), area, g, int(isrc))
    ref_si_tri_lap = g
  end function

  ! si_tri_stk (src/mod_singInt.f90:34-123); g = {Gxx,Gxy,Gxz,Gyy,Gyz,Gzz}
  subroutine ref_si_tri_stk(x, area, isrc, g) bind(C, name="ref_si_tri_stk")
    real(c_double) :: x(9), g(6)
    real(c_double), value :: area
    integer(c_int), value :: isrc
    call si_tri_stk(x(1), x(2), x(3), x(4), x(5), x(6), x(7), x(8), x(9), area, &
                    g(1), g(2), g(3), g(4), g(5), g(6), int(isrc))
  end subroutine
end module
