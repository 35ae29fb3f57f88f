!-----------------------------------------------------------------------
! Drop-in shims, Kelvin-Helmholtz boundaries: bc_for_U/F/G of 2D_KH/bc.f90
! (x periodic, reflecting walls in y with my/vy and By negated), implemented
! by libomhd.  Linked together with omhd_shims.f90, as 2D_KH links its own
! bc.f90 beside the 2D_basic routines; omhd_shims_recon.f90 holds the
! reconnection problem's routines of the same names instead.
!
!   bc_for_U  2D_KH/bc.f90:6   bc_for_F  2D_KH/bc.f90:42   bc_for_G  2D_KH/bc.f90:55
!-----------------------------------------------------------------------

! KH walls: my/vy and By change sign across the y faces (2D_KH/bc.f90:17-36)
subroutine kh_wall_sign(s9)
  implicit none
  real(8), intent(out) :: s9(9)
  s9 = 1.d0
  s9(2) = -1.d0
  s9(7) = -1.d0
end subroutine kh_wall_sign

subroutine bc_for_U(U, ix, jx)
  use omhd_iface
  implicit none
  integer, intent(in) :: ix, jx
  real(8), intent(inout) :: U(ix, jx, 9)
  integer(c_int32_t) :: t(6)
  real(8) :: s(9, 6)
  integer(c_int) :: rc
  t = (/ OMHD_BC_PERIODIC, OMHD_BC_PERIODIC, OMHD_BC_MIRROR, OMHD_BC_MIRROR, OMHD_BC_PERIODIC, OMHD_BC_PERIODIC /)
  s = 1.d0
  call kh_wall_sign(s(:, 3))
  call kh_wall_sign(s(:, 4))
  rc = omhd_bc_u(U, ix, jx, 1, t, s)
  if (rc /= OMHD_OK) call omhd_fail(rc, 'bc_for_U')
end subroutine bc_for_U

subroutine bc_for_F(VL, VR, ix, jx)
  use omhd_iface
  implicit none
  integer, intent(in) :: ix, jx
  real(8), intent(inout) :: VL(ix, jx, 9), VR(ix, jx, 9)
  real(8) :: one(9)
  integer(c_int) :: rc
  one = 1.d0
  rc = omhd_bc_face(VL, VR, ix, jx, 1, 1, OMHD_BC_PERIODIC, one, one)
  if (rc /= OMHD_OK) call omhd_fail(rc, 'bc_for_F')
end subroutine bc_for_F

subroutine bc_for_G(VL, VR, ix, jx)
  use omhd_iface
  implicit none
  integer, intent(in) :: ix, jx
  real(8), intent(inout) :: VL(ix, jx, 9), VR(ix, jx, 9)
  real(8) :: s(9)
  integer(c_int) :: rc
  call kh_wall_sign(s)
  rc = omhd_bc_face(VL, VR, ix, jx, 1, 2, OMHD_BC_MIRROR, s, s)
  if (rc /= OMHD_OK) call omhd_fail(rc, 'bc_for_G')
end subroutine bc_for_G
