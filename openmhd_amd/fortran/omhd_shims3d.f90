!-----------------------------------------------------------------------
! Drop-in shims, 3D forms: the reference's 3D step subroutines, same names
! and argument lists (kx after jx; the rk stages take the z flux wH after
! wG), implemented by libomhd on the GPU.  A driver written against
! 3D_basic/main.f90 links against these instead of the reference objects.
! The 2D forms are in omhd_shims.f90; a program links one set or the other,
! as it links either the 2D or the 3D reference routines.
!
!   u2v          3D_basic/u2v.f90:1          limiter      3D_basic/limiter.f90:1
!   flux_solver  3D_basic/flux_solver.f90:1  glm_ss2      3D_basic/glm_ss2.f90:1
!   rk_tvd21/22, rk_std21/22  3D_basic/rk.f90:9,46,81,118
!   set_dt       3D_basic/set_dt.f90:1       bc           3D_basic/bc.f90:1
!-----------------------------------------------------------------------
subroutine u2v(U, V, ix, jx, kx)
  use omhd_iface
  implicit none
  integer, intent(in) :: ix, jx, kx
  real(8), intent(in) :: U(ix, jx, kx, 9)
  real(8), intent(out) :: V(ix, jx, kx, 4)
  integer(c_int) :: rc
  rc = omhd_u2v(U, V, ix, jx, kx)
  if (rc /= OMHD_OK) call omhd_fail3(rc, 'u2v')
end subroutine u2v

subroutine limiter(wk, wL, wR, ix, jx, kx, dir, type)
  use omhd_iface
  implicit none
  integer, intent(in) :: ix, jx, kx, dir, type
  real(8), intent(in) :: wk(ix, jx, kx)
  real(8), intent(inout) :: wL(ix, jx, kx), wR(ix, jx, kx)
  integer(c_int) :: rc
  if (dir < 1 .or. dir > 3) return   ! limiter.f90 has no default case for dir
  rc = omhd_limiter(wk, wL, wR, ix, jx, kx, dir, type)
  if (rc /= OMHD_OK) call omhd_fail3(rc, 'limiter')
end subroutine limiter

subroutine flux_solver(F, VL, VR, ch, ix, jx, kx, dir, type)
  use omhd_iface
  implicit none
  integer, intent(in) :: ix, jx, kx, dir, type
  real(8), intent(out) :: F(ix, jx, kx, 9)
  real(8), intent(in) :: VL(ix, jx, kx, 9), VR(ix, jx, kx, 9)
  real(8), intent(in) :: ch
  integer(c_int) :: rc
  if (type < 0 .or. type > 4) return   ! no default case in flux_solver.f90:84-791: F stays as it is
  if (dir < 1 .or. dir > 3) return   ! nor for dir
  rc = omhd_flux_solver(F, VL, VR, ch, ix, jx, kx, dir, type)
  if (rc /= OMHD_OK) call omhd_fail3(rc, 'flux_solver')
end subroutine flux_solver

subroutine glm_ss2(U, ch, dt, ix, jx, kx)
  use omhd_iface
  implicit none
  integer, intent(in) :: ix, jx, kx
  real(8), intent(inout) :: U(ix, jx, kx, 9)
  real(8), intent(in) :: ch, dt
  integer(c_int) :: rc
  rc = omhd_glm_ss2(U, ch, dt, ix, jx, kx)
  if (rc /= OMHD_OK) call omhd_fail3(rc, 'glm_ss2')
end subroutine glm_ss2

! the four rk stages of 3D_basic/rk.f90 (mode 0..3 of omhd_rk)
subroutine rk_tvd21(wk, wk1, wF, wG, wH, dt, dx, ix, jx, kx)
  use omhd_iface
  implicit none
  integer, intent(in) :: ix, jx, kx
  real(8), intent(inout) :: wk(ix, jx, kx, 9), wk1(ix, jx, kx, 9)
  real(8), intent(in) :: wF(ix, jx, kx, 9), wG(ix, jx, kx, 9), dt, dx
  real(8), intent(in), target :: wH(ix, jx, kx, 9)
  integer(c_int) :: rc
  rc = omhd_rk(0, wk, wk1, wF, wG, c_loc(wH), dt, dx, ix, jx, kx)
  if (rc /= OMHD_OK) call omhd_fail3(rc, 'rk_tvd21')
end subroutine rk_tvd21

subroutine rk_tvd22(wk, wk1, wF, wG, wH, dt, dx, ix, jx, kx)
  use omhd_iface
  implicit none
  integer, intent(in) :: ix, jx, kx
  real(8), intent(inout) :: wk(ix, jx, kx, 9), wk1(ix, jx, kx, 9)
  real(8), intent(in) :: wF(ix, jx, kx, 9), wG(ix, jx, kx, 9), dt, dx
  real(8), intent(in), target :: wH(ix, jx, kx, 9)
  integer(c_int) :: rc
  rc = omhd_rk(1, wk, wk1, wF, wG, c_loc(wH), dt, dx, ix, jx, kx)
  if (rc /= OMHD_OK) call omhd_fail3(rc, 'rk_tvd22')
end subroutine rk_tvd22

subroutine rk_std21(wk, wk1, wF, wG, wH, dt, dx, ix, jx, kx)
  use omhd_iface
  implicit none
  integer, intent(in) :: ix, jx, kx
  real(8), intent(inout) :: wk(ix, jx, kx, 9), wk1(ix, jx, kx, 9)
  real(8), intent(in) :: wF(ix, jx, kx, 9), wG(ix, jx, kx, 9), dt, dx
  real(8), intent(in), target :: wH(ix, jx, kx, 9)
  integer(c_int) :: rc
  rc = omhd_rk(2, wk, wk1, wF, wG, c_loc(wH), dt, dx, ix, jx, kx)
  if (rc /= OMHD_OK) call omhd_fail3(rc, 'rk_std21')
end subroutine rk_std21

subroutine rk_std22(wk, wF, wG, wH, dt, dx, ix, jx, kx)
  use omhd_iface
  implicit none
  integer, intent(in) :: ix, jx, kx
  real(8), intent(inout) :: wk(ix, jx, kx, 9)
  real(8), intent(in) :: wF(ix, jx, kx, 9), wG(ix, jx, kx, 9), dt, dx
  real(8), intent(in), target :: wH(ix, jx, kx, 9)
  real(8) :: unused(1)
  integer(c_int) :: rc
  rc = omhd_rk(3, wk, unused, wF, wG, c_loc(wH), dt, dx, ix, jx, kx)
  if (rc /= OMHD_OK) call omhd_fail3(rc, 'rk_std22')
end subroutine rk_std22

subroutine set_dt(U, V, vmax, dt, dx, cfl, ix, jx, kx)
  use omhd_iface
  implicit none
  integer, intent(in) :: ix, jx, kx
  real(8), intent(in) :: U(ix, jx, kx, 9), V(ix, jx, kx, 4), dx, cfl
  real(8), intent(out) :: vmax, dt
  integer(c_int) :: rc
  rc = omhd_set_dt(U, V, vmax, dt, dx, cfl, ix, jx, kx, 3)
  if (rc /= OMHD_OK) call omhd_fail3(rc, 'set_dt')
end subroutine set_dt

! periodic in x, y and z (3D_basic/bc.f90)
subroutine bc(U, ix, jx, kx)
  use omhd_iface
  implicit none
  integer, intent(in) :: ix, jx, kx
  real(8), intent(inout) :: U(ix, jx, kx, 9)
  integer(c_int32_t) :: t(6)
  real(8) :: s(54)
  integer(c_int) :: rc
  t = OMHD_BC_PERIODIC
  s = 1.d0
  rc = omhd_bc_u(U, ix, jx, kx, t, s)
  if (rc /= OMHD_OK) call omhd_fail3(rc, 'bc')
end subroutine bc

! the 3D routines' stop messages (3D_basic/u2v.f90, set_dt.f90): three indices
subroutine omhd_fail3(rc, where)
  use omhd_iface
  implicit none
  integer(c_int), intent(in) :: rc
  character(*), intent(in) :: where
  type(omhd_error_info) :: e
  integer(c_int) :: dummy
  dummy = omhd_last_error(e)
  select case (rc)
  case (OMHD_ERR_PRESSURE)
     write(6,*) 'negative pressure at ', e%i, e%j, e%k, ' P: ', e%value
  case (OMHD_ERR_DENSITY)
     write(6,*) 'negative density at ', e%i, e%j, e%k, ' rho: ', e%value
  case default
     call omhd_fail(rc, where)
  end select
  stop
end subroutine omhd_fail3
