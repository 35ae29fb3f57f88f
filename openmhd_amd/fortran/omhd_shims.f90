!-----------------------------------------------------------------------
! Drop-in shims: the reference's step subroutines, same names and argument
! lists (2D forms), implemented by libomhd on the GPU.  A driver written
! against the reference (2D_basic/main.f90, 2D_KH/main.f90,
! 2D_reconnection/main.f90) links against these instead of the reference
! objects and runs unchanged.  The problem-specific routines sit in their own
! files, as in the reference: bc_for_U/F/G of 2D_KH in omhd_shims_kh.f90;
! bc_for_U/F/G, flux_resistive and set_dt_resistive of 2D_reconnection in
! omhd_shims_recon.f90.
!
!   u2v          2D_basic/u2v.f90:1          limiter      2D_basic/limiter.f90:1
!   flux_solver  2D_basic/flux_solver.f90:1  glm_ss2      2D_basic/glm_ss2.f90:1
!   rk_tvd21/22, rk_std21/22  2D_basic/rk.f90:9,42,73,106
!   set_dt       2D_basic/set_dt.f90:1       bc           2D_basic/bc.f90:1
!-----------------------------------------------------------------------
subroutine u2v(U, V, ix, jx)
  use omhd_iface
  implicit none
  integer, intent(in) :: ix, jx
  real(8), intent(in) :: U(ix, jx, 9)
  real(8), intent(out) :: V(ix, jx, 4)
  integer(c_int) :: rc
  rc = omhd_u2v(U, V, ix, jx, 1)
  if (rc /= OMHD_OK) call omhd_fail(rc, 'u2v')
end subroutine u2v

subroutine limiter(wk, wL, wR, ix, jx, dir, type)
  use omhd_iface
  implicit none
  integer, intent(in) :: ix, jx, dir, type
  real(8), intent(in) :: wk(ix, jx)
  real(8), intent(inout) :: wL(ix, jx), wR(ix, jx)
  integer(c_int) :: rc
  if (dir /= 1 .and. dir /= 2) return   ! the 2D limiter.f90 acts on dir 1 and 2 only (:30-46)
  rc = omhd_limiter(wk, wL, wR, ix, jx, 1, dir, type)
  if (rc /= OMHD_OK) call omhd_fail(rc, 'limiter')
end subroutine limiter

subroutine flux_solver(F, VL, VR, ch, ix, jx, dir, type)
  use omhd_iface
  implicit none
  integer, intent(in) :: ix, jx, dir, type
  real(8), intent(out) :: F(ix, jx, 9)
  real(8), intent(in) :: VL(ix, jx, 9), VR(ix, jx, 9)
  real(8), intent(in) :: ch
  integer(c_int) :: rc
  if (type < 0 .or. type > 4) return   ! no default case in flux_solver.f90:84-791: F stays as it is
  if (dir /= 1 .and. dir /= 2) return   ! nor for dir (:73-81)
  rc = omhd_flux_solver(F, VL, VR, ch, ix, jx, 1, dir, type)
  if (rc /= OMHD_OK) call omhd_fail(rc, 'flux_solver')
end subroutine flux_solver

subroutine glm_ss2(U, ch, dt, ix, jx)
  use omhd_iface
  implicit none
  integer, intent(in) :: ix, jx
  real(8), intent(inout) :: U(ix, jx, 9)
  real(8), intent(in) :: ch, dt
  integer(c_int) :: rc
  rc = omhd_glm_ss2(U, ch, dt, ix, jx, 1)
  if (rc /= OMHD_OK) call omhd_fail(rc, 'glm_ss2')
end subroutine glm_ss2

subroutine rk_tvd21(wk, wk1, wF, wG, dt, dx, ix, jx)
  use omhd_iface
  implicit none
  integer, intent(in) :: ix, jx
  real(8), intent(inout) :: wk(ix, jx, 9), wk1(ix, jx, 9)
  real(8), intent(in) :: wF(ix, jx, 9), wG(ix, jx, 9), dt, dx
  integer(c_int) :: rc
  rc = omhd_rk(0, wk, wk1, wF, wG, c_null_ptr, dt, dx, ix, jx, 1)
  if (rc /= OMHD_OK) call omhd_fail(rc, 'rk_tvd21')
end subroutine rk_tvd21

subroutine rk_tvd22(wk, wk1, wF, wG, dt, dx, ix, jx)
  use omhd_iface
  implicit none
  integer, intent(in) :: ix, jx
  real(8), intent(inout) :: wk(ix, jx, 9), wk1(ix, jx, 9)
  real(8), intent(in) :: wF(ix, jx, 9), wG(ix, jx, 9), dt, dx
  integer(c_int) :: rc
  rc = omhd_rk(1, wk, wk1, wF, wG, c_null_ptr, dt, dx, ix, jx, 1)
  if (rc /= OMHD_OK) call omhd_fail(rc, 'rk_tvd22')
end subroutine rk_tvd22

subroutine rk_std21(wk, wk1, wF, wG, dt, dx, ix, jx)
  use omhd_iface
  implicit none
  integer, intent(in) :: ix, jx
  real(8), intent(inout) :: wk(ix, jx, 9), wk1(ix, jx, 9)
  real(8), intent(in) :: wF(ix, jx, 9), wG(ix, jx, 9), dt, dx
  integer(c_int) :: rc
  rc = omhd_rk(2, wk, wk1, wF, wG, c_null_ptr, dt, dx, ix, jx, 1)
  if (rc /= OMHD_OK) call omhd_fail(rc, 'rk_std21')
end subroutine rk_std21

subroutine rk_std22(wk, wF, wG, dt, dx, ix, jx)
  use omhd_iface
  implicit none
  integer, intent(in) :: ix, jx
  real(8), intent(inout) :: wk(ix, jx, 9)
  real(8), intent(in) :: wF(ix, jx, 9), wG(ix, jx, 9), dt, dx
  real(8) :: unused(1)
  integer(c_int) :: rc
  rc = omhd_rk(3, wk, unused, wF, wG, c_null_ptr, dt, dx, ix, jx, 1)
  if (rc /= OMHD_OK) call omhd_fail(rc, 'rk_std22')
end subroutine rk_std22

subroutine set_dt(U, V, vmax, dt, dx, cfl, ix, jx)
  use omhd_iface
  implicit none
  integer, intent(in) :: ix, jx
  real(8), intent(in) :: U(ix, jx, 9), V(ix, jx, 4), dx, cfl
  real(8), intent(out) :: vmax, dt
  integer(c_int) :: rc
  rc = omhd_set_dt(U, V, vmax, dt, dx, cfl, ix, jx, 1, 2)
  if (rc /= OMHD_OK) call omhd_fail(rc, 'set_dt')
end subroutine set_dt

subroutine bc(U, ix, jx)
  use omhd_iface
  implicit none
  integer, intent(in) :: ix, jx
  real(8), intent(inout) :: U(ix, jx, 9)
  integer(c_int32_t) :: t(6)
  real(8) :: s(54)
  integer(c_int) :: rc
  t = OMHD_BC_PERIODIC
  s = 1.d0
  rc = omhd_bc_u(U, ix, jx, 1, t, s)
  if (rc /= OMHD_OK) call omhd_fail(rc, 'bc')
end subroutine bc
