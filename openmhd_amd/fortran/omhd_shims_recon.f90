!-----------------------------------------------------------------------
! Drop-in shims, 2D reconnection: the problem-specific routines of
! 2D_reconnection (bc.f90, flux_resistive.f90), same names and argument
! lists, implemented by libomhd.  Linked together with omhd_shims.f90, as
! 2D_reconnection links its own bc.f90 and flux_resistive.f90 beside the
! 2D_basic routines; omhd_shims_kh.f90 holds the KH routines of the same
! names instead.
!
!   bc_for_U  2D_reconnection/bc.f90:6    bc_for_F  2D_reconnection/bc.f90:54
!   bc_for_G  2D_reconnection/bc.f90:86
!   eta       2D_reconnection/flux_resistive.f90:1
!   set_dt_resistive  2D_reconnection/flux_resistive.f90:18
!   flux_resistive    2D_reconnection/flux_resistive.f90:27
!
! The resistivity profile eta(x, y) is evaluated here, on the host, with the
! host's libm (as the reference evaluates it), once per grid: flux_resistive
! keeps the face tables eta/dx and rebuilds them when x, y or dx change.  A
! caller with another profile edits eta and the two resistivity constants.
!-----------------------------------------------------------------------
module omhd_recon_resistivity
  use omhd_iface, only: OMHD_OK, c_int64_t, omhd_mirror_update
  implicit none
  ! magnetic Reynolds numbers: localized (the peak of eta, 1/Rm1) and
  ! background (1/Rm0); flux_resistive.f90:5-7, :22
  real(8), parameter :: rm1 = 60.d0, rm0 = 1000.d0
  ! face tables eta/dx: (:,:,1) x-faces (i+1/2, j), (:,:,2) y-faces (i, j+1/2)
  real(8), allocatable, save :: eta_dx(:,:,:), x_seen(:), y_seen(:)
  real(8), save :: dx_seen = 0.d0
contains
  subroutine eta_tables(x, y, dx, ix, jx)
    integer, intent(in) :: ix, jx
    real(8), intent(in) :: x(ix), y(jx), dx
    real(8), external :: eta
    real(8) :: odx
    integer :: i, j
    if (allocated(x_seen)) then
       if (size(x_seen) == ix .and. size(y_seen) == jx .and. dx_seen == dx) then
          if (all(x_seen == x) .and. all(y_seen == y)) return
       endif
       deallocate(eta_dx, x_seen, y_seen)
    endif
    allocate(eta_dx(ix, jx, 2), x_seen(ix), y_seen(jx))
    x_seen = x; y_seen = y; dx_seen = dx
    odx = 1.d0 / dx
    do j = 1, jx
       do i = 1, ix
          eta_dx(i, j, 1) = eta(x(i) + 0.5d0 * dx, y(j)) * odx   ! flux_resistive.f90:74
          eta_dx(i, j, 2) = eta(x(i), y(j) + 0.5d0 * dx) * odx   ! :97
       enddo
    enddo
    ! a new table may sit where an old one did: refresh a device mirror of it
    if (omhd_mirror_update(eta_dx, int(size(eta_dx), c_int64_t)) /= OMHD_OK) stop 'omhd_mirror_update'
  end subroutine eta_tables
end module omhd_recon_resistivity

! resistivity at (xx, yy): 1/Rm0 far away, 1/Rm1 at the origin, a cosh^-2
! profile in the radius (capped at 25, where cosh^-2 is below 1e-21)
function eta(xx, yy)
  use omhd_recon_resistivity, only: rm0, rm1
  implicit none
  real(8), intent(in) :: xx, yy
  real(8) :: eta, r
  r = min(sqrt(xx**2 + yy**2), 25.d0)
  eta = 1.d0 / rm0 + (rm0 - rm1) / (rm0 * rm1) * cosh(r)**(-2)
end function eta

! explicit diffusion limit at the resistivity peak
subroutine set_dt_resistive(dt, dx, cfl)
  use omhd_recon_resistivity, only: rm1
  implicit none
  real(8), intent(inout) :: dt
  real(8), intent(in) :: dx, cfl
  dt = min(dt, 0.5d0 * cfl * rm1 * dx**2)
end subroutine set_dt_resistive

subroutine flux_resistive(F, U, x, y, dx, ix, jx, dir)
  use omhd_iface
  use omhd_recon_resistivity, only: eta_tables, eta_dx
  implicit none
  integer, intent(in) :: ix, jx, dir
  real(8), intent(inout) :: F(ix, jx, 9)
  real(8), intent(in) :: U(ix, jx, 9), x(ix), y(jx), dx
  integer(c_int) :: rc
  if (dir /= 1 .and. dir /= 2) return   ! no 2D term in z (flux_resistive.f90:56-57)
  call eta_tables(x, y, dx, ix, jx)
  rc = omhd_flux_resistive(F, U, eta_dx(:, :, dir), ix, jx, dir)
  if (rc /= OMHD_OK) call omhd_fail(rc, 'flux_resistive')
end subroutine flux_resistive

! Reflecting faces, each with its own negated slots (bc.f90):
!   x- and x+  conducting walls: mx/vx, mz/vz, By, psi
!   y- (south) symmetry plane:   my/vy, mz/vz, Bx, psi
!   y+ (north) wall:             my/vy, By
subroutine recon_sign(s)
  implicit none
  real(8), intent(out) :: s(9, 6)
  s = 1.d0
  s((/1, 3, 7, 9/), 1) = -1.d0
  s((/1, 3, 7, 9/), 2) = -1.d0
  s((/2, 3, 6, 9/), 3) = -1.d0
  s((/2, 7/), 4) = -1.d0
end subroutine recon_sign

subroutine bc_for_U(U, ix, jx)
  use omhd_iface
  implicit none
  integer, intent(in) :: ix, jx
  real(8), intent(inout) :: U(ix, jx, 9)
  integer(c_int32_t) :: t(6)
  real(8) :: s(9, 6)
  integer(c_int) :: rc
  t = (/ OMHD_BC_MIRROR, OMHD_BC_MIRROR, OMHD_BC_MIRROR, OMHD_BC_MIRROR, OMHD_BC_PERIODIC, OMHD_BC_PERIODIC /)
  call recon_sign(s)
  rc = omhd_bc_u(U, ix, jx, 1, t, s)
  if (rc /= OMHD_OK) call omhd_fail(rc, 'bc_for_U')
end subroutine bc_for_U

subroutine bc_for_F(VL, VR, ix, jx)
  use omhd_iface
  implicit none
  integer, intent(in) :: ix, jx
  real(8), intent(inout) :: VL(ix, jx, 9), VR(ix, jx, 9)
  real(8) :: s(9, 6)
  integer(c_int) :: rc
  call recon_sign(s)
  rc = omhd_bc_face(VL, VR, ix, jx, 1, 1, OMHD_BC_MIRROR, s(:, 1), s(:, 2))
  if (rc /= OMHD_OK) call omhd_fail(rc, 'bc_for_F')
end subroutine bc_for_F

subroutine bc_for_G(VL, VR, ix, jx)
  use omhd_iface
  implicit none
  integer, intent(in) :: ix, jx
  real(8), intent(inout) :: VL(ix, jx, 9), VR(ix, jx, 9)
  real(8) :: s(9, 6)
  integer(c_int) :: rc
  call recon_sign(s)
  rc = omhd_bc_face(VL, VR, ix, jx, 1, 2, OMHD_BC_MIRROR, s(:, 3), s(:, 4))
  if (rc /= OMHD_OK) call omhd_fail(rc, 'bc_for_G')
end subroutine bc_for_G
