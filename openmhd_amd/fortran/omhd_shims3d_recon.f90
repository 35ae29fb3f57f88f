!-----------------------------------------------------------------------
! Drop-in shims, 3D reconnection: the problem-specific routines of
! 3D_reconnection_gpu (bc.cuf, flux_resistive.cuf) with the same names and
! argument lists, callable from the host and implemented by libomhd.  The
! reference has them as CUDA Fortran (device arrays, <<<>>> launches); a
! driver written as plain Fortran calls them like any other subroutine.
! Linked together with omhd_shims3d.f90.
!
!   bc_for_U  3D_reconnection_gpu/bc.cuf:7     bc_for_F  bc.cuf:103
!   bc_for_G  bc.cuf:147                       bc_for_H  bc.cuf:191
!   eta, flux_resistive  flux_resistive.cuf:12 (eta inline at :70-83)
!   set_dt_resistive     flux_resistive.cuf:2
!
! The resistivity profile eta(x, y) is evaluated here, on the host, with the
! host's libm, once per grid: flux_resistive keeps the three face tables (the
! profile does not vary in z) and rebuilds them when x, y or dx change.  A
! caller with another profile edits eta and the two resistivity constants.
!-----------------------------------------------------------------------
module omhd_recon3d_resistivity
  use omhd_iface, only: OMHD_OK, c_int64_t, omhd_mirror_update
  implicit none
  ! magnetic Reynolds numbers: localized (the peak of eta, 1/Rm1) and
  ! background (1/Rm0); flux_resistive.cuf:6, :41-42
  real(8), parameter :: rm1 = 60.d0, rm0 = 1000.d0
  ! eta at the face centres: (:,:,1) x-faces (x+dx/2, y), (:,:,2) y-faces
  ! (x, y+dx/2), (:,:,3) z-faces (x, y); flux_resistive.cuf:47-68
  real(8), allocatable, save :: eta_face(:,:,:), x_seen(:), y_seen(:)
  real(8), save :: dx_seen = 0.d0
contains
  subroutine eta_tables3(x, y, dx, ix, jx)
    integer, intent(in) :: ix, jx
    real(8), intent(in) :: x(ix), y(jx), dx
    real(8), external :: eta
    integer :: i, j
    if (allocated(x_seen)) then
       if (size(x_seen) == ix .and. size(y_seen) == jx .and. dx_seen == dx) then
          if (all(x_seen == x) .and. all(y_seen == y)) return
       endif
       deallocate(eta_face, x_seen, y_seen)
    endif
    allocate(eta_face(ix, jx, 3), x_seen(ix), y_seen(jx))
    x_seen = x; y_seen = y; dx_seen = dx
    do j = 1, jx
       do i = 1, ix
          eta_face(i, j, 1) = eta(x(i) + 0.5d0 * dx, y(j))
          eta_face(i, j, 2) = eta(x(i), y(j) + 0.5d0 * dx)
          eta_face(i, j, 3) = eta(x(i), y(j))
       enddo
    enddo
    ! a new table may sit where an old one did: refresh a device mirror of it
    if (omhd_mirror_update(eta_face, int(size(eta_face), c_int64_t)) /= OMHD_OK) stop 'omhd_mirror_update'
  end subroutine eta_tables3
end module omhd_recon3d_resistivity

! resistivity at (xx, yy): 1/Rm0 far away, 1/Rm1 on the z axis, a cosh^-2
! profile in the distance from it (capped at 25, where cosh^-2 is below 1e-21)
function eta(xx, yy)
  use omhd_recon3d_resistivity, only: rm0, rm1
  implicit none
  real(8), intent(in) :: xx, yy
  real(8) :: eta, r
  r = min(sqrt(xx**2 + yy**2), 25.d0)
  eta = 1.d0 / rm0 + (rm0 - rm1) / (rm0 * rm1) * cosh(r)**(-2)
end function eta

! explicit diffusion limit at the resistivity peak
subroutine set_dt_resistive(dt, dx, cfl)
  use omhd_recon3d_resistivity, only: rm1
  implicit none
  real(8), intent(inout) :: dt
  real(8), intent(in) :: dx, cfl
  dt = min(dt, 0.5d0 * cfl * rm1 * dx**2)
end subroutine set_dt_resistive

subroutine flux_resistive(F, U, x, y, z, dx, ix, jx, kx, dir)
  use omhd_iface
  use omhd_recon3d_resistivity, only: eta_tables3, eta_face
  implicit none
  integer, intent(in) :: ix, jx, kx, dir
  real(8), intent(inout) :: F(ix, jx, kx, 9)
  real(8), intent(in) :: U(ix, jx, kx, 9), x(ix), y(jx), z(kx), dx
  integer(c_int) :: rc
  if (dir < 1 .or. dir > 3) return   ! no default case in flux_resistive.cuf:47-68
  call eta_tables3(x, y, dx, ix, jx)
  rc = omhd_flux_resistive3d(F, U, eta_face(:, :, dir), 1, dx, ix, jx, kx, dir)
  if (rc /= OMHD_OK) call omhd_fail3(rc, 'flux_resistive')
end subroutine flux_resistive

! Reflecting faces in x and y, each with its own negated slots (bc.cuf:14-76),
! periodic in z (:78-96):
!   x- and x+  conducting walls: mx/vx, mz/vz, By, psi
!   y- (south) symmetry plane:   my/vy, mz/vz, Bx, psi
!   y+ (north) wall:             my/vy, By
subroutine recon3d_sign(s)
  implicit none
  real(8), intent(out) :: s(9, 6)
  s = 1.d0
  s((/1, 3, 7, 9/), 1) = -1.d0
  s((/1, 3, 7, 9/), 2) = -1.d0
  s((/2, 3, 6, 9/), 3) = -1.d0
  s((/2, 7/), 4) = -1.d0
end subroutine recon3d_sign

subroutine bc_for_U(U, ix, jx, kx)
  use omhd_iface
  implicit none
  integer, intent(in) :: ix, jx, kx
  real(8), intent(inout) :: U(ix, jx, kx, 9)
  integer(c_int32_t) :: t(6)
  real(8) :: s(9, 6)
  integer(c_int) :: rc
  t = (/ OMHD_BC_MIRROR, OMHD_BC_MIRROR, OMHD_BC_MIRROR, OMHD_BC_MIRROR, OMHD_BC_PERIODIC, OMHD_BC_PERIODIC /)
  call recon3d_sign(s)
  rc = omhd_bc_u(U, ix, jx, kx, t, s)
  if (rc /= OMHD_OK) call omhd_fail3(rc, 'bc_for_U')
end subroutine bc_for_U

subroutine bc_for_F(VL, VR, ix, jx, kx)
  use omhd_iface
  implicit none
  integer, intent(in) :: ix, jx, kx
  real(8), intent(inout) :: VL(ix, jx, kx, 9), VR(ix, jx, kx, 9)
  real(8) :: s(9, 6)
  integer(c_int) :: rc
  call recon3d_sign(s)
  rc = omhd_bc_face(VL, VR, ix, jx, kx, 1, OMHD_BC_MIRROR, s(:, 1), s(:, 2))
  if (rc /= OMHD_OK) call omhd_fail3(rc, 'bc_for_F')
end subroutine bc_for_F

subroutine bc_for_G(VL, VR, ix, jx, kx)
  use omhd_iface
  implicit none
  integer, intent(in) :: ix, jx, kx
  real(8), intent(inout) :: VL(ix, jx, kx, 9), VR(ix, jx, kx, 9)
  real(8) :: s(9, 6)
  integer(c_int) :: rc
  call recon3d_sign(s)
  rc = omhd_bc_face(VL, VR, ix, jx, kx, 2, OMHD_BC_MIRROR, s(:, 3), s(:, 4))
  if (rc /= OMHD_OK) call omhd_fail3(rc, 'bc_for_G')
end subroutine bc_for_G

subroutine bc_for_H(VL, VR, ix, jx, kx)
  use omhd_iface
  implicit none
  integer, intent(in) :: ix, jx, kx
  real(8), intent(inout) :: VL(ix, jx, kx, 9), VR(ix, jx, kx, 9)
  real(8) :: one(9)
  integer(c_int) :: rc
  one = 1.d0
  rc = omhd_bc_face(VL, VR, ix, jx, kx, 3, OMHD_BC_PERIODIC, one, one)
  if (rc /= OMHD_OK) call omhd_fail3(rc, 'bc_for_H')
end subroutine bc_for_H
