!-----------------------------------------------------------------------
! omhd_iface: iso_c_binding interfaces to libomhd (include/omhd.h).
! Fortran callers use the shims in omhd_shims.f90, which keep the
! reference subroutine names; this module is the binding underneath, and
! also what a caller of the resident step (omhd_create/step/...) uses.
!-----------------------------------------------------------------------
module omhd_iface
  use, intrinsic :: iso_c_binding
  implicit none

  integer(c_int), parameter :: OMHD_OK = 0, OMHD_ERR_PRESSURE = 1, OMHD_ERR_DENSITY = 2, &
       OMHD_ERR_DT = 3, OMHD_ERR_DIR = 4, OMHD_ERR_TYPE = 5, OMHD_ERR_EXTENT = 6, &
       OMHD_ERR_HIP = 7, OMHD_ERR_ARG = 8
  integer(c_int32_t), parameter :: OMHD_BC_PERIODIC = 0, OMHD_BC_MIRROR = 1

  type, bind(C) :: omhd_error_info
     integer(c_int32_t) :: code, i, j, k
     real(c_double) :: value
     character(kind=c_char) :: msg(160)
  end type omhd_error_info

  type, bind(C) :: omhd_config
     integer(c_int32_t) :: ndim, nx, ny, nz
     real(c_double) :: dx, cfl
     integer(c_int32_t) :: lm_type, flux_type, time_type
     integer(c_int32_t) :: bc_type(6)
     real(c_double) :: bc_sign(54)
     integer(c_int32_t) :: rows_per_wave, resistive
     real(c_double) :: x0, y0, eta0, eta1, eta_w
     integer(c_int32_t) :: reserved(8)
  end type omhd_config

  interface
     integer(c_int) function omhd_last_error(info) bind(C, name='omhd_last_error')
       import :: c_int, omhd_error_info
       type(omhd_error_info), intent(out) :: info
     end function
     integer(c_int) function omhd_u2v(U, V, ix, jx, kx) bind(C, name='omhd_u2v')
       import :: c_int, c_int32_t, c_double
       real(c_double), intent(in) :: U(*)
       real(c_double), intent(inout) :: V(*)
       integer(c_int32_t), value :: ix, jx, kx
     end function
     integer(c_int) function omhd_limiter(wk, wL, wR, ix, jx, kx, dir, type) bind(C, name='omhd_limiter')
       import :: c_int, c_int32_t, c_double
       real(c_double), intent(in) :: wk(*)
       real(c_double), intent(inout) :: wL(*), wR(*)
       integer(c_int32_t), value :: ix, jx, kx, dir, type
     end function
     integer(c_int) function omhd_flux_solver(F, VL, VR, ch, ix, jx, kx, dir, type) bind(C, name='omhd_flux_solver')
       import :: c_int, c_int32_t, c_double
       real(c_double), intent(inout) :: F(*)
       real(c_double), intent(in) :: VL(*), VR(*)
       real(c_double), value :: ch
       integer(c_int32_t), value :: ix, jx, kx, dir, type
     end function
     integer(c_int) function omhd_glm_ss2(U, ch, dt, ix, jx, kx) bind(C, name='omhd_glm_ss2')
       import :: c_int, c_int32_t, c_double
       real(c_double), intent(inout) :: U(*)
       real(c_double), value :: ch, dt
       integer(c_int32_t), value :: ix, jx, kx
     end function
     integer(c_int) function omhd_rk(mode, wk, wk1, wF, wG, wH, dt, dx, ix, jx, kx) bind(C, name='omhd_rk')
       import :: c_int, c_int32_t, c_double, c_ptr
       integer(c_int32_t), value :: mode
       real(c_double), intent(inout) :: wk(*), wk1(*)
       real(c_double), intent(in) :: wF(*), wG(*)
       type(c_ptr), value :: wH
       real(c_double), value :: dt, dx
       integer(c_int32_t), value :: ix, jx, kx
     end function
     integer(c_int) function omhd_set_dt(U, V, vmax, dt, dx, cfl, ix, jx, kx, ndim) bind(C, name='omhd_set_dt')
       import :: c_int, c_int32_t, c_double
       real(c_double), intent(in) :: U(*), V(*)
       real(c_double), intent(out) :: vmax, dt
       real(c_double), value :: dx, cfl
       integer(c_int32_t), value :: ix, jx, kx, ndim
     end function
     integer(c_int) function omhd_bc_u(U, ix, jx, kx, bc_type, sign) bind(C, name='omhd_bc_u')
       import :: c_int, c_int32_t, c_double
       real(c_double), intent(inout) :: U(*)
       integer(c_int32_t), value :: ix, jx, kx
       integer(c_int32_t), intent(in) :: bc_type(6)
       real(c_double), intent(in) :: sign(54)
     end function
     integer(c_int) function omhd_bc_face(VL, VR, ix, jx, kx, dir, bc_type, sign_lo, sign_hi) &
          bind(C, name='omhd_bc_face')
       import :: c_int, c_int32_t, c_double
       real(c_double), intent(inout) :: VL(*), VR(*)
       integer(c_int32_t), value :: ix, jx, kx, dir, bc_type
       real(c_double), intent(in) :: sign_lo(9), sign_hi(9)
     end function
     integer(c_int) function omhd_flux_resistive(F, U, eta_dx, ix, jx, dir) bind(C, name='omhd_flux_resistive')
       import :: c_int, c_int32_t, c_double
       real(c_double), intent(inout) :: F(*)
       real(c_double), intent(in) :: U(*), eta_dx(*)
       integer(c_int32_t), value :: ix, jx, dir
     end function
     integer(c_int) function omhd_flux_resistive3d(F, U, eta, eta_nz, dx, ix, jx, kx, dir) &
          bind(C, name='omhd_flux_resistive3d')
       import :: c_int, c_int32_t, c_double
       real(c_double), intent(inout) :: F(*)
       real(c_double), intent(in) :: U(*), eta(*)
       integer(c_int32_t), value :: eta_nz, ix, jx, kx, dir
       real(c_double), value :: dx
     end function
     integer(c_int) function omhd_create(cfg, h) bind(C, name='omhd_create')
       import :: c_int, c_ptr, omhd_config
       type(omhd_config), intent(in) :: cfg
       type(c_ptr), intent(out) :: h
     end function
     integer(c_int) function omhd_upload(h, U) bind(C, name='omhd_upload')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(in) :: U(*)
     end function
     integer(c_int) function omhd_step(h, nsteps, t, dt_trace, ch_trace, done) bind(C, name='omhd_step')
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: h
       integer(c_int32_t), value :: nsteps
       real(c_double), intent(inout) :: t
       real(c_double), intent(out) :: dt_trace(*), ch_trace(*)
       integer(c_int32_t), intent(out) :: done
     end function
     integer(c_int) function omhd_download(h, U, V) bind(C, name='omhd_download')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(out) :: U(*), V(*)
     end function
     integer(c_int) function omhd_destroy(h) bind(C, name='omhd_destroy')
       import :: c_int, c_ptr
       type(c_ptr), value :: h
     end function
     integer(c_int) function omhd_mirror_mode(on) bind(C, name='omhd_mirror_mode')
       import :: c_int, c_int32_t
       integer(c_int32_t), value :: on
     end function
     integer(c_int) function omhd_mirror_sync(a, n) bind(C, name='omhd_mirror_sync')
       import :: c_int, c_int64_t, c_double
       real(c_double), intent(in) :: a(*)
       integer(c_int64_t), value :: n
     end function
     integer(c_int) function omhd_mirror_update(a, n) bind(C, name='omhd_mirror_update')
       import :: c_int, c_int64_t, c_double
       real(c_double), intent(in) :: a(*)
       integer(c_int64_t), value :: n
     end function
  end interface

contains

  ! Reproduce the reference's diagnostic for a failed call, then stop as it
  ! does (u2v.f90:34-43, set_dt.f90:55-65, limiter.f90:310-312).
  subroutine omhd_fail(rc, where)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: where
    type(omhd_error_info) :: e
    integer(c_int) :: dummy
    dummy = omhd_last_error(e)
    select case (rc)
    case (OMHD_ERR_PRESSURE)
       write(6,*) 'negative pressure at ', e%i, e%j, ' P: ', e%value
    case (OMHD_ERR_DENSITY)
       write(6,*) 'negative density at ', e%i, e%j, ' rho: ', e%value
    case (OMHD_ERR_DT)
       write(6,*) ' dt is too small : ', e%value, ' < ', 1.0d-7
    case (OMHD_ERR_TYPE)
       write(6,*) 'unknown limiter'
    case default
       write(6,*) 'omhd error ', rc, ' in ', where
    end select
    stop
  end subroutine omhd_fail

end module omhd_iface
