!-----------------------------------------------------------------------
! main3d_recon: a reference-style driver for the 3D resistive reconnection
! problem.  It follows the loop of 3D_reconnection_gpu/main.cuf:135-303
! (conducting walls in x, a symmetry plane at the south boundary, a wall at
! the north, periodic in z, the resistive flux after each flux_solver call and
! the resistive dt limit after set_dt) as plain Fortran: every routine is a
! host call by its reference name and 3D argument list, implemented by the
! libomhd shims (omhd_shims3d.f90 + omhd_shims3d_recon.f90, the GPU).  The
! reference has this loop only in CUDA Fortran, so there is no CPU build of
! it; the oracle's 3D step is its check.
!
! usage: main3d_recon NX NY NZ STEPS --input F [--lm L] [--flux T] [--tt T]
!                     [--cfl C] [--output F] [--trace F] [--time]
!   --input  initial state in the snapshot format of 3D_basic/fileio.f90
!   --output final state in that format; --trace dt/ch per step (raw f8)
!   defaults: minmod, HLLD, TVD-RK2, CFL 0.3 (main.cuf:19-27)
!-----------------------------------------------------------------------
program main3d_recon
  implicit none
  integer, parameter :: vx = 1, vy = 2, vz = 3, pr = 4
  integer :: nx, ny, nz, ix, jx, kx, nsteps, lm_type, flux_type, time_type, step, i, m, narg
  logical :: do_time
  character(len=512) :: arg, fin, fout, ftrace
  real(8) :: dx, cfl, t, dt, ch, secs
  real(8), allocatable :: U(:,:,:,:), U1(:,:,:,:), V(:,:,:,:), VL(:,:,:,:), VR(:,:,:,:)
  real(8), allocatable :: F(:,:,:,:), G(:,:,:,:), H(:,:,:,:), x(:), y(:), z(:), dts(:), chs(:)
  integer(8) :: c0, c1, rate

  lm_type = 1; flux_type = 3; time_type = 0; cfl = 0.3d0; do_time = .false.
  fin = ''; fout = ''; ftrace = ''
  narg = command_argument_count()
  if (narg < 4) then
     write(6,*) 'usage: main3d_recon NX NY NZ STEPS --input F [--lm L] [--flux T] [--tt T] [--cfl C] ', &
                '[--output F] [--trace F] [--time]'
     stop 2
  endif
  call get_command_argument(1, arg); read(arg, *) nx
  call get_command_argument(2, arg); read(arg, *) ny
  call get_command_argument(3, arg); read(arg, *) nz
  call get_command_argument(4, arg); read(arg, *) nsteps
  i = 5
  do while (i <= narg)
     call get_command_argument(i, arg)
     select case (trim(arg))
     case ('--lm');     i = i + 1; call get_command_argument(i, arg); read(arg, *) lm_type
     case ('--flux');   i = i + 1; call get_command_argument(i, arg); read(arg, *) flux_type
     case ('--tt');     i = i + 1; call get_command_argument(i, arg); read(arg, *) time_type
     case ('--cfl');    i = i + 1; call get_command_argument(i, arg); read(arg, *) cfl
     case ('--input');  i = i + 1; call get_command_argument(i, fin)
     case ('--output'); i = i + 1; call get_command_argument(i, fout)
     case ('--trace');  i = i + 1; call get_command_argument(i, ftrace)
     case ('--time');   do_time = .true.
     end select
     i = i + 1
  enddo
  if (len_trim(fin) == 0) then
     write(6,*) 'main3d_recon needs --input F (a snapshot in the format of 3D_basic/fileio.f90)'
     stop 2
  endif

  ix = nx + 2; jx = ny + 2; kx = nz + 2
  allocate(U(ix,jx,kx,9), U1(ix,jx,kx,9), V(ix,jx,kx,4), VL(ix,jx,kx,9), VR(ix,jx,kx,9))
  allocate(F(ix,jx,kx,9), G(ix,jx,kx,9), H(ix,jx,kx,9), x(ix), y(jx), z(kx), dts(nsteps), chs(nsteps))
  VL = 0.d0; VR = 0.d0; F = 0.d0; G = 0.d0; H = 0.d0

  call read_snapshot(trim(fin))
  dx = x(2) - x(1)
  call bc_for_U(U, ix, jx, kx)
  U1 = U

  call system_clock(c0, rate)
  do step = 1, nsteps
     if (step == 2) call system_clock(c0, rate)
     call u2v(U, V, ix, jx, kx)
     call set_dt(U, V, ch, dt, dx, cfl, ix, jx, kx)
     call set_dt_resistive(dt, dx, cfl)
     call glm_ss2(U, ch, dt, ix, jx, kx)
     call sweep(U, 1, F)
     call sweep(U, 2, G)
     call sweep(U, 3, H)
     if (time_type == 0) then
        call rk_tvd21(U, U1, F, G, H, dt, dx, ix, jx, kx)
     else
        call rk_std21(U, U1, F, G, H, dt, dx, ix, jx, kx)
     endif
     call bc_for_U(U1, ix, jx, kx)
     call u2v(U1, V, ix, jx, kx)
     call sweep(U1, 1, F)
     call sweep(U1, 2, G)
     call sweep(U1, 3, H)
     if (time_type == 0) then
        call rk_tvd22(U, U1, F, G, H, dt, dx, ix, jx, kx)
     else
        call rk_std22(U, F, G, H, dt, dx, ix, jx, kx)
     endif
     call bc_for_U(U, ix, jx, kx)
     call glm_ss2(U, ch, dt, ix, jx, kx)
     t = t + dt
     dts(step) = dt; chs(step) = ch
  enddo
  call system_clock(c1, rate)

  if (do_time .and. nsteps >= 2) then
     secs = dble(c1 - c0) / dble(rate)
     write(6,'(a,es16.8)') 'seconds_steps_2_to_N ', secs
     write(6,'(a,es16.8)') 'zone_updates_per_s ', dble(nx) * dble(ny) * dble(nz) * dble(nsteps - 1) / secs
  endif
  if (len_trim(ftrace) > 0) then
     open(17, file=trim(ftrace), form='unformatted', access='stream', status='replace')
     write(17) dts, chs
     close(17)
  endif
  if (len_trim(fout) > 0) then
     call u2v(U, V, ix, jx, kx)
     call write_snapshot(trim(fout))
  endif

contains

  ! 9 limiter calls, the face patch, flux_solver, then the resistive flux on
  ! the same faces (main.cuf:160-215)
  subroutine sweep(Us, dir, FX)
    real(8), intent(in) :: Us(ix,jx,kx,9)
    integer, intent(in) :: dir
    real(8), intent(inout) :: FX(ix,jx,kx,9)
    do m = 1, 4
       call limiter(V(:,:,:,m), VL(:,:,:,m), VR(:,:,:,m), ix, jx, kx, dir, lm_type)
    enddo
    do m = 5, 9
       call limiter(Us(:,:,:,m), VL(:,:,:,m), VR(:,:,:,m), ix, jx, kx, dir, lm_type)
    enddo
    select case (dir)
    case (1)
       call bc_for_F(VL, VR, ix, jx, kx)
    case (2)
       call bc_for_G(VL, VR, ix, jx, kx)
    case (3)
       call bc_for_H(VL, VR, ix, jx, kx)
    end select
    call flux_solver(FX, VL, VR, ch, ix, jx, kx, dir, flux_type)
    call flux_resistive(FX, Us, x, y, z, dx, ix, jx, kx, dir)
  end subroutine sweep

  ! snapshot format of 3D_basic/fileio.f90:16-54
  subroutine read_snapshot(fname)
    character(*), intent(in) :: fname
    integer :: ixf, jxf, kxf
    open(16, file=fname, form='unformatted', access='stream', status='old')
    read(16) t
    read(16) ixf
    read(16) jxf
    read(16) kxf
    if (ixf /= ix .or. jxf /= jx .or. kxf /= kx) then
       write(6,*) 'parameter mismatch'
       stop
    endif
    read(16) x
    read(16) y
    read(16) z
    do m = 1, 9
       read(16) U(:,:,:,m)
    enddo
    close(16)
  end subroutine read_snapshot

  subroutine write_snapshot(fname)
    character(*), intent(in) :: fname
    open(16, file=fname, form='unformatted', access='stream', status='replace')
    write(16) t
    write(16) ix
    write(16) jx
    write(16) kx
    write(16) x
    write(16) y
    write(16) z
    do m = 1, 9
       write(16) U(:,:,:,m)
    enddo
    do m = 1, 4
       write(16) V(:,:,:,m)
    enddo
    close(16)
  end subroutine write_snapshot

end program main3d_recon
