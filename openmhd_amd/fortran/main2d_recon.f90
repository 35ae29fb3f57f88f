!-----------------------------------------------------------------------
! main2d_recon: a reference-style driver for the 2D resistive reconnection
! problem.  It follows the loop of 2D_reconnection/main.f90:56-188 (conducting
! walls in x, a symmetry plane at the south boundary, a wall at the north,
! the resistive flux after each flux_solver call and the resistive dt limit
! after set_dt) and calls only the reference's routines by their reference
! names, so the same source links either against the reference's
! 2D_basic + 2D_reconnection routines (the reference CPU build, see
! oracle/Makefile) or against the libomhd shims (omhd_shims.f90 +
! omhd_shims_recon.f90, the GPU).
!
! usage: main2d_recon NX NY STEPS --input F [--lm L] [--flux T] [--tt T]
!                     [--cfl C] [--output F] [--trace F] [--time]
!   --input  initial state in the snapshot format of 2D_basic/fileio.f90
!   --output final state in that format; --trace dt/ch per step (raw f8)
!   defaults: minmod limiter, HLLD, TVD-RK2, CFL 0.35 (main.f90:12-19)
!-----------------------------------------------------------------------
program main2d_recon
  implicit none
  integer :: nx, ny, ix, jx, nsteps, lm_type, flux_type, time_type, step, i, m, narg
  logical :: do_time
  character(len=512) :: arg, fin, fout, ftrace
  real(8) :: dx, cfl, t, dt, ch, secs
  real(8), allocatable :: U(:,:,:), U1(:,:,:), V(:,:,:), VL(:,:,:), VR(:,:,:), F(:,:,:), G(:,:,:)
  real(8), allocatable :: x(:), y(:), dts(:), chs(:)
  integer(8) :: c0, c1, rate

  lm_type = 1; flux_type = 3; time_type = 0; cfl = 0.35d0; do_time = .false.
  fin = ''; fout = ''; ftrace = ''
  narg = command_argument_count()
  if (narg < 3) then
     write(6,*) 'usage: main2d_recon NX NY STEPS --input F [--lm L] [--flux T] [--tt T] [--cfl C] ', &
                '[--output F] [--trace F] [--time]'
     stop 2
  endif
  call get_command_argument(1, arg); read(arg, *) nx
  call get_command_argument(2, arg); read(arg, *) ny
  call get_command_argument(3, arg); read(arg, *) nsteps
  i = 4
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
     write(6,*) 'main2d_recon needs --input F (a snapshot in the format of 2D_basic/fileio.f90)'
     stop 2
  endif

  ix = nx + 2; jx = ny + 2
  allocate(U(ix,jx,9), U1(ix,jx,9), V(ix,jx,4), VL(ix,jx,9), VR(ix,jx,9), F(ix,jx,9), G(ix,jx,9))
  allocate(x(ix), y(jx), dts(nsteps), chs(nsteps))
  VL = 0.d0; VR = 0.d0; F = 0.d0; G = 0.d0

  call read_snapshot(trim(fin))
  dx = x(2) - x(1)
  call bc_for_U(U, ix, jx)
  U1 = U

  call system_clock(c0, rate)
  do step = 1, nsteps
     if (step == 2) call system_clock(c0, rate)
     call u2v(U, V, ix, jx)
     call set_dt(U, V, ch, dt, dx, cfl, ix, jx)
     call set_dt_resistive(dt, dx, cfl)
     call glm_ss2(U, ch, dt, ix, jx)
     call sweep(U, 1, F)
     call sweep(U, 2, G)
     if (time_type == 0) then
        call rk_tvd21(U, U1, F, G, dt, dx, ix, jx)
     else
        call rk_std21(U, U1, F, G, dt, dx, ix, jx)
     endif
     call bc_for_U(U1, ix, jx)
     call u2v(U1, V, ix, jx)
     call sweep(U1, 1, F)
     call sweep(U1, 2, G)
     if (time_type == 0) then
        call rk_tvd22(U, U1, F, G, dt, dx, ix, jx)
     else
        call rk_std22(U, F, G, dt, dx, ix, jx)
     endif
     call bc_for_U(U, ix, jx)
     call glm_ss2(U, ch, dt, ix, jx)
     t = t + dt
     dts(step) = dt; chs(step) = ch
  enddo
  call system_clock(c1, rate)

  if (do_time .and. nsteps >= 2) then
     secs = dble(c1 - c0) / dble(rate)
     write(6,'(a,es16.8)') 'seconds_steps_2_to_N ', secs
     write(6,'(a,es16.8)') 'zone_updates_per_s ', dble(nx) * dble(ny) * dble(nsteps - 1) / secs
  endif
  if (len_trim(ftrace) > 0) then
     open(17, file=trim(ftrace), form='unformatted', access='stream', status='replace')
     write(17) dts, chs
     close(17)
  endif
  if (len_trim(fout) > 0) then
     call u2v(U, V, ix, jx)
     call write_snapshot(trim(fout))
  endif

contains

  ! 9 limiter calls, the face patch, flux_solver, then the resistive flux
  ! on the same faces (2D_reconnection/main.f90:88-120)
  subroutine sweep(Us, dir, FX)
    real(8), intent(in) :: Us(ix,jx,9)
    integer, intent(in) :: dir
    real(8), intent(inout) :: FX(ix,jx,9)
    do m = 1, 4
       call limiter(V(:,:,m), VL(:,:,m), VR(:,:,m), ix, jx, dir, lm_type)
    enddo
    do m = 5, 9
       call limiter(Us(:,:,m), VL(:,:,m), VR(:,:,m), ix, jx, dir, lm_type)
    enddo
    if (dir == 1) then
       call bc_for_F(VL, VR, ix, jx)
    else
       call bc_for_G(VL, VR, ix, jx)
    endif
    call flux_solver(FX, VL, VR, ch, ix, jx, dir, flux_type)
    call flux_resistive(FX, Us, x, y, dx, ix, jx, dir)
  end subroutine sweep

  ! snapshot format of 2D_basic/fileio.f90:16-49
  subroutine read_snapshot(fname)
    character(*), intent(in) :: fname
    integer :: ixf, jxf
    open(16, file=fname, form='unformatted', access='stream', status='old')
    read(16) t
    read(16) ixf
    read(16) jxf
    if (ixf /= ix .or. jxf /= jx) then
       write(6,*) 'parameter mismatch'
       stop
    endif
    read(16) x
    read(16) y
    do m = 1, 9
       read(16) U(:,:,m)
    enddo
    close(16)
  end subroutine read_snapshot

  subroutine write_snapshot(fname)
    character(*), intent(in) :: fname
    open(16, file=fname, form='unformatted', access='stream', status='replace')
    write(16) t
    write(16) ix
    write(16) jx
    write(16) x
    write(16) y
    do m = 1, 9
       write(16) U(:,:,m)
    enddo
    do m = 1, 4
       write(16) V(:,:,m)
    enddo
    close(16)
  end subroutine write_snapshot

end program main2d_recon
