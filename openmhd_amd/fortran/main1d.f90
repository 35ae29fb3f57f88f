!-----------------------------------------------------------------------
! main1d: a reference-style 1D shock-tube driver.  It follows the loop of
! 1D_basic/main.f90:38-149 (outflow boundaries, no GLM: ch = 0, the RK
! stages written out in the main program) and calls u2v, set_dt, limiter
! and flux_solver by their reference names, so the same source links either
! against the reference routines (the reference CPU build, see
! oracle/Makefile) or against the libomhd shims (omhd_shims.f90, the GPU).
! The 1D routines of the reference are the 2D_basic ones with jx = 1.
!
! usage: main1d [--nx N] [--tend T] [--cfl C] [--lm L] [--flux F] [--tt T]
!               [--steps MAX] [--output F] [--trace F]
!   defaults: the Brio & Wu (1988) shock tube of 1D_basic/model.f90 on
!   [-0.5, 0.5], N = 300, T = 0.2, C = 0.8, minmod, HLLD, TVD-RK2
!   --output final state in the snapshot format of 2D_basic/fileio.f90
!            (jx = 1, y = 0); --trace dt per step (raw f8)
!-----------------------------------------------------------------------
program main1d
  implicit none
  integer, parameter :: mx = 1, my = 2, mz = 3, en = 4, vx = 1, vy = 2, vz = 3, pr = 4
  integer, parameter :: ro = 5, bx = 6, by = 7, bz = 8, ps = 9
  real(8), parameter :: gamma = 5.d0 / 3.d0
  real(8), parameter :: zero = 0.d0
  integer :: n, ix, jx, lm_type, flux_type, time_type, loop_max, nstep, i, m, narg
  character(len=512) :: arg, fout, ftrace
  real(8) :: tend, cfl, dx, t, dt, dtx, vmax
  real(8), allocatable :: U(:,:,:), U1(:,:,:), V(:,:,:), VL(:,:,:), VR(:,:,:), F(:,:,:)
  real(8), allocatable :: x(:), y(:), dts(:)

  n = 300; tend = 0.2d0; cfl = 0.8d0
  lm_type = 1; flux_type = 3; time_type = 0; loop_max = 30000
  fout = ''; ftrace = ''
  narg = command_argument_count()
  i = 1
  do while (i <= narg)
     call get_command_argument(i, arg)
     select case (trim(arg))
     case ('--nx');     i = i + 1; call get_command_argument(i, arg); read(arg, *) n
     case ('--tend');   i = i + 1; call get_command_argument(i, arg); read(arg, *) tend
     case ('--cfl');    i = i + 1; call get_command_argument(i, arg); read(arg, *) cfl
     case ('--lm');     i = i + 1; call get_command_argument(i, arg); read(arg, *) lm_type
     case ('--flux');   i = i + 1; call get_command_argument(i, arg); read(arg, *) flux_type
     case ('--tt');     i = i + 1; call get_command_argument(i, arg); read(arg, *) time_type
     case ('--steps');  i = i + 1; call get_command_argument(i, arg); read(arg, *) loop_max
     case ('--output'); i = i + 1; call get_command_argument(i, fout)
     case ('--trace');  i = i + 1; call get_command_argument(i, ftrace)
     case default
        write(6,*) 'usage: main1d [--nx N] [--tend T] [--cfl C] [--lm L] [--flux F] [--tt T] ', &
                   '[--steps MAX] [--output F] [--trace F]'
        stop 2
     end select
     i = i + 1
  enddo

  ix = n + 2; jx = 1
  allocate(U(ix,jx,9), U1(ix,jx,9), V(ix,jx,4), VL(ix,jx,9), VR(ix,jx,9), F(ix,jx,9))
  allocate(x(ix), y(jx), dts(loop_max))
  U1 = 0.d0; VL = 0.d0; VR = 0.d0; F = 0.d0

  t = 0.d0
  call brio_wu(U, V, x, y, dx)
  call outflow(U)

  nstep = 0
  do
     call u2v(U, V, ix, jx)
     if (t >= tend) exit
     if (nstep + 1 >= loop_max) exit
     call set_dt(U, V, vmax, dt, dx, cfl, ix, jx)
     dtx = dt / dx

     call sweep(U)
     if (time_type == 0) then
        do i = 2, ix - 1
           U1(i,1,:) = U(i,1,:) + dtx * (F(i-1,1,:) - F(i,1,:))
        enddo
     else
        do i = 2, ix - 1
           U1(i,1,:) = U(i,1,:) + 0.5d0 * dtx * (F(i-1,1,:) - F(i,1,:))
        enddo
     endif
     call outflow(U1)
     call u2v(U1, V, ix, jx)

     call sweep(U1)
     if (time_type == 0) then
        do i = 2, ix - 1
           U(i,1,:) = 0.5d0 * (U(i,1,:) + U1(i,1,:) + dtx * (F(i-1,1,:) - F(i,1,:)))
        enddo
     else
        do i = 2, ix - 1
           U(i,1,:) = U(i,1,:) + dtx * (F(i-1,1,:) - F(i,1,:))
        enddo
     endif
     call outflow(U)
     t = t + dt
     nstep = nstep + 1
     dts(nstep) = dt
  enddo

  write(6,'(a,i8,a,es24.16)') 'steps ', nstep, ' t ', t
  if (len_trim(ftrace) > 0) then
     open(17, file=trim(ftrace), form='unformatted', access='stream', status='replace')
     if (nstep > 0) write(17) dts(1:nstep)
     close(17)
  endif
  if (len_trim(fout) > 0) call write_snapshot(trim(fout))

contains

  ! 9 limiter calls (V(vx..pr), then U(ro..ps)) and one flux_solver call in x
  ! with ch = 0 (1D_basic/main.f90:85-98)
  subroutine sweep(Us)
    real(8), intent(in) :: Us(ix,jx,9)
    do m = 1, 4
       call limiter(V(:,:,m), VL(:,:,m), VR(:,:,m), ix, jx, 1, lm_type)
    enddo
    do m = 5, 9
       call limiter(Us(:,:,m), VL(:,:,m), VR(:,:,m), ix, jx, 1, lm_type)
    enddo
    call flux_solver(F, VL, VR, zero, ix, jx, 1, flux_type)
  end subroutine sweep

  ! zero-gradient ghosts (1D_basic/main.f90:42-43)
  subroutine outflow(A)
    real(8), intent(inout) :: A(ix,jx,9)
    A(ix,:,:) = A(ix-1,:,:)
    A(1,:,:)  = A(2,:,:)
  end subroutine outflow

  ! Brio & Wu (1988) left/right states, the active case of 1D_basic/model.f90
  subroutine brio_wu(Uo, Vo, xo, yo, dxo)
    real(8), intent(out) :: Uo(ix,jx,9), Vo(ix,jx,4), xo(ix), yo(jx), dxo
    real(8) :: f1, v2, B2
    integer :: ii
    dxo = (0.5d0 - (-0.5d0)) / dble(ix - 2)
    xo(1) = -0.5d0 - dxo / 2
    do ii = 2, ix
       xo(ii) = xo(ii-1) + dxo
    enddo
    yo(1) = 0.d0
    do ii = 1, ix
       Vo(ii,1,vx:vz) = 0.d0
       Uo(ii,1,ps) = 0.d0
       Uo(ii,1,bx) = 0.75d0
       Uo(ii,1,bz) = 0.d0
       if (xo(ii) < 0.d0) then
          Uo(ii,1,ro) = 1.d0;     Vo(ii,1,pr) = 1.d0;  Uo(ii,1,by) = 1.d0
       else
          Uo(ii,1,ro) = 0.125d0;  Vo(ii,1,pr) = 0.1d0; Uo(ii,1,by) = -1.d0
       endif
       f1 = 1.d0 / (gamma - 1)
       v2 = Vo(ii,1,vx)**2 + Vo(ii,1,vy)**2 + Vo(ii,1,vz)**2
       B2 = Uo(ii,1,bx)**2 + Uo(ii,1,by)**2 + Uo(ii,1,bz)**2
       Uo(ii,1,mx:mz) = Uo(ii,1,ro) * Vo(ii,1,vx:vz)
       Uo(ii,1,en) = 0.5d0 * (Uo(ii,1,ro) * v2 + B2) + f1 * Vo(ii,1,pr)
    enddo
  end subroutine brio_wu

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

end program main1d
