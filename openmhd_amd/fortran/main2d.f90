!-----------------------------------------------------------------------
! main2d: a reference-style 2D driver.  It calls only the reference's step
! subroutines by their reference names, in the order of 2D_basic/main.f90:
! 56-188 (periodic) and 2D_KH/main.f90 (KH walls), so the same source links
! either against the reference routines (the reference CPU build, see
! oracle/Makefile) or against the libomhd shims (omhd_shims.f90, the GPU).
!
! usage: main2d (--ot | --kh) N STEPS [--lm L] [--input F] [--output F]
!               [--trace F] [--time]
!   --ot     Orszag-Tang, N x N, periodic        (2D_basic/model.f90)
!   --kh     Kelvin-Helmholtz layer, N x N, x periodic, y walls
!   --input  initial state in the snapshot format of 2D_basic/fileio.f90
!   --output final state in that format; --trace dt/ch per step (raw f8)
!   --time   time steps 2..STEPS and print zone updates per second
!-----------------------------------------------------------------------
program main2d
  implicit none
  integer, parameter :: mx = 1, my = 2, mz = 3, en = 4, vx = 1, vy = 2, vz = 3, pr = 4
  integer, parameter :: ro = 5, bx = 6, by = 7, bz = 8, ps = 9
  real(8), parameter :: gamma = 5.d0 / 3.d0
  integer :: ix, jx, n, nsteps, lm_type, step, i, j, k, narg
  integer, parameter :: flux_type = 3
  logical :: kh, do_time
  character(len=512) :: arg, fin, fout, ftrace
  real(8) :: dx, cfl, t, dt, ch
  real(8), allocatable :: U(:,:,:), U1(:,:,:), V(:,:,:), VL(:,:,:), VR(:,:,:), F(:,:,:), G(:,:,:)
  real(8), allocatable :: x(:), y(:), dts(:), chs(:)
  integer(8) :: c0, c1, rate
  real(8) :: secs

  kh = .false.; do_time = .false.; lm_type = -1
  fin = ''; fout = ''; ftrace = ''
  cfl = 0.4d0
  narg = command_argument_count()
  if (narg < 3) then
     write(6,*) 'usage: main2d (--ot|--kh) N STEPS [--lm L] [--input F] [--output F] [--trace F] [--time]'
     stop 2
  endif
  call get_command_argument(1, arg); kh = (trim(arg) == '--kh')
  call get_command_argument(2, arg); read(arg, *) n
  call get_command_argument(3, arg); read(arg, *) nsteps
  i = 4
  do while (i <= narg)
     call get_command_argument(i, arg)
     select case (trim(arg))
     case ('--lm');     i = i + 1; call get_command_argument(i, arg); read(arg, *) lm_type
     case ('--input');  i = i + 1; call get_command_argument(i, fin)
     case ('--output'); i = i + 1; call get_command_argument(i, fout)
     case ('--trace');  i = i + 1; call get_command_argument(i, ftrace)
     case ('--time');   do_time = .true.
     end select
     i = i + 1
  enddo
  if (lm_type < 0) then
     if (kh) then
        lm_type = 1   ! minmod
     else
        lm_type = 2   ! MC
     endif
  endif

  ix = n + 2; jx = n + 2
  allocate(U(ix,jx,9), U1(ix,jx,9), V(ix,jx,4), VL(ix,jx,9), VR(ix,jx,9), F(ix,jx,9), G(ix,jx,9))
  allocate(x(ix), y(jx), dts(nsteps), chs(nsteps))
  VL = 0.d0; VR = 0.d0; F = 0.d0; G = 0.d0

  if (len_trim(fin) > 0) then
     call read_snapshot(trim(fin), ix, jx, t, x, y, U)
     dx = x(2) - x(1)
  else if (kh) then
     call init_kh(U, x, y, dx, ix, jx)
  else
     call init_ot(U, x, y, dx, ix, jx)
  endif
  t = 0.d0
  if (kh) then
     call bc_for_U(U, ix, jx)
  else
     call bc(U, ix, jx)
  endif
  U1 = U

  call system_clock(c0, rate)
  do step = 1, nsteps
     if (step == 2) call system_clock(c0, rate)
     call u2v(U, V, ix, jx)
     call set_dt(U, V, ch, dt, dx, cfl, ix, jx)
     call glm_ss2(U, ch, dt, ix, jx)
     do k = 1, 2
        if (k == 1) then
           call sweep(U, 1, F)
           call sweep(U, 2, G)
           call rk_tvd21(U, U1, F, G, dt, dx, ix, jx)
           if (kh) then
              call bc_for_U(U1, ix, jx)
           else
              call bc(U1, ix, jx)
           endif
           call u2v(U1, V, ix, jx)
        else
           call sweep(U1, 1, F)
           call sweep(U1, 2, G)
           call rk_tvd22(U, U1, F, G, dt, dx, ix, jx)
           if (kh) then
              call bc_for_U(U, ix, jx)
           else
              call bc(U, ix, jx)
           endif
        endif
     enddo
     call glm_ss2(U, ch, dt, ix, jx)
     t = t + dt
     dts(step) = dt; chs(step) = ch
  enddo
  call system_clock(c1, rate)

  if (do_time .and. nsteps >= 2) then
     secs = dble(c1 - c0) / dble(rate)
     write(6,'(a,es16.8)') 'seconds_steps_2_to_N ', secs
     write(6,'(a,es16.8)') 'zone_updates_per_s ', dble(n) * dble(n) * dble(nsteps - 1) / secs
  endif
  if (len_trim(ftrace) > 0) then
     open(17, file=trim(ftrace), form='unformatted', access='stream', status='replace')
     write(17) dts, chs
     close(17)
  endif
  if (len_trim(fout) > 0) then
     call u2v(U, V, ix, jx)
     call write_snapshot(trim(fout), ix, jx, t, x, y, U, V)
  endif

contains

  ! 9 limiter calls, the face patch, one flux_solver call (main.f90:91-105)
  subroutine sweep(Us, dir, FX)
    real(8), intent(in) :: Us(ix,jx,9)
    integer, intent(in) :: dir
    real(8), intent(inout) :: FX(ix,jx,9)
    integer :: m
    do m = 1, 4
       call limiter(V(:,:,m), VL(:,:,m), VR(:,:,m), ix, jx, dir, lm_type)
    enddo
    do m = 5, 9
       call limiter(Us(:,:,m), VL(:,:,m), VR(:,:,m), ix, jx, dir, lm_type)
    enddo
    if (dir == 1) then
       if (kh) then
          call bc_for_F(VL, VR, ix, jx)
       else
          VR(ix-1,:,:) = VR(1,:,:)
          VL(1,:,:)    = VL(ix-1,:,:)
       endif
    else
       if (kh) then
          call bc_for_G(VL, VR, ix, jx)
       else
          VL(:,1,:)    = VL(:,jx-1,:)
          VR(:,jx-1,:) = VR(:,1,:)
       endif
    endif
    call flux_solver(FX, VL, VR, ch, ix, jx, dir, flux_type)
  end subroutine sweep

  subroutine prim2cons(W, Uc)
    real(8), intent(in) :: W(9)
    real(8), intent(out) :: Uc(9)
    real(8) :: v2, B2
    v2 = W(vx)**2 + W(vy)**2 + W(vz)**2
    B2 = W(bx)**2 + W(by)**2 + W(bz)**2
    Uc(mx:mz) = W(ro) * W(vx:vz)
    Uc(en) = 0.5d0 * (W(ro) * v2 + B2) + W(pr) / (gamma - 1)
    Uc(ro:ps) = W(ro:ps)
  end subroutine prim2cons

  subroutine init_ot(Uo, xo, yo, dxo, ixo, jxo)
    integer, intent(in) :: ixo, jxo
    real(8), intent(out) :: Uo(ixo,jxo,9), xo(ixo), yo(jxo), dxo
    real(8) :: W(9), pi
    integer :: ii, jj
    pi = 4.d0 * atan(1.d0)
    dxo = 2.d0 * pi / dble(ixo - 2)
    do ii = 1, ixo
       xo(ii) = (dble(ii) - 1.5d0) * dxo
    enddo
    do jj = 1, jxo
       yo(jj) = (dble(jj) - 1.5d0) * dxo
    enddo
    do jj = 1, jxo
       do ii = 1, ixo
          W = 0.d0
          W(ro) = gamma**2; W(pr) = gamma
          W(vx) = -sin(yo(jj)); W(vy) = sin(xo(ii))
          W(bx) = -sin(yo(jj)); W(by) = sin(2.d0 * xo(ii))
          call prim2cons(W, Uo(ii,jj,:))
       enddo
    enddo
  end subroutine init_ot

  ! KH layer: Lx = 40, y in [-20,20], vx = 0.5 tanh(y), rho = 1, B = (0.2,0,1),
  ! p = beta |B|^2 / 2 with beta = 2, small deterministic vy ripple near y = 0
  subroutine init_kh(Uo, xo, yo, dxo, ixo, jxo)
    integer, intent(in) :: ixo, jxo
    real(8), intent(out) :: Uo(ixo,jxo,9), xo(ixo), yo(jxo), dxo
    real(8) :: W(9), pi
    integer :: ii, jj
    pi = 4.d0 * atan(1.d0)
    dxo = 40.d0 / dble(ixo - 2)
    do ii = 1, ixo
       xo(ii) = (dble(ii) - 1.5d0) * dxo
    enddo
    do jj = 1, jxo
       yo(jj) = -0.5d0 * dxo * dble(jxo - 2) + (dble(jj) - 1.5d0) * dxo
    enddo
    do jj = 1, jxo
       do ii = 1, ixo
          W = 0.d0
          W(ro) = 1.d0; W(bx) = 0.2d0; W(bz) = 1.d0
          W(pr) = 2.d0 * (0.2d0**2 + 1.d0) / 2.d0
          W(vx) = 0.5d0 * tanh(yo(jj))
          if (abs(yo(jj)) < 2.d0) W(vy) = 1.d-3 * sin(2.d0 * pi * xo(ii) / 40.d0 * 7.d0 + 3.d0 * yo(jj))
          call prim2cons(W, Uo(ii,jj,:))
       enddo
    enddo
  end subroutine init_kh

  ! snapshot format of 2D_basic/fileio.f90:27-49
  subroutine read_snapshot(fname, ixo, jxo, to, xo, yo, Uo)
    character(*), intent(in) :: fname
    integer, intent(in) :: ixo, jxo
    real(8), intent(out) :: to, xo(ixo), yo(jxo), Uo(ixo,jxo,9)
    integer :: ixf, jxf, m
    open(16, file=fname, form='unformatted', access='stream', status='old')
    read(16) to
    read(16) ixf
    read(16) jxf
    if (ixf /= ixo .or. jxf /= jxo) then
       write(6,*) 'parameter mismatch'
       write(6,*) ' ix= ', ixo, ' file: ', ixf
       write(6,*) ' jx= ', jxo, ' file: ', jxf
       stop
    endif
    read(16) xo
    read(16) yo
    do m = 1, 9
       read(16) Uo(:,:,m)
    enddo
    close(16)
  end subroutine read_snapshot

  subroutine write_snapshot(fname, ixo, jxo, to, xo, yo, Uo, Vo)
    character(*), intent(in) :: fname
    integer, intent(in) :: ixo, jxo
    real(8), intent(in) :: to, xo(ixo), yo(jxo), Uo(ixo,jxo,9), Vo(ixo,jxo,4)
    integer :: m
    open(16, file=fname, form='unformatted', access='stream', status='replace')
    write(16) to
    write(16) ixo
    write(16) jxo
    write(16) xo
    write(16) yo
    do m = 1, 9
       write(16) Uo(:,:,m)
    enddo
    do m = 1, 4
       write(16) Vo(:,:,m)
    enddo
    close(16)
  end subroutine write_snapshot

end program main2d
