!-----------------------------------------------------------------------
! main2d_resident: the same runs as main2d, driven through libomhd's
! resident step (state kept in HBM, two fused launches per step) from
! Fortran via iso_c_binding.  Reads and writes the snapshot format of
! 2D_basic/fileio.f90.
!
! usage: main2d_resident (--ot | --kh) N STEPS --input F [--lm L]
!                        [--output F] [--trace F] [--time]
!-----------------------------------------------------------------------
program main2d_resident
  use omhd_iface
  implicit none
  integer :: ix, jx, n, nsteps, lm_type, i, narg, m
  logical :: kh, do_time
  character(len=512) :: arg, fin, fout, ftrace
  real(8) :: t, secs
  real(8), allocatable :: U(:,:,:), V(:,:,:), x(:), y(:), dts(:), chs(:)
  type(omhd_config) :: cfg
  type(c_ptr) :: h
  integer(c_int) :: rc
  integer(c_int32_t) :: done
  integer(8) :: c0, c1, rate
  integer :: ixf, jxf

  kh = .false.; do_time = .false.; lm_type = -1
  fin = ''; fout = ''; ftrace = ''
  narg = command_argument_count()
  if (narg < 3) then
     write(6,*) 'usage: main2d_resident (--ot|--kh) N STEPS --input F [--lm L] [--output F] [--trace F] [--time]'
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
     lm_type = 2
     if (kh) lm_type = 1
  endif
  if (len_trim(fin) == 0) then
     write(6,*) 'main2d_resident needs --input F (a snapshot in the format of 2D_basic/fileio.f90)'
     stop 2
  endif
  ix = n + 2; jx = n + 2
  allocate(U(ix,jx,9), V(ix,jx,4), x(ix), y(jx), dts(nsteps), chs(nsteps))

  open(16, file=trim(fin), form='unformatted', access='stream', status='old')
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
  t = 0.d0

  cfg%ndim = 2; cfg%nx = n; cfg%ny = n; cfg%nz = 1
  cfg%dx = x(2) - x(1); cfg%cfl = 0.4d0
  cfg%lm_type = lm_type; cfg%flux_type = 3; cfg%time_type = 0
  cfg%bc_type = OMHD_BC_PERIODIC
  cfg%bc_sign = 1.d0
  if (kh) then   ! walls in y with the sign table of 2D_KH/bc.f90:17-36
     cfg%bc_type(3:4) = OMHD_BC_MIRROR
     cfg%bc_sign(18 + 2) = -1.d0; cfg%bc_sign(18 + 7) = -1.d0
     cfg%bc_sign(27 + 2) = -1.d0; cfg%bc_sign(27 + 7) = -1.d0
  endif
  cfg%rows_per_wave = 0
  cfg%resistive = 0
  cfg%x0 = 0.d0; cfg%y0 = 0.d0; cfg%eta0 = 0.d0; cfg%eta1 = 0.d0; cfg%eta_w = 1.d0
  cfg%reserved = 0

  rc = omhd_create(cfg, h)
  if (rc /= OMHD_OK) call omhd_fail(rc, 'omhd_create')
  rc = omhd_upload(h, U)
  if (rc /= OMHD_OK) call omhd_fail(rc, 'omhd_upload')
  rc = omhd_step(h, 1, t, dts(1), chs(1), done)
  if (rc /= OMHD_OK) call omhd_fail(rc, 'omhd_step')
  call system_clock(c0, rate)
  if (nsteps > 1) then
     rc = omhd_step(h, nsteps - 1, t, dts(2), chs(2), done)
     if (rc /= OMHD_OK) call omhd_fail(rc, 'omhd_step')
  endif
  call system_clock(c1, rate)
  if (do_time .and. nsteps >= 2) then
     secs = dble(c1 - c0) / dble(rate)
     write(6,'(a,es16.8)') 'seconds_steps_2_to_N ', secs
     write(6,'(a,es16.8)') 'zone_updates_per_s ', dble(n) * dble(n) * dble(nsteps - 1) / secs
  endif
  rc = omhd_download(h, U, V)
  if (rc /= OMHD_OK) call omhd_fail(rc, 'omhd_download')
  rc = omhd_destroy(h)
  if (len_trim(ftrace) > 0) then
     open(17, file=trim(ftrace), form='unformatted', access='stream', status='replace')
     write(17) dts, chs
     close(17)
  endif
  if (len_trim(fout) > 0) then
     open(16, file=trim(fout), form='unformatted', access='stream', status='replace')
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
  endif
end program main2d_resident
