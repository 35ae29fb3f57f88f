!-----------------------------------------------------------------------
! Snapshot routines with the argument lists of 2D_basic/fileio.f90:16,58
! (fileio_output, fileio_input) and the same file layout, which the
! reference's readers expect: a stream of t, ix, jx, x(ix), y(jx), the 9
! planes of U and (output only) the 4 planes of V, each plane x fastest.
!
! Linked in place of the reference's fileio.f90, they are the points where an
! unchanged driver touches its arrays on the host, so they also carry the
! device mirrors of libomhd's drop-in calls (omhd_mirror_mode, OMHD_MIRROR=1
! in the environment): fileio_output first brings the host copies of U and V
! up to date, fileio_input refreshes the mirror of U after reading it.  With
! mirroring off both are plain host I/O.
!-----------------------------------------------------------------------
subroutine fileio_output(filename, ix, jx, t, x, y, U, V)
  use omhd_iface
  implicit none
  character*256 :: filename
  integer, intent(in) :: ix, jx
  real(8), intent(in) :: t, x(ix), y(jx)
  real(8), intent(in) :: U(ix, jx, 9), V(ix, jx, 4)
  integer :: m, io
  if (omhd_mirror_sync(U, int(ix, c_int64_t) * jx * 9) /= OMHD_OK) call omhd_fail(OMHD_ERR_HIP, 'fileio_output')
  if (omhd_mirror_sync(V, int(ix, c_int64_t) * jx * 4) /= OMHD_OK) call omhd_fail(OMHD_ERR_HIP, 'fileio_output')
  open(newunit=io, file=filename, form='unformatted', access='stream')
  write(io) t, ix, jx, x, y
  do m = 1, 9
     write(io) U(:, :, m)
  enddo
  do m = 1, 4
     write(io) V(:, :, m)
  enddo
  close(io)
end subroutine fileio_output

subroutine fileio_input(filename, ix, jx, t, x, y, U)
  use omhd_iface
  implicit none
  character*256 :: filename
  integer, intent(in) :: ix, jx
  real(8), intent(out) :: t, x(ix), y(jx)
  real(8), intent(out) :: U(ix, jx, 9)
  integer :: m, io, ix0, jx0
  U = 0.d0
  open(newunit=io, file=filename, form='unformatted', access='stream')
  read(io) t, ix0, jx0
  if (ix0 /= ix .or. jx0 /= jx) then   ! the reference's check and message (fileio.f90:79-85)
     write(6,*) 'parameter mismatch'
     write(6,*) ' ix= ', ix, ' ix0= ', ix0
     write(6,*) ' jx= ', jx, ' jx0= ', jx0
     stop
  endif
  read(io) x, y
  do m = 1, 9
     read(io) U(:, :, m)
  enddo
  close(io)
  if (omhd_mirror_update(U, int(ix, c_int64_t) * jx * 9) /= OMHD_OK) call omhd_fail(OMHD_ERR_HIP, 'fileio_input')
end subroutine fileio_input
