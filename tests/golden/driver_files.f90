! Writes the reference driver's record formats for a small synthetic state
! with Fortran's own I/O library, as golden files for the C writers:
!   efmt.txt         '(6e16.7)' records of edge-case values
!   go2d_00007_00001.dat   the model_write records (gocean2d_io_mod.f90:124-144)
!   tmask_00001.dat        dump_tmask, internal points (model_mod.f90:278-286)
!   tmask_raw_00001.dat    dump_tmask, raw (model_mod.f90:270-275)
! Arrays are (0:W-1, 0:H-1) so that indices match the C layout jj*W + ji.
program driver_files
  use, intrinsic :: ieee_arithmetic
  implicit none
  integer, parameter :: M = 6, N = 5, W = 8, H = 6
  double precision :: ht(0:W-1,0:H-1), sshn(0:W-1,0:H-1), un(0:W-1,0:H-1), vn(0:W-1,0:H-1)
  integer :: tmask(0:W-1,0:H-1)
  double precision, parameter :: dx = 1000.0d0, dy = 1500.0d0
  double precision :: v(20), rtmp1, rtmp2
  integer :: ji, jj, i

  v = (/ 0.0d0, 500.0d0, -500.0d0, 1.5d3, 123456789.0d0, 0.5d0, -1.0d-5, 9.99999996d0, &
         1.0d-120, -2.5d150, 1.23456785d0, 1.23456775d0, 0.99999999d0, 1.0d0, -0.0d0, &
         2.2250738585072014d-308, 1.7976931348623157d308, 1d99, 9.9999999d98, 1d-99 /)
  open(21, file='efmt.txt', status='unknown', action='write')
  do i = 1, 20
     write(21,'(6e16.7)') v(i), v(i) * 3.0d0, -v(i) / 7.0d0, v(i) * 1.0d-3, v(i) + 1.0d0, v(i) * 1.0d5
  end do
  write(21,'(6e16.7)') ieee_value(rtmp1, ieee_quiet_nan), ieee_value(rtmp1, ieee_positive_inf), &
       ieee_value(rtmp1, ieee_negative_inf), 0.0d0, 0.0d0, 0.0d0
  close(21)

  do jj = 0, H-1
     do ji = 0, W-1
        ht(ji,jj) = 100.0d0 + ji + 10.0d0 * jj
        sshn(ji,jj) = 0.01d0 * (ji - jj) + 1.0d-3 / (1 + ji + jj)
        un(ji,jj) = 0.1d0 * ji - 0.05d0 * jj
        vn(ji,jj) = -0.02d0 * ji * jj + 1.0d-7
        tmask(ji,jj) = 0
        if (ji >= 1 .and. ji <= M-2 .and. jj >= 1 .and. jj <= N-2) tmask(ji,jj) = 1
        if (jj == 0 .and. ji < M) tmask(ji,jj) = -1
     end do
  end do

  open(21, file='go2d_00007_00001.dat', status='unknown', action='write')
  do jj = 1, N-2
     do ji = 1, M-2
        rtmp1 = 0.5d0 * (un(ji-1,jj) + un(ji,jj))
        rtmp2 = 0.5d0 * (vn(ji,jj-1) + vn(ji,jj))
        write(21,'(6e16.7)') ji * dx, jj * dy, ht(ji,jj), sshn(ji,jj), rtmp1, rtmp2
     end do
     write(21,*)
  end do
  close(21)

  open(21, file='tmask_00001.dat', status='unknown', action='write')
  do jj = 1, N-2
     do ji = 1, M-2
        write(21,'(2e16.7,1x,I3)') ji * dx, jj * dy, tmask(ji,jj)
     end do
     write(21,*)
  end do
  close(21)

  open(21, file='tmask_raw_00001.dat', status='unknown', action='write')
  do jj = 0, N-1
     do ji = 0, M-1
        write(21,'(3(I4,1x))') ji + 1, jj + 1, tmask(ji,jj)
     end do
     write(21,*)
  end do
  close(21)
end program driver_files
