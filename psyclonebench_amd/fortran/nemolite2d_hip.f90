! NEMOLite2D driven from Fortran through libnemo_hip.so's C ABI.
!
! The binding below is the one the reference's device-resident PSy layer
! declares for its C step (manual_versions/psykal_c_opencl/
! time_step_mod.f90:11-58: c_invoke_time_step with 25 host arrays, 25
! device-handle slots and the scalars; c_read_from_device), so this program
! shows a Fortran host driving the HIP step through exactly that interface.
! It reads the reference's namelist (&namctl with the defaults of
! common/gocean2d_io_mod.f90:58-69), sets the model up through
! libnemo_setup.so (dl_esm_inf and model_mod are not in this repository),
! steps nit000..nitend and logs the checksums of the reference driver
! (psykal_c_opencl/nemolite2d.f90:116-120).
module nemo_hip_abi
  use iso_c_binding
  implicit none
  interface
     subroutine c_invoke_time_step(ssha_t, sshn_t, sshn_u, sshn_v, hu, hv, un, vn, ua, ht, &
          ssha_u, va, ssha_v, tmask, area_t, area_u, area_v, dx_u, dx_v, dx_t, dy_u, dy_v, &
          dy_t, gphiu, gphiv, &
          ssha_t_d, sshn_t_d, sshn_u_d, sshn_v_d, hu_d, hv_d, un_d, vn_d, ua_d, ht_d, &
          ssha_u_d, va_d, ssha_v_d, tmask_d, area_t_d, area_u_d, area_v_d, dx_u_d, dx_v_d, &
          dx_t_d, dy_u_d, dy_v_d, dy_t_d, gphiu_d, gphiv_d, &
          istp, xstart, xstop, ystart, ystop, width, total_size, rdt, cbfr, visc, omega, &
          d2r, g) bind(C, name="c_invoke_time_step")
       import :: c_double, c_int, c_ptr
       real(c_double), intent(inout), dimension(*) :: ssha_t, sshn_t, sshn_u, sshn_v, hu, &
            hv, un, vn, ua, ht, ssha_u, va, ssha_v, area_t, area_u, area_v, dx_u, dx_v, &
            dx_t, dy_u, dy_v, dy_t, gphiu, gphiv
       integer(c_int), intent(inout), dimension(*) :: tmask
       type(c_ptr), intent(inout) :: ssha_t_d, sshn_t_d, sshn_u_d, sshn_v_d, hu_d, hv_d, &
            un_d, vn_d, ua_d, ht_d, ssha_u_d, va_d, ssha_v_d, tmask_d, area_t_d, area_u_d, &
            area_v_d, dx_u_d, dx_v_d, dx_t_d, dy_u_d, dy_v_d, dy_t_d, gphiu_d, gphiv_d
       integer(c_int), value :: istp, xstart, xstop, ystart, ystop, width, total_size
       real(c_double), value :: rdt, cbfr, visc, omega, d2r, g
     end subroutine c_invoke_time_step

     subroutine c_read_from_device(from, to, startx, starty, nx, ny, blocking) &
          bind(C, name="c_read_from_device")
       import :: c_ptr, c_int, c_bool
       type(c_ptr), value :: from, to
       integer(c_int), value :: startx, starty, nx, ny
       logical(c_bool), value :: blocking
     end subroutine c_read_from_device

     subroutine nemo_grid_dims(m, n, align, w, h) bind(C, name="nemo_grid_dims")
       import :: c_int
       integer(c_int), value :: m, n, align
       integer(c_int), intent(out) :: w, h
     end subroutine nemo_grid_dims

     integer(c_int) function nemo_setup_tmask(mode, m, n, w, h, tmask) &
          bind(C, name="nemo_setup_tmask")
       import :: c_int
       integer(c_int), value :: mode, m, n, w, h
       integer(c_int), intent(out), dimension(*) :: tmask
     end function nemo_setup_tmask

     subroutine nemo_setup_metrics(w, h, dx, dy, gphi, area_t, area_u, area_v, dx_u, dx_v, &
          dx_t, dy_u, dy_v, dy_t, gphiu, gphiv) bind(C, name="nemo_setup_metrics")
       import :: c_int, c_double
       integer(c_int), value :: w, h
       real(c_double), value :: dx, dy, gphi
       real(c_double), intent(out), dimension(*) :: area_t, area_u, area_v, dx_u, dx_v, &
            dx_t, dy_u, dy_v, dy_t, gphiu, gphiv
     end subroutine nemo_setup_metrics

     subroutine nemo_setup_depth_uniform(w, h, depth, ht, hu, hv) &
          bind(C, name="nemo_setup_depth_uniform")
       import :: c_int, c_double
       integer(c_int), value :: w, h
       real(c_double), value :: depth
       real(c_double), intent(out), dimension(*) :: ht, hu, hv
     end subroutine nemo_setup_depth_uniform

     subroutine nemo_initialise(m, n, w, h, area_t, area_u, area_v, sshn_t, sshn_u, sshn_v, &
          un, vn) bind(C, name="nemo_initialise")
       import :: c_int, c_double
       integer(c_int), value :: m, n, w, h
       real(c_double), intent(in), dimension(*) :: area_t, area_u, area_v, sshn_t
       real(c_double), intent(inout), dimension(*) :: sshn_u, sshn_v, un, vn
     end subroutine nemo_initialise

     real(c_double) function nemo_field_checksum(x, w, x0, x1, y0, y1) &
          bind(C, name="nemo_field_checksum")
       import :: c_int, c_double
       real(c_double), intent(in), dimension(*) :: x
       integer(c_int), value :: w, x0, x1, y0, y1
     end function nemo_field_checksum
  end interface
end module nemo_hip_abi

program nemolite2d_hip
  use iso_c_binding
  use nemo_hip_abi
  implicit none
  ! common/physical_params_mod.f90:9-15
  real(c_double), parameter :: pi = 3.1415926535897932_c_double
  real(c_double), parameter :: g = 9.80665_c_double, omega = 7.292116e-05_c_double
  real(c_double), parameter :: d2r = pi / 180.0_c_double
  integer :: jpiglo, jpjglo, jphgr_msh, nit000, nitend, irecord
  real(c_double) :: dx, dy, dep_const, rdt, cbfr, visc
  namelist /namctl/ jpiglo, jpjglo, jphgr_msh, dx, dy, dep_const, nit000, nitend, &
       irecord, rdt, cbfr, visc
  real(c_double), allocatable, target, dimension(:,:) :: ssha_t, sshn_t, sshn_u, sshn_v, &
       hu, hv, un, vn, ua, ht, ssha_u, va, ssha_v, area_t, area_u, area_v, dx_u, dx_v, &
       dx_t, dy_u, dy_v, dy_t, gphiu, gphiv
  integer(c_int), allocatable, target :: tmask(:,:)
  type(c_ptr) :: d(25)
  integer(c_int) :: w, h
  integer :: istp, ios

  jpiglo = 50; jpjglo = 100; jphgr_msh = 1; dx = 1000.0_c_double; dy = 1000.0_c_double
  dep_const = 100.0_c_double; nit000 = 1; nitend = 1000; irecord = 1
  rdt = 10.0_c_double; cbfr = 0.001_c_double; visc = 100.0_c_double
  open(99, file='namelist', status='old', iostat=ios)
  if (ios /= 0) stop 'no namelist file'
  read(99, nml=namctl, iostat=ios)
  if (ios /= 0) stop 'err found in reading namelist file'
  close(99)

  call nemo_grid_dims(jpiglo, jpjglo, 64, w, h)
  allocate(ssha_t(w,h), sshn_t(w,h), sshn_u(w,h), sshn_v(w,h), hu(w,h), hv(w,h), un(w,h), &
       vn(w,h), ua(w,h), ht(w,h), ssha_u(w,h), va(w,h), ssha_v(w,h), area_t(w,h), &
       area_u(w,h), area_v(w,h), dx_u(w,h), dx_v(w,h), dx_t(w,h), dy_u(w,h), dy_v(w,h), &
       dy_t(w,h), gphiu(w,h), gphiv(w,h), tmask(w,h))
  ssha_t = 0; sshn_t = 0; sshn_u = 0; sshn_v = 0; un = 0; vn = 0; ua = 0; va = 0
  ssha_u = 0; ssha_v = 0
  if (nemo_setup_tmask(jphgr_msh, jpiglo, jpjglo, w, h, tmask) /= 0) stop 'jphgr_msh'
  call nemo_setup_metrics(w, h, dx, dy, 50.0_c_double, area_t, area_u, area_v, dx_u, dx_v, &
       dx_t, dy_u, dy_v, dy_t, gphiu, gphiv)
  call nemo_setup_depth_uniform(w, h, dep_const, ht, hu, hv)
  call nemo_initialise(jpiglo, jpjglo, w, h, area_t, area_u, area_v, sshn_t, sshn_u, sshn_v, &
       un, vn)
  write(*, "('Simulation domain = (',I4,':',I4,',',I4,':',I4,')')") 2, jpiglo - 1, 2, jpjglo - 1

  d = c_null_ptr
  do istp = nit000, nitend
     call c_invoke_time_step(ssha_t, sshn_t, sshn_u, sshn_v, hu, hv, un, vn, ua, ht, ssha_u, &
          va, ssha_v, tmask, area_t, area_u, area_v, dx_u, dx_v, dx_t, dy_u, dy_v, dy_t, &
          gphiu, gphiv, d(1), d(2), d(3), d(4), d(5), d(6), d(7), d(8), d(9), d(10), d(11), &
          d(12), d(13), d(14), d(15), d(16), d(17), d(18), d(19), d(20), d(21), d(22), &
          d(23), d(24), d(25), istp, 1, jpiglo - 2, 1, jpjglo - 2, w, w * h, rdt, cbfr, &
          visc, omega, d2r, g)
  end do

  ! slots 9 and 12: ua, va (each shares its buffer with un, vn after a step)
  call c_read_from_device(d(9), c_loc(ua), 0, 0, w, h, .true._c_bool)
  call c_read_from_device(d(12), c_loc(va), 0, 0, w, h, .true._c_bool)
  call c_read_from_device(d(2), c_loc(sshn_t), 0, 0, w, h, .true._c_bool)
  write(*, "('ua checksum = ', E16.8)") nemo_field_checksum(ua, w, 1, jpiglo - 2, 1, jpjglo - 2)
  write(*, "('va checksum = ', E16.8)") nemo_field_checksum(va, w, 1, jpiglo - 2, 1, jpjglo - 2)
  write(*, "('sshn checksum = ', E16.8)") &
       nemo_field_checksum(sshn_t, w, 1, jpiglo - 2, 1, jpjglo - 2)
  write(*, "(A)") 'Simulation finished!!'
end program nemolite2d_hip
