! NEMOLite2D driven from Fortran through libnemo_hip.so's C ABI.
!
! The binding below is the one the reference's device-resident PSy layer
! declares for its C step (manual_versions/psykal_c_opencl/
! time_step_mod.f90:11-58: c_invoke_time_step with 25 host arrays, 25
! device-handle slots and the scalars; c_read_from_device), so this program
! shows a Fortran host driving the HIP step through exactly that interface.
!
! Module hip_fields runs the field protocol of that PSy layer in this
! program's own types (dl_esm_inf's field_mod and grid_mod are not in this
! repository): each field carries its data, a device-handle slot, a
! data_on_device flag and a read_from_device_c procedure pointer;
! invoke_time_step passes the data and the slots to c_invoke_time_step, and
! after its first call marks the fields device-resident and installs
! c_read_from_device as their read-back (time_step_mod.f90:76-192);
! field_get_data reads a device-resident field back through that pointer
! from its current slot, as dl_esm_inf's get_data does.
!
! The program reads the reference's namelist (&namctl with the defaults of
! common/gocean2d_io_mod.f90:58-69), sets the model up through
! libnemo_setup.so, steps nit000..nitend, logs the ssh checksum every
! irecord steps through field_get_data, and at the end the checksums of the
! reference driver (psykal_c_opencl/nemolite2d.f90:116-120).
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

module hip_fields
  use iso_c_binding
  use nemo_hip_abi
  implicit none

  abstract interface
     subroutine read_from_device_iface(from, to, startx, starty, nx, ny, blocking) bind(C)
       import :: c_ptr, c_int, c_bool
       type(c_ptr), value :: from, to
       integer(c_int), value :: startx, starty, nx, ny
       logical(c_bool), value :: blocking
     end subroutine read_from_device_iface
  end interface

  ! One field: host data (width x height, first index fastest), the device
  ! handle the library keeps in its slot, and the read-back hook.
  type :: hip_field
     real(c_double), allocatable :: data(:,:)
     type(c_ptr) :: device_ptr = c_null_ptr
     logical :: data_on_device = .false.
     procedure(read_from_device_iface), nopass, pointer :: read_from_device_c => null()
  end type hip_field

  ! The grid's arrays and their device slots; internal bounds 1-based.
  type :: hip_grid
     integer :: xstart, xstop, ystart, ystop
     integer(c_int), allocatable :: tmask(:,:)
     real(c_double), allocatable, dimension(:,:) :: area_t, area_u, area_v, dx_u, dx_v, &
          dx_t, dy_u, dy_v, dy_t, gphiu, gphiv
     type(c_ptr) :: tmask_device = c_null_ptr, area_t_device = c_null_ptr, &
          area_u_device = c_null_ptr, area_v_device = c_null_ptr, dx_u_device = c_null_ptr, &
          dx_v_device = c_null_ptr, dx_t_device = c_null_ptr, dy_u_device = c_null_ptr, &
          dy_v_device = c_null_ptr, dy_t_device = c_null_ptr, gphiu_device = c_null_ptr, &
          gphiv_device = c_null_ptr
  end type hip_grid

contains

  subroutine field_alloc(f, w, h)
    type(hip_field), intent(inout) :: f
    integer, intent(in) :: w, h
    allocate(f%data(w, h))
    f%data = 0.0_c_double
  end subroutine field_alloc

  ! Host copy of a field: read back from the device when the data lives
  ! there and a read-back is installed, else the host data as it is.
  subroutine field_get_data(f)
    type(hip_field), intent(inout), target :: f
    if (f%data_on_device .and. associated(f%read_from_device_c)) then
       call f%read_from_device_c(f%device_ptr, c_loc(f%data), 0_c_int, 0_c_int, &
            int(size(f%data, 1), c_int), int(size(f%data, 2), c_int), .true._c_bool)
    end if
  end subroutine field_get_data

  ! One step through the C ABI with the device-resident layer's protocol:
  ! the host data and the device slots of every field and grid array, the
  ! 0-based internal bounds, the row pitch and the total size.
  subroutine invoke_time_step(istp, grid, ssha_t, ssha_u, ssha_v, sshn_t, sshn_u, sshn_v, &
       hu, hv, ht, ua, va, un, vn, rdt, cbfr, visc, omega, d2r, g)
    integer, intent(in) :: istp
    type(hip_grid), intent(inout) :: grid
    type(hip_field), intent(inout) :: ssha_t, ssha_u, ssha_v, sshn_t, sshn_u, sshn_v, hu, &
         hv, ht, ua, va, un, vn
    real(c_double), intent(in) :: rdt, cbfr, visc, omega, d2r, g
    logical, save :: first_time = .true.

    call c_invoke_time_step(ssha_t%data, sshn_t%data, sshn_u%data, sshn_v%data, hu%data, &
         hv%data, un%data, vn%data, ua%data, ht%data, ssha_u%data, va%data, ssha_v%data, &
         grid%tmask, grid%area_t, grid%area_u, grid%area_v, grid%dx_u, grid%dx_v, &
         grid%dx_t, grid%dy_u, grid%dy_v, grid%dy_t, grid%gphiu, grid%gphiv, &
         ssha_t%device_ptr, sshn_t%device_ptr, sshn_u%device_ptr, sshn_v%device_ptr, &
         hu%device_ptr, hv%device_ptr, un%device_ptr, vn%device_ptr, ua%device_ptr, &
         ht%device_ptr, ssha_u%device_ptr, va%device_ptr, ssha_v%device_ptr, &
         grid%tmask_device, grid%area_t_device, grid%area_u_device, grid%area_v_device, &
         grid%dx_u_device, grid%dx_v_device, grid%dx_t_device, grid%dy_u_device, &
         grid%dy_v_device, grid%dy_t_device, grid%gphiu_device, grid%gphiv_device, &
         int(istp, c_int), int(grid%xstart - 1, c_int), int(grid%xstop - 1, c_int), &
         int(grid%ystart - 1, c_int), int(grid%ystop - 1, c_int), &
         int(size(ssha_t%data, 1), c_int), int(size(ssha_t%data), c_int), &
         rdt, cbfr, visc, omega, d2r, g)

    if (first_time) then
       first_time = .false.
       ssha_t%data_on_device = .true.
       sshn_t%data_on_device = .true.
       sshn_u%data_on_device = .true.
       sshn_v%data_on_device = .true.
       hu%data_on_device = .true.
       hv%data_on_device = .true.
       un%data_on_device = .true.
       vn%data_on_device = .true.
       ua%data_on_device = .true.
       ht%data_on_device = .true.
       ssha_u%data_on_device = .true.
       va%data_on_device = .true.
       ssha_v%data_on_device = .true.
       ! the fields the step writes get a read-back; the depths never change
       ssha_t%read_from_device_c => c_read_from_device
       sshn_t%read_from_device_c => c_read_from_device
       sshn_u%read_from_device_c => c_read_from_device
       sshn_v%read_from_device_c => c_read_from_device
       un%read_from_device_c => c_read_from_device
       vn%read_from_device_c => c_read_from_device
       ua%read_from_device_c => c_read_from_device
       ssha_u%read_from_device_c => c_read_from_device
       va%read_from_device_c => c_read_from_device
       ssha_v%read_from_device_c => c_read_from_device
    end if
  end subroutine invoke_time_step

end module hip_fields

program nemolite2d_hip
  use iso_c_binding
  use nemo_hip_abi
  use hip_fields
  implicit none
  ! common/physical_params_mod.f90:9-15
  real(c_double), parameter :: pi = 3.1415926535897932_c_double
  real(c_double), parameter :: g = 9.80665_c_double, omega = 7.292116e-05_c_double
  real(c_double), parameter :: d2r = pi / 180.0_c_double
  integer :: jpiglo, jpjglo, jphgr_msh, nit000, nitend, irecord
  real(c_double) :: dx, dy, dep_const, rdt, cbfr, visc
  namelist /namctl/ jpiglo, jpjglo, jphgr_msh, dx, dy, dep_const, nit000, nitend, &
       irecord, rdt, cbfr, visc
  type(hip_field), target :: ssha_t, ssha_u, ssha_v, sshn_t, sshn_u, sshn_v, hu, hv, ht, &
       ua, va, un, vn
  type(hip_grid) :: grid
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
  call field_alloc(ssha_t, w, h); call field_alloc(ssha_u, w, h)
  call field_alloc(ssha_v, w, h); call field_alloc(sshn_t, w, h)
  call field_alloc(sshn_u, w, h); call field_alloc(sshn_v, w, h)
  call field_alloc(hu, w, h); call field_alloc(hv, w, h); call field_alloc(ht, w, h)
  call field_alloc(ua, w, h); call field_alloc(va, w, h)
  call field_alloc(un, w, h); call field_alloc(vn, w, h)
  allocate(grid%tmask(w,h), grid%area_t(w,h), grid%area_u(w,h), grid%area_v(w,h), &
       grid%dx_u(w,h), grid%dx_v(w,h), grid%dx_t(w,h), grid%dy_u(w,h), grid%dy_v(w,h), &
       grid%dy_t(w,h), grid%gphiu(w,h), grid%gphiv(w,h))
  grid%xstart = 2; grid%xstop = jpiglo - 1; grid%ystart = 2; grid%ystop = jpjglo - 1
  if (nemo_setup_tmask(jphgr_msh, jpiglo, jpjglo, w, h, grid%tmask) /= 0) stop 'jphgr_msh'
  call nemo_setup_metrics(w, h, dx, dy, 50.0_c_double, grid%area_t, grid%area_u, &
       grid%area_v, grid%dx_u, grid%dx_v, grid%dx_t, grid%dy_u, grid%dy_v, grid%dy_t, &
       grid%gphiu, grid%gphiv)
  call nemo_setup_depth_uniform(w, h, dep_const, ht%data, hu%data, hv%data)
  call nemo_initialise(jpiglo, jpjglo, w, h, grid%area_t, grid%area_u, grid%area_v, &
       sshn_t%data, sshn_u%data, sshn_v%data, un%data, vn%data)
  write(*, "('Simulation domain = (',I4,':',I4,',',I4,':',I4,')')") grid%xstart, &
       grid%xstop, grid%ystart, grid%ystop

  do istp = nit000, nitend
     call invoke_time_step(istp, grid, ssha_t, ssha_u, ssha_v, sshn_t, sshn_u, sshn_v, hu, &
          hv, ht, ua, va, un, vn, rdt, cbfr, visc, omega, d2r, g)
     if (mod(istp, irecord) == 0) then
        call field_get_data(sshn_t)
        write(*, "('step ',I6,' sshn checksum = ', E26.18)") istp, &
             nemo_field_checksum(sshn_t%data, w, 1, jpiglo - 2, 1, jpjglo - 2)
     end if
  end do

  call field_get_data(ua)
  call field_get_data(va)
  call field_get_data(sshn_t)
  write(*, "('ua checksum = ', E16.8)") nemo_field_checksum(ua%data, w, 1, jpiglo - 2, 1, &
       jpjglo - 2)
  write(*, "('va checksum = ', E16.8)") nemo_field_checksum(va%data, w, 1, jpiglo - 2, 1, &
       jpjglo - 2)
  write(*, "('sshn checksum = ', E16.8)") nemo_field_checksum(sshn_t%data, w, 1, jpiglo - 2, &
       1, jpjglo - 2)
  write(*, "('ua full checksum = ', E26.18)") nemo_field_checksum(ua%data, w, 1, jpiglo - 2, &
       1, jpjglo - 2)
  write(*, "('va full checksum = ', E26.18)") nemo_field_checksum(va%data, w, 1, jpiglo - 2, &
       1, jpjglo - 2)
  write(*, "(A)") 'Simulation finished!!'
end program nemolite2d_hip
