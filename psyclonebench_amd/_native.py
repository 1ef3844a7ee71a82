"""ctypes bindings of the in-tree native libraries.

``libnemo_setup.so`` (host set-up) always loads. ``libnemo_hip.so`` loads
wherever ROCm's runtime is installed; its compute entry points need a GPU
and fail loudly without one (they print the HIP error and exit, the
reference's convention). Nothing here falls back to a CPU implementation.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

PKG = Path(__file__).resolve().parent

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_vpp = C.POINTER(C.c_void_p)

N_ARRAYS = 25

# The 25 arrays of the reference C ABI, in argument order
# (manual_versions/psykal_cpp/time_step.cpp:12-37).
ARRAY_NAMES = (
    "ssha_t", "sshn_t", "sshn_u", "sshn_v", "hu", "hv", "un", "vn", "ua",
    "ht", "ssha_u", "va", "ssha_v", "tmask", "area_t", "area_u", "area_v",
    "dx_u", "dx_v", "dx_t", "dy_u", "dy_v", "dy_t", "gphiu", "gphiv",
)

_HOST = [_ip if n == "tmask" else _dp for n in ARRAY_NAMES]
_SCAL_DEV = [C.c_int] * 7 + [C.c_double] * 6     # istep, 4 bounds, width, total
_SCAL_HOST = [C.c_int] * 6 + [C.c_double] * 6    # istep, 4 bounds, width

_setup = None
_hip = None
_hip_cpp = None


def _load(name: str) -> C.CDLL:
    path = PKG / name
    if not path.exists():
        raise OSError(
            f"{path} is not built; run `python -c 'import __graft_entry__ as g; g.build()'`"
        )
    return C.CDLL(str(path))


def setup_lib() -> C.CDLL:
    global _setup
    if _setup is None:
        lib = _load("libnemo_setup.so")
        lib.nemo_grid_dims.argtypes = [C.c_int, C.c_int, C.c_int, _ip, _ip]
        lib.nemo_setup_tmask.argtypes = [C.c_int] * 5 + [_ip]
        lib.nemo_setup_tmask.restype = C.c_int
        lib.nemo_setup_metrics.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double,
                                           C.c_double] + [_dp] * 11
        lib.nemo_setup_depth_uniform.argtypes = [C.c_int, C.c_int, C.c_double] + [_dp] * 3
        lib.nemo_setup_bathymetry.argtypes = [C.c_uint64] + [C.c_int] * 4 + \
            [C.c_double] * 4 + [C.c_int] + [_dp] * 3
        lib.nemo_initialise.argtypes = [C.c_int] * 4 + [_dp] * 7
        lib.nemo_field_checksum.argtypes = [_dp] + [C.c_int] * 5
        lib.nemo_field_checksum.restype = C.c_double
        lib.nemo_uniforms.argtypes = [C.c_uint64, C.c_int64, _dp]
        lib.nemo_fortran_e.argtypes = [C.c_double, C.c_int, C.c_int, C.c_char_p]
        lib.nemo_write_results.argtypes = [C.c_char_p] + [_dp] * 4 + [C.c_int] * 5 + \
            [C.c_double] * 2
        lib.nemo_write_results.restype = C.c_int
        lib.nemo_dump_tmask.argtypes = [C.c_char_p, _ip] + [C.c_int] * 7 + [C.c_double] * 2 + \
            [C.c_int]
        lib.nemo_dump_tmask.restype = C.c_int
        _setup = lib
    return _setup


def hip_lib() -> C.CDLL:
    """libnemo_hip.so with argument types set. Raises OSError if not built
    or if the HIP runtime cannot be loaded."""
    global _hip
    if _hip is None:
        lib = _load("libnemo_hip.so")
        lib.c_invoke_time_step.argtypes = _HOST + [_vpp] * N_ARRAYS + _SCAL_DEV
        lib.c_invoke_time_steps.argtypes = [C.c_int] + _HOST + [_vpp] * N_ARRAYS + _SCAL_DEV
        lib.c_invoke_time_step_host.argtypes = _HOST + _SCAL_HOST
        lib.c_invoke_time_step_noalloc.argtypes = [C.c_void_p] * N_ARRAYS + _SCAL_DEV + \
            [C.c_void_p]
        lib.c_read_from_device.argtypes = [C.c_void_p, C.c_void_p] + [C.c_int] * 4 + [C.c_bool]
        lib.c_write_to_device.argtypes = [C.c_void_p, C.c_void_p] + [C.c_int] * 4 + [C.c_bool]
        lib.nemo_hip_set_tide.argtypes = [C.c_double, C.c_double]
        lib.nemo_hip_set_mode.argtypes = [C.c_int]
        lib.nemo_hip_context_mode.argtypes = [_vpp]
        lib.nemo_hip_context_mode.restype = C.c_int
        lib.nemo_hip_set_kernel_timing.argtypes = [C.c_int]
        lib.nemo_hip_fused_kernel_ms.restype = C.c_double
        lib.nemo_hip_set_fused_variant.argtypes = [C.c_int]
        lib.nemo_hip_fused_variant.restype = C.c_int
        lib.nemo_hip_fused_variant_count.restype = C.c_int
        lib.nemo_hip_fused_variant_shape.argtypes = [C.c_int, _ip, _ip, _ip]
        lib.nemo_hip_fused_variant_kernel.argtypes = [C.c_int]
        lib.nemo_hip_fused_variant_kernel.restype = C.c_char_p
        lib.nemo_mp_create.argtypes = _HOST + [C.c_int] * 6 + [C.c_double] * 8
        lib.nemo_mp_create.restype = C.c_void_p
        lib.nemo_mp_steps.argtypes = [C.c_void_p, C.c_int, C.c_int]
        lib.nemo_mp_kernel_ms.argtypes = [C.c_void_p, C.c_int, C.c_int]
        lib.nemo_mp_kernel_ms.restype = C.c_double
        lib.nemo_mp_set_rows.argtypes = [C.c_void_p, C.c_int]
        lib.nemo_mp_set_columns.argtypes = [C.c_void_p, C.c_int]
        lib.nemo_mp_set_prefetch.argtypes = [C.c_void_p, C.c_int]
        lib.nemo_mp_read.argtypes = [C.c_void_p, C.c_int, _dp]
        lib.nemo_mp_destroy.argtypes = [C.c_void_p]
        vp = C.c_void_p
        lib.nemo_ens_create.argtypes = [C.c_int, C.POINTER(vp), _ip, _dp, _dp] + \
            [C.c_double] * 6 + [C.c_int]
        lib.nemo_ens_create.restype = vp
        lib.nemo_ens_shape_count.argtypes = [vp]
        lib.nemo_ens_shape_count.restype = C.c_int
        lib.nemo_ens_shape_info.argtypes = [vp, C.c_int, _ip, _ip, _ip]
        lib.nemo_ens_set_shape.argtypes = [vp, C.c_int]
        lib.nemo_ens_shape_columns.argtypes = [vp, C.c_int]
        lib.nemo_ens_shape_columns.restype = C.c_int
        lib.nemo_ens_shape_prefetch.argtypes = [vp, C.c_int]
        lib.nemo_ens_shape_prefetch.restype = C.c_int
        lib.nemo_ens_shape.argtypes = [vp]
        lib.nemo_ens_shape.restype = C.c_int
        lib.nemo_ens_autotune.argtypes = [vp, C.c_int, _dp]
        lib.nemo_ens_autotune.restype = C.c_int
        lib.nemo_ens_steps.argtypes = [vp, C.c_int, C.c_int]
        lib.nemo_ens_kernel_ms.argtypes = [vp, C.c_int, C.c_int]
        lib.nemo_ens_kernel_ms.restype = C.c_double
        lib.nemo_ens_interior_points.argtypes = [vp]
        lib.nemo_ens_interior_points.restype = C.c_longlong
        lib.nemo_ens_read.argtypes = [vp, C.c_int, C.c_int, _dp]
        lib.nemo_ens_destroy.argtypes = [vp]
        fp = C.POINTER(C.c_float)
        st = [vp] * 10  # coarse state: ssh, u, v, sshu, sshv, ht, tmask, area_t/u/v
        lib.nemo_cl_create.argtypes = [C.c_int] * 4 + [fp]
        lib.nemo_cl_create.restype = vp
        lib.nemo_cl_param_count.restype = C.c_int
        lib.nemo_cl_layer_info.argtypes = [C.c_int, _ip, _ip, _ip]
        lib.nemo_cl_get_params.argtypes = [vp, fp]
        lib.nemo_cl_set_params.argtypes = [vp, fp]
        lib.nemo_cl_set_norm.argtypes = [vp, _dp, _dp, _dp]
        lib.nemo_cl_set_optimizer.argtypes = [vp, C.c_int] + [C.c_double] * 4
        lib.nemo_cl_set_guard.argtypes = [vp] + [C.c_double] * 3
        lib.nemo_cl_filter.argtypes = [vp, vp, vp, vp, C.c_int, C.c_int]
        lib.nemo_cl_read_filtered.argtypes = [vp, C.c_int, _dp]
        lib.nemo_cl_set_state.argtypes = [vp] + st
        lib.nemo_cl_train_step.argtypes = [vp] + st
        lib.nemo_cl_last_loss.argtypes = [vp]
        lib.nemo_cl_last_loss.restype = C.c_double
        lib.nemo_cl_set_clip.argtypes = [vp, C.c_double]
        lib.nemo_cl_set_loss_weights.argtypes = [vp] + [C.c_double] * 3
        lib.nemo_cl_set_gate.argtypes = [vp, C.c_double]
        lib.nemo_cl_set_ssh_cells.argtypes = [vp, C.c_int]
        lib.nemo_cl_set_guard_limit.argtypes = [vp, C.c_double]
        lib.nemo_cl_last_stats.argtypes = [vp, _dp]
        lib.nemo_cl_correct.argtypes = [vp, C.c_int] + st
        lib.nemo_cl_steps.argtypes = [vp, _vpp, C.c_int, C.c_int, C.c_int]
        lib.nemo_cl_guard_trips.argtypes = [vp]
        lib.nemo_cl_guard_trips.restype = C.c_int
        lib.nemo_cl_set_timing.argtypes = [vp, C.c_int]
        lib.nemo_cl_forward_ms.argtypes = [vp]
        lib.nemo_cl_forward_ms.restype = C.c_double
        lib.nemo_cl_set_conv_form.argtypes = [C.c_int]
        lib.nemo_cl_conv_form.restype = C.c_int
        lib.nemo_cl_forward_host.argtypes = [vp, fp, fp]
        lib.nemo_cl_grad_host.argtypes = [vp, fp, fp, fp]
        lib.nemo_cl_grad_host.restype = C.c_double
        lib.nemo_cl_train_host.argtypes = [vp, fp, fp]
        lib.nemo_cl_train_host.restype = C.c_double
        lib.nemo_cl_destroy.argtypes = [vp]
        lib.nemo_hip_release.argtypes = [_vpp]
        lib.nemo_hip_device_info.argtypes = [C.c_char_p, C.c_int]
        lib.nemo_hip_device_info.restype = C.c_int
        _hip = lib
    return _hip


def hip_cpp_lib() -> C.CDLL:
    """libnemo_hip_cpp.so: c_invoke_time_step with the psykal_cpp signature."""
    global _hip_cpp
    if _hip_cpp is None:
        hip_lib()
        lib = _load("libnemo_hip_cpp.so")
        lib.c_invoke_time_step.argtypes = _HOST + _SCAL_HOST
        _hip_cpp = lib
    return _hip_cpp


def device_info() -> tuple[int, str]:
    buf = C.create_string_buffer(256)
    n = hip_lib().nemo_hip_device_info(buf, 256)
    return n, buf.value.decode()
