"""TEST INFRASTRUCTURE ONLY -- the oracle and the reference build.

Loads ``oracle/libnemo_oracle.so`` (the plain-C restatement in
``nemo_oracle.c``) and, where built, the reference's own C++ PSy layer in
``oracle/_ref/``. Only tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may use this package, and only as the
checker or the timed CPU baseline; the product never imports it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REF_DIR = HERE / "_ref"

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)

ARRAY_NAMES = (
    "ssha_t", "sshn_t", "sshn_u", "sshn_v", "hu", "hv", "un", "vn", "ua",
    "ht", "ssha_u", "va", "ssha_v", "tmask", "area_t", "area_u", "area_v",
    "dx_u", "dx_v", "dx_t", "dy_u", "dy_v", "dy_t", "gphiu", "gphiv",
)
_STEP_ARGTYPES = ([_ip if n == "tmask" else _dp for n in ARRAY_NAMES]
                  + [C.c_int] * 6 + [C.c_double] * 6)

# per-kernel sweeps: (name, argtypes after the region+width prefix)
_R = [C.c_int] * 5
KERNEL_ARGTYPES = {
    "continuity": _R + [_dp] * 8 + [C.c_double, _dp],
    "momentum_u": _R + [_dp] * 10 + [_ip] + [_dp] * 7 + [C.c_double] * 6,
    "momentum_v": _R + [_dp] * 10 + [_ip] + [_dp] * 7 + [C.c_double] * 6,
    "bc_ssh": _R + [C.c_int, _dp, _ip, C.c_double],
    "bc_solid_u": _R + [_dp, _ip],
    "bc_solid_v": _R + [_dp, _ip],
    "bc_flather_u": _R + [_dp, _dp, _dp, _ip, C.c_double],
    "bc_flather_v": _R + [_dp, _dp, _dp, _ip, C.c_double],
    "next_sshu": _R + [_dp, _dp, _ip, _dp, _dp],
    "next_sshv": _R + [_dp, _dp, _ip, _dp, _dp],
}

_oracle = None


def build(with_reference: bool = True) -> None:
    """make -C oracle: the restatement always, the reference build when the
    reference tree is present."""
    target = "all" if with_reference else "oracle"
    subprocess.run(["make", "-s", "-C", str(HERE), target], check=True)


def oracle_lib() -> C.CDLL:
    global _oracle
    if _oracle is None:
        path = HERE / "libnemo_oracle.so"
        if not path.exists():
            build(with_reference=False)
        lib = C.CDLL(str(path))
        lib.oracle_invoke_time_step.argtypes = _STEP_ARGTYPES
        lib.oracle_invoke_time_step_omp.argtypes = _STEP_ARGTYPES
        lib.oracle_set_tide.argtypes = [C.c_double, C.c_double]
        lib.oracle_checksum.argtypes = [_dp] + [C.c_int] * 5
        lib.oracle_checksum.restype = C.c_double
        for k, at in KERNEL_ARGTYPES.items():
            getattr(lib, "oracle_" + k).argtypes = at
        _oracle = lib
    return _oracle


def ref_available() -> bool:
    return (REF_DIR / "libref_time_step.so").exists()


def ref_lib(name: str = "libref_time_step.so") -> C.CDLL:
    """A reference build from oracle/_ref (c_invoke_time_step, or the
    per-kernel ref_* sweeps for libref_kernels.so)."""
    lib = C.CDLL(str(REF_DIR / name))
    if name == "libref_kernels.so":
        for k, at in KERNEL_ARGTYPES.items():
            getattr(lib, "ref_" + k).argtypes = at
    else:
        lib.c_invoke_time_step.argtypes = _STEP_ARGTYPES
    return lib


def pointers(arrays: list[np.ndarray]) -> list:
    out = []
    for a in arrays:
        assert a.flags["C_CONTIGUOUS"]
        out.append(a.ctypes.data_as(_ip if a.dtype == np.int32 else _dp))
    return out


def step_fn(kind: str = "oracle"):
    """Callable(arrays, istep, bounds, width, scalars) for one step:
    kind = oracle | oracle_omp | ref | ref_omp | ref_ofast."""
    if kind == "oracle":
        f = oracle_lib().oracle_invoke_time_step
    elif kind == "oracle_omp":
        f = oracle_lib().oracle_invoke_time_step_omp
    elif kind == "ref":
        f = ref_lib().c_invoke_time_step
    elif kind == "ref_omp":
        f = ref_lib("libref_time_step_omp.so").c_invoke_time_step
    elif kind == "ref_ofast":
        f = ref_lib("libref_time_step_ofast.so").c_invoke_time_step
    else:
        raise ValueError(kind)

    def run(arrays, istep, bounds, width, scalars):
        f(*pointers(arrays), istep, *bounds, width, *scalars)
    return run


def set_tide(amp: float, omega_tide: float) -> None:
    oracle_lib().oracle_set_tide(amp, omega_tide)


def run_steps(arrays: list[np.ndarray], istep0: int, nsteps: int, bounds, width,
              scalars, kind: str = "oracle") -> None:
    """nsteps steps in place on the 25 arrays (C-ABI order)."""
    f = step_fn(kind)
    for k in range(nsteps):
        f(arrays, istep0 + k, bounds, width, scalars)


def checksum(a: np.ndarray, width: int, bounds) -> float:
    xs, xe, ys, ye = bounds
    return oracle_lib().oracle_checksum(a.ctypes.data_as(_dp), width, xs, xe, ys, ye)
