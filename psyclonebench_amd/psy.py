"""Host-side mirror of the reference's PSy-layer call for the time step.

:func:`invoke_time_step` has the argument list and behaviour of the
device-resident Fortran shim ``invoke_time_step``
(``manual_versions/psykal_c_opencl/time_step_mod.f90:76-192``): it hands
the field data, the grid arrays and the device-handle slots to the C ABI
``c_invoke_time_step``; on the first call the library uploads everything
and fills the slots, and the fields are marked as living on the device with
``c_read_from_device`` as their read-back hook.

:class:`TimeStepper` wraps the same call for a :class:`~.model.Model` and
adds the multi-step hipGraph replay (``c_invoke_time_steps``).
"""
from __future__ import annotations

import ctypes as C

from ._native import ARRAY_NAMES, hip_lib
from .model import GRID_NAMES, Field, Grid, Model, _ptr


class GridSlots:
    """Device-handle slots of the 12 grid arrays (the shim keeps them with
    the grid object, like field%device_ptr for fields)."""

    def __init__(self, grid: Grid):
        self.grid = grid
        self.slots = {n: C.c_void_p() for n in GRID_NAMES}


def _call(nsteps: int | None, istp: int, fields: dict[str, Field], grid: Grid,
          gslots: GridSlots, scalars) -> None:
    lib = hip_lib()
    host = []
    slots = []
    for n in ARRAY_NAMES:
        if n in fields:
            host.append(_ptr(fields[n].data))
            slots.append(C.cast(C.byref(fields[n].device_ptr), C.POINTER(C.c_void_p)))
        else:
            host.append(_ptr(grid.array(n)))
            slots.append(C.cast(C.byref(gslots.slots[n]), C.POINTER(C.c_void_p)))
    tail = (istp, grid.xs, grid.xe, grid.ys, grid.ye, grid.W, grid.total, *scalars)
    if nsteps is None:
        lib.c_invoke_time_step(*host, *slots, *tail)
    else:
        lib.c_invoke_time_steps(nsteps, *host, *slots, *tail)
    for f in fields.values():
        f.data_on_device = True


def invoke_time_step(istp: int, ssha_t: Field, ssha_u: Field, ssha_v: Field,
                     sshn_t: Field, sshn_u: Field, sshn_v: Field, hu: Field,
                     hv: Field, ht: Field, ua: Field, va: Field, un: Field,
                     vn: Field, *, grid: Grid, grid_slots: GridSlots,
                     scalars) -> None:
    """One time step on the device, argument order of the reference shim.
    ``scalars`` = (rdt, cbfr, visc, omega, d2r, g)."""
    fields = dict(ssha_t=ssha_t, ssha_u=ssha_u, ssha_v=ssha_v, sshn_t=sshn_t,
                  sshn_u=sshn_u, sshn_v=sshn_v, hu=hu, hv=hv, ht=ht, ua=ua,
                  va=va, un=un, vn=vn)
    _call(None, istp, fields, grid, grid_slots, scalars)


class TimeStepper:
    """Drives one :class:`Model` through the device-resident C ABI."""

    def __init__(self, model: Model, mode: int = 0):
        self.model = model
        self.grid_slots = GridSlots(model.grid)
        self.istp = model.cfg.nit000
        self.mode = mode
        self._configured = False

    def _configure(self):
        lib = hip_lib()
        lib.nemo_hip_set_tide(self.model.cfg.tide_amp, self.model.cfg.tide_omega)
        lib.nemo_hip_set_mode(self.mode)
        self._configured = True

    def step(self) -> None:
        """One c_invoke_time_step call (plain kernel launches)."""
        if not self._configured:
            self._configure()
        m = self.model
        _call(None, self.istp, m.fields, m.grid, self.grid_slots, m.scalars)
        self.istp += 1

    def run(self, nsteps: int) -> None:
        """nsteps steps in one c_invoke_time_steps call (hipGraph replay)."""
        if not self._configured:
            self._configure()
        m = self.model
        _call(nsteps, self.istp, m.fields, m.grid, self.grid_slots, m.scalars)
        self.istp += nsteps

    @property
    def context_mode(self) -> int:
        """1 unfused, 2 fused, 0 before the first step."""
        slot = C.cast(C.byref(self.model.fields["ssha_t"].device_ptr),
                      C.POINTER(C.c_void_p))
        return hip_lib().nemo_hip_context_mode(slot)

    def fetch(self, name: str):
        """Host copy of a field (blocking read-back)."""
        return self.model.fields[name].get_data().copy()

    def synchronize(self) -> None:
        hip_lib().nemo_hip_synchronize()

    def release(self) -> None:
        slot = C.cast(C.byref(self.model.fields["ssha_t"].device_ptr),
                      C.POINTER(C.c_void_p))
        hip_lib().nemo_hip_release(slot)
        for f in self.model.fields.values():
            f.data_on_device = False
            f.device_ptr.value = None


class MixedStepper:
    """The mixed-precision form (nemo_mp_*): FP32 fields, compensated
    hi/lo FP32 ssh with FP64 accumulation, FP64 tide. Starts from the
    model's FP64 host arrays."""

    FIELDS = {"sshn_t": 0, "un": 1, "vn": 2, "sshn_u": 3, "sshn_v": 4}

    def __init__(self, model: Model, rows: int = 32, cols: int = 61, prefetch: int = 0):
        lib = hip_lib()
        g = model.grid
        c = model.cfg
        self.model = model
        self.istp = c.nit000
        self.handle = lib.nemo_mp_create(*model.host_pointers(), g.xs, g.xe, g.ys, g.ye,
                                         g.W, g.total, *model.scalars, c.tide_amp,
                                         c.tide_omega)
        lib.nemo_mp_set_rows(self.handle, rows)
        lib.nemo_mp_set_columns(self.handle, cols)
        lib.nemo_mp_set_prefetch(self.handle, prefetch)

    def run(self, nsteps: int) -> None:
        hip_lib().nemo_mp_steps(self.handle, nsteps, self.istp)
        self.istp += nsteps

    def timed_run(self, nsteps: int) -> float:
        """nsteps event-timed steps; mean kernel ms."""
        ms = hip_lib().nemo_mp_kernel_ms(self.handle, nsteps, self.istp)
        self.istp += nsteps
        return ms

    def fetch(self, name: str):
        import numpy as np
        g = self.model.grid
        out = np.empty((g.H, g.W), np.float64)
        hip_lib().nemo_mp_read(self.handle, self.FIELDS[name],
                               out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def release(self) -> None:
        if self.handle:
            hip_lib().nemo_mp_destroy(self.handle)
            self.handle = None


def invoke_time_step_host(model: Model, istp: int) -> None:
    """Host-pointer step (psykal_cpp/time_step.cpp:10-51) through
    c_invoke_time_step_host: host arrays in, host arrays updated."""
    lib = hip_lib()
    g = model.grid
    lib.nemo_hip_set_tide(model.cfg.tide_amp, model.cfg.tide_omega)
    lib.c_invoke_time_step_host(*model.host_pointers(), istp, g.xs, g.xe, g.ys,
                                g.ye, g.W, *model.scalars)
