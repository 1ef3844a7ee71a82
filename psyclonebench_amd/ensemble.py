"""Batched ensembles of NEMOLite2D domains (no reference counterpart).

:func:`ensemble_configs` draws the domains of a seeded ensemble: per domain
k, five SplitMix64 uniforms u0..u4 from the ensemble seed give

* interior sides ``nx, ny = round(exp(ln n_min + u * (ln n_max - ln n_min)))``
  (log-uniform in [n_min, n_max]),
* mean depth ``50 + 450 u2`` m with +-50 % smooth relief (sigma 16 cells,
  bathymetry seed ``seed * 1000003 + k + 1``), clipped to >= 20 m,
* tidal amplitude ``0.5 + 1.5 u3`` m and period ``40000 + 6000 u4`` s,

each with an open west edge and the physics of the 128^2 case (dx = dy =
1 km, dt = 20 s, cbfr 1.5e-3, visc 0.1).

:class:`EnsembleStepper` advances a list of :class:`~.model.Model` together
through ``nemo_ens_*`` (one launch per step over a tile -> domain work list).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import replace

import numpy as np

from ._native import N_ARRAYS, hip_lib, setup_lib
from .model import Bathymetry, Model, NemoConfig, c1_config


def ensemble_configs(ndom: int = 256, seed: int = 0, n_min: int = 64,
                     n_max: int = 1024, nsteps: int = 2000) -> list[NemoConfig]:
    u = np.empty(5 * ndom, np.float64)
    setup_lib().nemo_uniforms(seed, 5 * ndom, u.ctypes.data_as(C.POINTER(C.c_double)))
    lo, hi = math.log(n_min), math.log(n_max)
    out = []
    for k in range(ndom):
        a = u[5 * k:5 * k + 5]
        nx = int(round(math.exp(lo + a[0] * (hi - lo))))
        ny = int(round(math.exp(lo + a[1] * (hi - lo))))
        mean = 50.0 + 450.0 * a[2]
        amp = 0.5 + 1.5 * a[3]
        period = 40000.0 + 6000.0 * a[4]
        out.append(replace(c1_config(nsteps), jpiglo=nx + 2, jpjglo=ny + 2,
                           bathymetry=Bathymetry(seed=seed * 1000003 + k + 1, sigma=16.0,
                                                 mean=mean, relief=0.5 * mean,
                                                 min_depth=20.0),
                           tide_amp=amp, tide_omega=2.0 * math.pi / period))
    return out


class EnsembleStepper:
    """Every model of `models` advanced together; precision 64 (FP64 step)
    or 32 (mixed form). The models' host arrays are read at creation only."""

    FIELDS = {"sshn_t": 0, "un": 1, "vn": 2, "sshn_u": 3, "sshn_v": 4}

    def __init__(self, models: list[Model], precision: int = 64):
        lib = hip_lib()
        self.models = models
        self.precision = precision
        nd = len(models)
        scal = models[0].scalars
        for m in models:
            if m.scalars != scal:
                raise ValueError("ensemble domains must share rdt, cbfr, visc and constants")
        ptrs = (C.c_void_p * (nd * N_ARRAYS))()
        geom = np.empty(6 * nd, np.int32)
        amp = np.empty(nd, np.float64)
        omg = np.empty(nd, np.float64)
        for d, m in enumerate(models):
            for a, arr in enumerate(m.arrays()):
                ptrs[d * N_ARRAYS + a] = arr.ctypes.data
            g = m.grid
            geom[6 * d:6 * d + 6] = (g.xs, g.xe, g.ys, g.ye, g.W, g.total)
            amp[d] = m.cfg.tide_amp
            omg[d] = m.cfg.tide_omega
        self.istp = models[0].cfg.nit000
        self.handle = lib.nemo_ens_create(
            nd, ptrs, geom.ctypes.data_as(C.POINTER(C.c_int)),
            amp.ctypes.data_as(C.POINTER(C.c_double)),
            omg.ctypes.data_as(C.POINTER(C.c_double)), *scal, precision)

    @property
    def interior_points(self) -> int:
        return int(hip_lib().nemo_ens_interior_points(self.handle))

    def shapes(self) -> list[tuple[int, int, int, int, int]]:
        """(rows per band, wavefronts per workgroup, workgroups (0: one per
        tile, else a persistent launch), columns per wavefront, prefetch
        depth) per shape."""
        lib = hip_lib()
        out = []
        for k in range(lib.nemo_ens_shape_count(self.handle)):
            r, w, b = C.c_int(), C.c_int(), C.c_int()
            lib.nemo_ens_shape_info(self.handle, k, C.byref(r), C.byref(w), C.byref(b))
            out.append((r.value, w.value, b.value, lib.nemo_ens_shape_columns(self.handle, k),
                        lib.nemo_ens_shape_prefetch(self.handle, k)))
        return out

    @property
    def shape(self) -> int:
        return hip_lib().nemo_ens_shape(self.handle)

    def set_shape(self, k: int) -> None:
        hip_lib().nemo_ens_set_shape(self.handle, k)

    def autotune(self, steps_per_shape: int = 5) -> tuple[int, list[float]]:
        lib = hip_lib()
        n = lib.nemo_ens_shape_count(self.handle)
        ms = np.zeros(n, np.float64)
        best = lib.nemo_ens_autotune(self.handle, steps_per_shape,
                                     ms.ctypes.data_as(C.POINTER(C.c_double)))
        return best, [float(x) for x in ms]

    def run(self, nsteps: int) -> None:
        hip_lib().nemo_ens_steps(self.handle, nsteps, self.istp)
        self.istp += nsteps

    def timed_run(self, nsteps: int) -> float:
        ms = hip_lib().nemo_ens_kernel_ms(self.handle, nsteps, self.istp)
        self.istp += nsteps
        return ms

    def synchronize(self) -> None:
        hip_lib().nemo_hip_synchronize()

    def fetch(self, d: int, name: str) -> np.ndarray:
        g = self.models[d].grid
        out = np.empty((g.H, g.W), np.float64)
        hip_lib().nemo_ens_read(self.handle, d, self.FIELDS[name],
                                out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def release(self) -> None:
        if self.handle:
            hip_lib().nemo_ens_destroy(self.handle)
            self.handle = None
