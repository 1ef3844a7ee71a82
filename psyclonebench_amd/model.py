"""NEMOLite2D model set-up in the dl_esm_inf layout.

Stands in for the reference's Fortran model layer and the dl_esm_inf grid
and field types (the dl_esm_inf submodule is absent from the reference
tree; its layout is inferred from its users):

* :class:`Grid`   -- ``grid_type``: T-mask, metric arrays, internal region
  (``common/model_mod.f90:26-69``; metrics as ``original/nemolite2d.f90:
  218-231,274-290``)
* :class:`Field`  -- ``r2d_field``: a W x H array plus the device-handle
  slot and read-back hook of the device-resident PSy layer
  (``psykal_c_opencl/time_step_mod.f90:162-190``)
* :func:`make_model` -- ``model_init`` + ``initialisation``
  (``common/initialisation_mod.f90:7-63``)

Arrays are numpy ``(H, W)`` C-order, so ``a.ravel()[jj*W + ji]`` is point
(ji, jj), the C index of the reference kernels.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field, replace
from typing import Optional

import numpy as np

from ._native import ARRAY_NAMES, setup_lib

# common/physical_params_mod.f90:9-15
PI = 3.1415926535897932
G = 9.80665
OMEGA = 7.292116e-05
D2R = PI / 180.0

# boundary_conditions_kern.c:185-186 (the reference's hard-wired tide)
REF_TIDE_AMP = 0.2
REF_TIDE_OMEGA = 2.0 * 3.14159 / (12.42 * 3600.0)

MASK_CLOSED, MASK_SOUTH_OPEN, MASK_EAST_OPEN, MASK_SOUTH_EAST, MASK_WEST_OPEN = range(5)

FIELD_NAMES = ("ssha_t", "sshn_t", "sshn_u", "sshn_v", "hu", "hv", "un", "vn",
               "ua", "ht", "ssha_u", "va", "ssha_v")
GRID_NAMES = ("tmask", "area_t", "area_u", "area_v", "dx_u", "dx_v", "dx_t",
              "dy_u", "dy_v", "dy_t", "gphiu", "gphiv")


@dataclass
class Bathymetry:
    """Seeded smooth random depth: mean + relief * (low-passed noise in
    [-1,1]), clipped to >= min_depth (see nemo_setup_bathymetry)."""
    seed: int = 0
    sigma: float = 32.0
    mean: float = 100.0
    relief: float = 50.0
    min_depth: float = 20.0


@dataclass
class NemoConfig:
    """The namelist ``&namctl`` (common/gocean2d_io_mod.f90:50-68) plus the
    extensions this project adds (open west edge, tide parameters, seeded
    bathymetry)."""
    jpiglo: int = 130
    jpjglo: int = 130
    jphgr_msh: int = MASK_SOUTH_OPEN
    dx: float = 1000.0
    dy: float = 1000.0
    dep_const: float = 100.0
    nit000: int = 1
    nitend: int = 100
    rdt: float = 20.0
    cbfr: float = 0.00015
    visc: float = 0.1
    tide_amp: float = REF_TIDE_AMP
    tide_omega: float = REF_TIDE_OMEGA
    bathymetry: Optional[Bathymetry] = None
    gphi: float = 50.0
    align: int = 64

    @property
    def nsteps(self) -> int:
        return self.nitend - self.nit000 + 1

    @property
    def interior_points(self) -> int:
        return (self.jpiglo - 2) * (self.jpjglo - 2)


def reference_config(n_interior: int, nsteps: int, mask: int = MASK_SOUTH_OPEN,
                     **kw) -> NemoConfig:
    """The reference namelists (original/namelist.*): south-open channel,
    100 m deep, rdt 20 s, cbfr 1.5e-4, visc 0.1, reference tide literals."""
    return NemoConfig(jpiglo=n_interior + 2, jpjglo=n_interior + 2,
                      jphgr_msh=mask, nitend=nsteps, **kw)


def c1_config(nsteps: int = 2000) -> NemoConfig:
    """128^2 interior, dx=dy=1 km, depth 100 m, dt 20 s, cbfr 1.5e-3,
    visc 0.1; open west edge forced by 1.0 m * sin(2 pi t / 44712 s)."""
    return NemoConfig(jpiglo=130, jpjglo=130, jphgr_msh=MASK_WEST_OPEN,
                      nitend=nsteps, cbfr=0.0015, tide_amp=1.0,
                      tide_omega=2.0 * math.pi / 44712.0)


def c2_config(n_interior: int = 4096, nsteps: int = 1000, seed: int = 0) -> NemoConfig:
    """C1's physics at n^2 with seeded bathymetry 100 m + 50 m * smooth
    noise (sigma 32 cells), clipped >= 20 m."""
    return replace(c1_config(nsteps), jpiglo=n_interior + 2, jpjglo=n_interior + 2,
                   bathymetry=Bathymetry(seed=seed))


def _ptr(a: np.ndarray):
    if a.dtype == np.int32:
        return a.ctypes.data_as(C.POINTER(C.c_int))
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Grid:
    """dl_esm_inf ``grid_type`` restated: M x N T-mask extent (one boundary
    ring included), W x H allocation, 0-based internal region
    xs..xe x ys..ye = 1..M-2 x 1..N-2."""

    def __init__(self, cfg: NemoConfig):
        lib = setup_lib()
        self.M, self.N = cfg.jpiglo, cfg.jpjglo
        if self.M < 5 or self.N < 5:
            raise ValueError("jpiglo and jpjglo must be >= 5")
        w, h = C.c_int(), C.c_int()
        lib.nemo_grid_dims(self.M, self.N, cfg.align, C.byref(w), C.byref(h))
        self.W, self.H = w.value, h.value
        self.xs, self.xe, self.ys, self.ye = 1, self.M - 2, 1, self.N - 2
        self.dx, self.dy = cfg.dx, cfg.dy
        shape = (self.H, self.W)
        self.tmask = np.zeros(shape, np.int32)
        if lib.nemo_setup_tmask(cfg.jphgr_msh, self.M, self.N, self.W, self.H,
                                _ptr(self.tmask)) != 0:
            raise ValueError(f"unknown mask mode {cfg.jphgr_msh}")
        self.metrics = {n: np.empty(shape, np.float64) for n in GRID_NAMES[1:]}
        m = self.metrics
        lib.nemo_setup_metrics(self.W, self.H, cfg.dx, cfg.dy, cfg.gphi,
                               *(_ptr(m[n]) for n in GRID_NAMES[1:]))

    @property
    def total(self) -> int:
        return self.W * self.H

    @property
    def bounds(self) -> tuple[int, int, int, int]:
        return self.xs, self.xe, self.ys, self.ye

    def array(self, name: str) -> np.ndarray:
        return self.tmask if name == "tmask" else self.metrics[name]


@dataclass
class Field:
    """``r2d_field``: host data, the device-handle slot that the
    device-resident PSy layer fills (field%device_ptr), and whether the
    device copy is the authoritative one (field%data_on_device)."""
    name: str
    data: np.ndarray
    device_ptr: C.c_void_p = field(default_factory=C.c_void_p)
    data_on_device: bool = False

    def get_data(self) -> np.ndarray:
        """dl_esm_inf get_data(): fetches the device copy first when the
        device holds the data (psykal_c_opencl/time_step_mod.f90:179-189,
        gocean2d_io_mod.f90:115-118)."""
        if self.data_on_device and self.device_ptr.value:
            from ._native import hip_lib
            h, w = self.data.shape
            hip_lib().c_read_from_device(self.device_ptr, self.data.ctypes.data,
                                         1, 1, w, h, True)
        return self.data


class Model:
    """Grid, the 13 fields and the run parameters of one NEMOLite2D model."""

    def __init__(self, cfg: NemoConfig):
        self.cfg = cfg
        self.grid = Grid(cfg)
        g = self.grid
        shape = (g.H, g.W)
        self.fields = {n: Field(n, np.zeros(shape, np.float64)) for n in FIELD_NAMES}
        f = {n: self.fields[n].data for n in FIELD_NAMES}
        lib = setup_lib()
        if cfg.bathymetry is None:
            lib.nemo_setup_depth_uniform(g.W, g.H, cfg.dep_const, _ptr(f["ht"]),
                                         _ptr(f["hu"]), _ptr(f["hv"]))
        else:
            b = cfg.bathymetry
            lib.nemo_setup_bathymetry(b.seed, g.M, g.N, g.W, g.H, b.sigma, b.mean,
                                      b.relief, b.min_depth, 0, _ptr(f["ht"]),
                                      _ptr(f["hu"]), _ptr(f["hv"]))
        m = g.metrics
        lib.nemo_initialise(g.M, g.N, g.W, g.H, _ptr(m["area_t"]), _ptr(m["area_u"]),
                            _ptr(m["area_v"]), _ptr(f["sshn_t"]), _ptr(f["sshn_u"]),
                            _ptr(f["sshn_v"]), _ptr(f["un"]), _ptr(f["vn"]))

    # the physical scalars of the C ABI: rdt, cbfr, visc, omega, d2r, g
    @property
    def scalars(self) -> tuple[float, ...]:
        c = self.cfg
        return (c.rdt, c.cbfr, c.visc, OMEGA, D2R, G)

    def array(self, name: str) -> np.ndarray:
        if name in self.fields:
            return self.fields[name].data
        return self.grid.array(name)

    def arrays(self) -> list[np.ndarray]:
        """The 25 arrays in C-ABI order."""
        return [self.array(n) for n in ARRAY_NAMES]

    def host_pointers(self) -> list:
        return [_ptr(a) for a in self.arrays()]

    def copy(self) -> "Model":
        """Deep copy of the host state (new Field objects, no device slots)."""
        other = object.__new__(Model)
        other.cfg = self.cfg
        other.grid = self.grid
        other.fields = {n: Field(n, f.data.copy()) for n, f in self.fields.items()}
        return other

    def checksum(self, name: str) -> float:
        """field_checksum: sum of |x| over the internal T region."""
        g = self.grid
        a = self.fields[name].get_data()
        return setup_lib().nemo_field_checksum(_ptr(a), g.W, g.xs, g.xe, g.ys, g.ye)


def make_model(cfg: NemoConfig) -> Model:
    return Model(cfg)
