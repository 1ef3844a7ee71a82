"""Learned coarse-grid closure (no reference counterpart).

A coarse NEMOLite2D model (the fine grid coarsened by `factor`: f times the
spacing and the time step) is stepped with the plain FP64 step; after each
coarse step a small convolutional network (``nemo_cl_*``, 4 layers of 5x5
convolutions, 32 channels, GELU, bf16 MFMA with FP32 accumulation) adds
corrections to u, v and ssh that move the coarse state toward the box
filter of a fine FP64 run. An energy check on the device keeps the plain
step whenever a correction would be non-finite or raise the energy by more
than a set fraction in one step.

* :func:`coarse_config`, :func:`coarse_model` -- the coarse model of a fine
  configuration: spacing and time step times f, depth = f x f box filter
  of the fine depth
* :class:`ClosureNet` -- the network, its optimizer and the fine <-> coarse
  plumbing, over device pointers of :class:`~.psy.TimeStepper` models
"""
from __future__ import annotations

import ctypes as C
from dataclasses import replace

import numpy as np

from ._native import hip_lib
from .model import Model, NemoConfig
from .psy import TimeStepper

TAPS = 25
LAYERS = ((16, 32), (32, 32), (32, 32), (32, 16))   # (cin, cout) as stored
USED_IN, USED_OUT = 4, 3
# forward FLOP per coarse point over the used channels:
# 25 taps x (4*32 + 32*32 + 32*32 + 32*3) multiply-adds x 2
FLOP_PER_POINT = 25 * (4 * 32 + 32 * 32 + 32 * 32 + 32 * 3) * 2


def layer_offsets() -> list[int]:
    off, out = 0, []
    for cin, cout in LAYERS:
        out.append(off)
        off += TAPS * cin * cout + cout
    out.append(off)
    return out


def init_params(seed: int) -> np.ndarray:
    """He-normal weights over the used channels (numpy PCG64 normals from
    `seed`), zero biases and padded channels; the output layer scaled by 0.1
    so that the first corrections are small."""
    rng = np.random.default_rng(seed)
    offs = layer_offsets()
    p = np.zeros(offs[-1], np.float32)
    for l, (cin, cout) in enumerate(LAYERS):
        ci_used = USED_IN if l == 0 else cin
        co_used = USED_OUT if l == 3 else cout
        std = np.sqrt(2.0 / (TAPS * ci_used)) * (0.1 if l == 3 else 1.0)
        w = np.zeros((TAPS, cout, cin), np.float32)
        w[:, :co_used, :ci_used] = rng.standard_normal((TAPS, co_used, ci_used)) * std
        p[offs[l]:offs[l] + w.size] = w.ravel()
    return p


def coarse_config(fine: NemoConfig, factor: int) -> NemoConfig:
    nxf, nyf = fine.jpiglo - 2, fine.jpjglo - 2
    if nxf % factor or nyf % factor:
        raise ValueError(f"fine interior {nxf}x{nyf} not divisible by {factor}")
    return replace(fine, jpiglo=nxf // factor + 2, jpjglo=nyf // factor + 2,
                   dx=fine.dx * factor, dy=fine.dy * factor, rdt=fine.rdt * factor,
                   bathymetry=None, nitend=fine.nit000 + (fine.nsteps // factor) - 1)


def coarse_model(fine: Model, factor: int) -> Model:
    """Coarse model whose interior depth is the f x f mean of the fine
    interior depth; points outside the interior take the nearest interior
    value and hu, hv follow from ht as in the set-up library."""
    cfg = coarse_config(fine.cfg, factor)
    m = Model(cfg)
    g, gf = m.grid, fine.grid
    nx, ny = g.M - 2, g.N - 2
    htf = fine.array("ht")[gf.ys:gf.ye + 1, gf.xs:gf.xe + 1]
    htc = htf.reshape(ny, factor, nx, factor).mean(axis=(1, 3))
    jj = np.clip(np.arange(g.H) - 1, 0, ny - 1)
    ii = np.clip(np.arange(g.W) - 1, 0, nx - 1)
    ht = htc[jj[:, None], ii[None, :]]
    hu = ht.copy()
    hu[:, :-1] = 0.5 * (ht[:, :-1] + ht[:, 1:])
    hv = ht.copy()
    hv[:-1, :] = 0.5 * (ht[:-1, :] + ht[1:, :])
    np.copyto(m.array("ht"), ht)
    np.copyto(m.array("hu"), hu)
    np.copyto(m.array("hv"), hv)
    return m


def _state_ptrs(st: TimeStepper) -> tuple:
    """Device pointers of a coarse model after at least one step: ssh, u, v,
    face ssh (current buffers), ht, tmask, area_t/u/v."""
    f = st.model.fields
    gs = st.grid_slots.slots
    vals = (f["sshn_t"].device_ptr, f["un"].device_ptr, f["vn"].device_ptr,
            f["sshn_u"].device_ptr, f["sshn_v"].device_ptr, f["ht"].device_ptr,
            gs["tmask"], gs["area_t"], gs["area_u"], gs["area_v"])
    if not all(v.value for v in vals):
        raise RuntimeError("coarse model has no device state yet (step it once first)")
    return tuple(v.value for v in vals)


def set_conv_form(form: int) -> None:
    """Launch form of the 32 -> 32 convolutions for every network in the
    process (nemo_cl_set_conv_form): 0 two 32-pixel tiles per wavefront
    pass, 1 one tile with twice the wavefronts per workgroup. Same bits."""
    hip_lib().nemo_cl_set_conv_form(int(form))


def conv_form() -> int:
    return hip_lib().nemo_cl_conv_form()


class ClosureNet:
    """The network and its buffers for one coarse grid."""

    def __init__(self, coarse: Model, params: np.ndarray | None = None, seed: int = 0):
        lib = hip_lib()
        g = coarse.grid
        self.nx, self.ny = g.M - 2, g.N - 2
        p = init_params(seed) if params is None else np.ascontiguousarray(params, np.float32)
        if p.size != lib.nemo_cl_param_count():
            raise ValueError(f"{p.size} parameters, the network has {lib.nemo_cl_param_count()}")
        self.handle = lib.nemo_cl_create(self.nx, self.ny, g.W, g.H,
                                         p.ctypes.data_as(C.POINTER(C.c_float)))
        self.norm = None

    # ---- parameters and settings ------------------------------------------
    def params(self) -> np.ndarray:
        out = np.empty(hip_lib().nemo_cl_param_count(), np.float32)
        hip_lib().nemo_cl_get_params(self.handle, out.ctypes.data_as(C.POINTER(C.c_float)))
        return out

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        hip_lib().nemo_cl_set_params(self.handle, p.ctypes.data_as(C.POINTER(C.c_float)))

    def set_norm(self, in_shift, in_scale, out_scale) -> None:
        a = [np.ascontiguousarray(x, np.float64) for x in (in_shift, in_scale, out_scale)]
        self.norm = tuple(x.copy() for x in a)
        hip_lib().nemo_cl_set_norm(self.handle, *(x.ctypes.data_as(C.POINTER(C.c_double))
                                                  for x in a))

    def set_optimizer(self, kind: str = "adam", lr: float = 1e-3, beta1: float = 0.9,
                      beta2: float = 0.999, eps: float = 1e-8) -> None:
        hip_lib().nemo_cl_set_optimizer(self.handle, {"sgd": 0, "adam": 1}[kind], lr, beta1,
                                        beta2, eps)

    def set_guard(self, rel: float, abs_tol: float, g: float = 9.80665) -> None:
        hip_lib().nemo_cl_set_guard(self.handle, rel, abs_tol, g)

    # ---- fine <-> coarse ---------------------------------------------------
    def filter_fine(self, fine: TimeStepper, factor: int) -> None:
        f = fine.model.fields
        hip_lib().nemo_cl_filter(self.handle, f["sshn_t"].device_ptr, f["un"].device_ptr,
                                 f["vn"].device_ptr, fine.model.grid.W, factor)

    def filtered(self, name: str, shape) -> np.ndarray:
        out = np.empty(shape, np.float64)
        hip_lib().nemo_cl_read_filtered(self.handle, {"sshn_t": 0, "un": 1, "vn": 2}[name],
                                        out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def set_state(self, coarse: TimeStepper) -> None:
        hip_lib().nemo_cl_set_state(self.handle, *_state_ptrs(coarse))

    def train_step(self, coarse: TimeStepper) -> None:
        hip_lib().nemo_cl_train_step(self.handle, *_state_ptrs(coarse))

    def last_loss(self) -> float:
        return hip_lib().nemo_cl_last_loss(self.handle)

    def set_gate(self, eps: float) -> None:
        hip_lib().nemo_cl_set_gate(self.handle, eps)

    def set_guard_limit(self, max_out: float) -> None:
        """Reject corrections with any component beyond max_out normalised
        units as well (0: no bound)."""
        hip_lib().nemo_cl_set_guard_limit(self.handle, max_out)

    def set_ssh_cells(self, all_wet: bool) -> None:
        """Correct and learn the ssh of every wet cell (True) or only of the
        cells the step does not clamp to the tide (False)."""
        hip_lib().nemo_cl_set_ssh_cells(self.handle, 1 if all_wet else 0)

    def set_clip(self, max_norm: float) -> None:
        hip_lib().nemo_cl_set_clip(self.handle, max_norm)

    def set_loss_weights(self, wu: float, wv: float, wssh: float) -> None:
        hip_lib().nemo_cl_set_loss_weights(self.handle, wu, wv, wssh)

    def last_stats(self) -> dict:
        o = np.zeros(7, np.float64)
        hip_lib().nemo_cl_last_stats(self.handle, o.ctypes.data_as(C.POINTER(C.c_double)))
        return {"loss_u": o[0], "loss_v": o[1], "loss_ssh": o[2], "grad_norm": o[3],
                "energy_before": o[4], "energy_after": o[5], "nonfinite": o[6]}

    def correct(self, coarse: TimeStepper, guard: bool = True) -> None:
        hip_lib().nemo_cl_correct(self.handle, 1 if guard else 0, *_state_ptrs(coarse))

    def steps(self, coarse: TimeStepper, nsteps: int, guard: bool = True) -> None:
        """nsteps coarse steps each followed by the correction, as one graph
        replay (nemo_cl_steps); the same as alternating coarse.step() and
        correct(coarse, guard)."""
        slot = C.cast(C.byref(coarse.model.fields["ssha_t"].device_ptr),
                      C.POINTER(C.c_void_p))
        hip_lib().nemo_cl_steps(self.handle, slot, nsteps, coarse.istp, 1 if guard else 0)
        coarse.istp += nsteps

    def guard_trips(self) -> int:
        return hip_lib().nemo_cl_guard_trips(self.handle)

    def set_timing(self, on: bool) -> None:
        hip_lib().nemo_cl_set_timing(self.handle, 1 if on else 0)

    def forward_ms(self) -> float:
        return hip_lib().nemo_cl_forward_ms(self.handle)

    # ---- test hooks --------------------------------------------------------
    def forward_host(self, x4: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x4, np.float32)
        out = np.empty((self.ny, self.nx, 3), np.float32)
        hip_lib().nemo_cl_forward_host(self.handle, x.ctypes.data_as(C.POINTER(C.c_float)),
                                       out.ctypes.data_as(C.POINTER(C.c_float)))
        return out

    def grad_host(self, x4: np.ndarray, target: np.ndarray) -> tuple[float, np.ndarray]:
        x = np.ascontiguousarray(x4, np.float32)
        t = np.ascontiguousarray(target, np.float32)
        g = np.empty(hip_lib().nemo_cl_param_count(), np.float32)
        loss = hip_lib().nemo_cl_grad_host(self.handle, x.ctypes.data_as(C.POINTER(C.c_float)),
                                           t.ctypes.data_as(C.POINTER(C.c_float)),
                                           g.ctypes.data_as(C.POINTER(C.c_float)))
        return loss, g

    def train_host(self, x4: np.ndarray, target: np.ndarray) -> float:
        x = np.ascontiguousarray(x4, np.float32)
        t = np.ascontiguousarray(target, np.float32)
        return hip_lib().nemo_cl_train_host(self.handle, x.ctypes.data_as(C.POINTER(C.c_float)),
                                            t.ctypes.data_as(C.POINTER(C.c_float)))

    def release(self) -> None:
        if self.handle:
            hip_lib().nemo_cl_destroy(self.handle)
            self.handle = None
