"""The reference driver's files: namelist in, results and mask out.

* :func:`read_namelist` -- the ``&namctl`` group of a ``namelist`` file with
  the reference's defaults (common/gocean2d_io_mod.f90:21-80), as a
  :class:`~.model.NemoConfig` plus the output interval ``irecord``
* :func:`write_results` -- ``go2d_<istp>_<rank>.dat`` (model_write,
  gocean2d_io_mod.f90:98-149), written by ``nemo_write_results``
* :func:`dump_tmask` -- ``tmask_<rank>.dat`` (model_mod.f90:251-291),
  written by ``nemo_dump_tmask``

The files are byte-compatible with the reference's (Fortran ``6e16.7``
records and an empty record after each row), so tools such as
manual_versions/regent/test_result.py read them unchanged. The T-point
coordinates follow dl_esm_inf's grid as restated by the Regent version
(manual_versions/regent/model_init.rg:77-80): x = (ji - 1) dx for the 1-based
column ji, so the 0-based column times dx.
"""
from __future__ import annotations

import ctypes as C
import re
from dataclasses import replace
from pathlib import Path

import numpy as np

from ._native import setup_lib
from .model import Model, NemoConfig

# read_namelist's defaults (gocean2d_io_mod.f90:58-69)
NAMELIST_DEFAULTS = {
    "jpiglo": 50, "jpjglo": 100, "jphgr_msh": 1, "dx": 1000.0, "dy": 1000.0,
    "dep_const": 100.0, "nit000": 1, "nitend": 1000, "irecord": 1, "rdt": 10.0,
    "cbfr": 0.001, "visc": 100.0,
}
_INTS = {"jpiglo", "jpjglo", "jphgr_msh", "nit000", "nitend", "irecord"}


def _fortran_number(tok: str, integer: bool):
    t = tok.strip().rstrip(",")
    if integer:
        return int(t)
    return float(re.sub(r"[dD]", "e", t))


def parse_namctl(text: str) -> dict:
    """Values of the first ``&namctl`` group in `text` over the defaults.
    Comments start with ``!``; names are case-insensitive; the group ends at
    ``/``. An unknown name raises ValueError, as Fortran's read does."""
    body = []
    for line in text.splitlines():
        body.append(line.split("!", 1)[0])
    s = "\n".join(body)
    m = re.search(r"&namctl\b(.*?)(?:^|\s)/", s, re.IGNORECASE | re.DOTALL | re.MULTILINE)
    if not m:
        raise ValueError("no &namctl group")
    vals = dict(NAMELIST_DEFAULTS)
    for name, tok in re.findall(r"(\w+)\s*=\s*([^\s,/]+)", m.group(1)):
        key = name.lower()
        if key not in vals:
            raise ValueError(f"unknown namelist entry {name}")
        vals[key] = _fortran_number(tok, key in _INTS)
    return vals


def read_namelist(path="namelist", base: NemoConfig | None = None) -> tuple[NemoConfig, int]:
    """(config, irecord) from a namelist file. Fields the namelist does not
    carry (tide, bathymetry, alignment) come from `base` (default: the
    reference's south-open channel and tide literals)."""
    vals = parse_namctl(Path(path).read_text())
    cfg = replace(base or NemoConfig(), jpiglo=vals["jpiglo"], jpjglo=vals["jpjglo"],
                  jphgr_msh=vals["jphgr_msh"], dx=vals["dx"], dy=vals["dy"],
                  dep_const=vals["dep_const"], nit000=vals["nit000"],
                  nitend=vals["nitend"], rdt=vals["rdt"], cbfr=vals["cbfr"],
                  visc=vals["visc"])
    return cfg, vals["irecord"]


def fortran_e(x: float, w: int = 16, d: int = 7) -> str:
    """One Fortran Ew.d field as the library writes it."""
    buf = C.create_string_buffer(w)
    lib = setup_lib()
    lib.nemo_fortran_e(C.c_double(x), w, d, buf)
    return buf.raw[:w].decode()


def results_name(istp: int, rank: int = 1) -> str:
    """go2d_<istp>_<rank>.dat: '(I5.5,"_",I5.5)' (gocean2d_io_mod.f90:124)."""
    return f"go2d_{istp:05d}_{rank:05d}.dat"


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def write_results(model: Model, istp: int, fields: dict | None = None, rank: int = 1,
                  directory=".") -> Path:
    """go2d_<istp>_<rank>.dat of the model's current state. `fields`
    (sshn_t, un, vn as W x H arrays) replaces the host arrays, e.g. with
    values read back from the device."""
    g = model.grid
    arr = {k: np.ascontiguousarray((fields or {}).get(k, model.array(k)), np.float64)
           for k in ("ht", "sshn_t", "un", "vn")}
    for k, a in arr.items():
        if a.shape != (g.H, g.W):
            raise ValueError(f"{k}: shape {a.shape}, the grid is {(g.H, g.W)}")
    path = Path(directory) / results_name(istp, rank)
    rc = setup_lib().nemo_write_results(str(path).encode(), _dp(arr["ht"]), _dp(arr["sshn_t"]),
                                        _dp(arr["un"]), _dp(arr["vn"]), g.W, g.xs, g.xe,
                                        g.ys, g.ye, model.cfg.dx, model.cfg.dy)
    if rc != 0:
        raise OSError(f"cannot write {path}")
    return path


def dump_tmask(model: Model, raw: bool = False, rank: int = 1, directory=".") -> Path:
    """tmask_<rank>.dat ('(I5.5)' rank, model_mod.f90:261-263)."""
    g = model.grid
    tm = np.ascontiguousarray(model.array("tmask"), np.int32)
    path = Path(directory) / f"tmask_{rank:05d}.dat"
    rc = setup_lib().nemo_dump_tmask(str(path).encode(),
                                     tm.ctypes.data_as(C.POINTER(C.c_int)), g.W, g.M, g.N,
                                     g.xs, g.xe, g.ys, g.ye, model.cfg.dx, model.cfg.dy,
                                     1 if raw else 0)
    if rc != 0:
        raise OSError(f"cannot write {path}")
    return path
