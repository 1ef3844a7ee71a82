"""Shared test helpers: golden fixtures, models rebuilt from fixture inputs,
comparison utilities."""
from __future__ import annotations

import json
from functools import lru_cache
from pathlib import Path

import numpy as np

from psyclonebench_amd.model import Bathymetry, Model, reference_config

GOLDEN = Path(__file__).resolve().parent / "golden"

# the reference scalar arguments used by every fixture: rdt and cbfr come
# from the case; visc 0.1; omega, d2r, g from physical_params_mod.f90
SCALAR_NAMES = ("rdt", "cbfr", "visc", "omega", "d2r", "g")


@lru_cache(maxsize=None)
def manifest() -> dict:
    return json.loads((GOLDEN / "manifest.json").read_text())


@lru_cache(maxsize=None)
def load_npz(name: str) -> dict:
    with np.load(GOLDEN / name) as z:
        return {k: z[k] for k in z.files}


def step_case_names() -> list[str]:
    return list(manifest()["steps"])


def model_for_case(name: str, nsteps: int | None = None) -> Model:
    """The model of a multi-step fixture, with tmask/ht/hu/hv taken from the
    fixture (so the check does not depend on the set-up library)."""
    c = manifest()["steps"][name]
    cfg = reference_config(c["n"], nsteps or max(c["checkpoints"]), mask=c["mask"],
                           cbfr=c["cbfr"])
    if c.get("bathy"):
        cfg.bathymetry = Bathymetry(**c["bathy"])
    m = Model(cfg)
    fx = load_npz(f"steps_{name}.npz")
    for k in ("tmask", "ht", "hu", "hv"):
        np.copyto(m.array(k), fx[f"in_{k}"])
    return m


def kernel_fixture() -> tuple[dict, dict]:
    return manifest()["kernels"], load_npz("kernels.npz")


def kernel_inputs(mask_name: str) -> dict[str, np.ndarray]:
    _, z = kernel_fixture()
    pre = f"{mask_name}/in/"
    return {k[len(pre):]: v.copy() for k, v in z.items() if k.startswith(pre)}


def live_state_model(n: int, seed: int, random_mask: bool = False) -> Model:
    """An n^2 model whose 24 FP64 arrays are seeded random and nonzero over
    the whole allocation (padding included), in the value ranges of the
    per-kernel fixtures: ssh and velocities in [-1, 1], depths in [50, 150],
    areas in [5e5, 2e6], lengths in [500, 1500], latitudes in [-80, 80].
    Exact +-0 velocities exercise the upwind selectors. The mask is the
    open-west one of the 4096^2 workload, or (random_mask) seeded {-1, 0, 1}
    with no wet cell on row 0, column 0 or the last T row (which the fused
    form requires)."""
    from psyclonebench_amd.model import c2_config
    m = Model(c2_config(n, 1, seed=0))
    g = m.grid
    rng = np.random.default_rng(seed)
    ranges = dict(ssha_t=(-1, 1), sshn_t=(-1, 1), sshn_u=(-1, 1), sshn_v=(-1, 1),
                  hu=(50, 150), hv=(50, 150), un=(-1, 1), vn=(-1, 1), ua=(-1, 1),
                  ht=(50, 150), ssha_u=(-1, 1), va=(-1, 1), ssha_v=(-1, 1),
                  area_t=(5e5, 2e6), area_u=(5e5, 2e6), area_v=(5e5, 2e6),
                  dx_u=(500, 1500), dx_v=(500, 1500), dx_t=(500, 1500),
                  dy_u=(500, 1500), dy_v=(500, 1500), dy_t=(500, 1500),
                  gphiu=(-80, 80), gphiv=(-80, 80))
    for name, (lo, hi) in ranges.items():
        a = m.array(name)
        a[:] = rng.uniform(lo, hi, size=a.shape)
    m.array("un")[g.N // 2, 100:400] = 0.0
    m.array("vn")[g.N // 3, 200:500] = -0.0
    if random_mask:
        t = g.tmask
        t[:] = 0
        r = rng.choice(np.array([-1, 0, 1], np.int32), size=(g.N, g.M),
                       p=[0.1, 0.1, 0.8])
        r[0, :] = np.minimum(r[0, :], 0)
        r[:, 0] = np.minimum(r[:, 0], 0)
        r[g.N - 1, :] = np.minimum(r[g.N - 1, :], 0)
        t[:g.N, :g.M] = r
    return m


def assert_bitwise(a: np.ndarray, b: np.ndarray, what: str = "") -> None:
    """Equal as bit patterns (so -0.0 != 0.0 and NaN payloads count)."""
    assert a.shape == b.shape, (what, a.shape, b.shape)
    va, vb = a.view(np.uint64) if a.dtype == np.float64 else a, \
        b.view(np.uint64) if b.dtype == np.float64 else b
    bad = np.flatnonzero(va.ravel() != vb.ravel())
    if bad.size:
        k = bad[0]
        raise AssertionError(
            f"{what}: {bad.size} points differ; first at flat index {k}: "
            f"{a.ravel()[k]!r} vs {b.ravel()[k]!r}")


def rel_err(a: np.ndarray, ref: np.ndarray) -> tuple[float, float]:
    """(max|a-ref| / max|ref|, ||a-ref||_2 / ||ref||_2)."""
    d = a.astype(np.float64) - ref.astype(np.float64)
    mref = float(np.max(np.abs(ref))) or 1.0
    nref = float(np.linalg.norm(ref)) or 1.0
    return float(np.max(np.abs(d))) / mref, float(np.linalg.norm(d)) / nref
