"""psyclonebench_amd -- the NEMOLite2D time step of PSycloneBench on one
AMD Instinct MI355X.

The product is the native library ``libnemo_hip.so`` (hand-written HIP for
gfx950 behind the reference PSy layer's C ABI, declared in
``include/nemolite2d_hip.h``). This package builds it, sets up models in the
dl_esm_inf layout and drives the C ABI from Python.
"""
from .model import (  # noqa: F401
    Bathymetry, Field, Grid, Model, NemoConfig, c1_config, c2_config,
    make_model, reference_config,
)
from .psy import TimeStepper, invoke_time_step, invoke_time_step_host  # noqa: F401

__all__ = [
    "Bathymetry", "Field", "Grid", "Model", "NemoConfig", "TimeStepper",
    "c1_config", "c2_config", "invoke_time_step", "invoke_time_step_host",
    "make_model", "reference_config",
]
