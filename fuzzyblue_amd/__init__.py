"""fuzzyblue_amd — MI355X-native precomputed atmospheric scattering.

Public API mirrors the reference crate root (src/lib.rs:8-12): Builder,
Atmosphere, PendingAtmosphere, Parameters, Renderer, DrawParameters.  The
compute path is libfuzzyblue_hip.so (HIP kernels for gfx950 behind the C ABI
in include/fuzzyblue_hip.h); importing this package does not load it, the
first API call does, and fails loudly if it is missing.
"""
from .params import DensityProfile, DensityProfileLayer, Parameters  # noqa: F401
from .api import (  # noqa: F401
    Atmosphere,
    Builder,
    DeviceBuffer,
    DrawParameters,
    PendingAtmosphere,
    Renderer,
    Stream,
    device_count,
    ray_direction,
    synchronize,
)
