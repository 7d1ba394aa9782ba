"""shrinkers_amd -- MI355X (gfx950) Stieltjes-transform hot path of shrinkers.

The package mirrors the reference's ``src/stieltjes`` method API and its
``stieltjes_transform*`` Python bindings; every evaluation runs in
``libshrinkers_hip.so`` (hand-written HIP kernels behind a C ABI, see
``include/shrinkers_stieltjes.h``).  There is no CPU implementation: if the
library is missing, the first call raises.
"""
from .stieltjes import (  # noqa: F401
    CHEB_PRESETS,
    GPU_METHODS,
    METHODS,
    ChebCodeBatch,
    all_points_dev,
    at_points_dev,
    chebcode_all_points_ws_dev,
    chebcode_at_points_ws_dev,
    chebcode_workspace_bytes,
    chebcode_workspace_status,
    compute_all_stieltjes,
    compute_all_stieltjes_f32,
    compute_all_stieltjes_with_deriv,
    compute_stieltjes_at_points,
    compute_stieltjes_scaled_ray,
    default_eta,
    default_eta_bulk,
    parse_method,
    stieltjes_transform,
    stieltjes_transform_with_deriv,
    symmetric_range,
)
from ._lib import ShkError  # noqa: F401

__version__ = "0.1.0"
