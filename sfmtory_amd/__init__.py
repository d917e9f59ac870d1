"""MI355X-native Schur-complement Levenberg-Marquardt bundle adjustment.

Drop-in for the `sfm-ba` crate of sfmtory: the C ABI in include/sfmba.h
mirrors the crate's public API, and this package mirrors it in Python over
that ABI.  All compute runs in hand-written HIP kernels for gfx950.
"""
from .ba import (  # noqa: F401
    BaInput,
    BaOutput,
    BaParams,
    CameraModel,
    IntrinsicsRefineParams,
    Observation,
    Pose,
    Problem,
    RobustLoss,
    bundle_adjust,
    bundle_adjust_batch,
    cameras_from_array,
    cameras_to_array,
    default_fixed_params_mask,
    device_count,
    filter_and_reoptimize,
    mean_reprojection_error,
    observation_jacobians,
    observations_to_array,
    poses_to_array,
    refine_intrinsics,
    run_bundle_adjustment,
    total_cost,
)
from .colmap import refine_model, rewrite_model  # noqa: F401
from .scope import Reconstruction, Window, bundle_adjust_scope, bundle_adjust_scope_batch  # noqa: F401
