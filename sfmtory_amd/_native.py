"""ctypes view of include/sfmba.h and loading of the HIP engine library."""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

LIB_DIR = Path(__file__).resolve().parent / "lib"
ENGINE_PATH = LIB_DIR / "libsfmba.so"
SCENE_PATH = LIB_DIR / "libsfmscene.so"

SFMBA_OK = 0
SFMBA_ERR_INVALID_ARGUMENT = -1
SFMBA_ERR_DEVICE = -2
SFMBA_ERR_OUT_OF_MEMORY = -3
SFMBA_ERR_UNSUPPORTED = -4

NOALLOC_MAX_OBSERVATIONS = 1 << 26

PRECISION_FP64 = 0
MIXED_MIN_LAMBDA = 1e-5  # SFMBA_MIXED_MIN_LAMBDA
PRECISION_MIXED = 1
SOLVER_CHOLESKY = 0
SOLVER_PCG = 1
SOLVER_PCG_MIXED = 2
INTRINSICS_FIXED = 0
INTRINSICS_FREE = 1
INTRINSICS_FREE_GUARDED = 2

CAMERA_DTYPE = np.dtype([("model_id", "<i4"), ("reserved", "<i4"), ("params", "<f8", (8,))])
POSE_DTYPE = np.dtype([("qw", "<f8"), ("qx", "<f8"), ("qy", "<f8"), ("qz", "<f8"),
                       ("tx", "<f8"), ("ty", "<f8"), ("tz", "<f8")])
OBSERVATION_DTYPE = np.dtype([("image_idx", "<u8"), ("point_idx", "<u8"), ("x", "<f8"), ("y", "<f8")])


class Camera(C.Structure):
    _fields_ = [("model_id", C.c_int32), ("reserved", C.c_int32), ("params", C.c_double * 8)]


class Pose(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("qw", "qx", "qy", "qz", "tx", "ty", "tz")]


class Observation(C.Structure):
    _fields_ = [("image_idx", C.c_uint64), ("point_idx", C.c_uint64), ("x", C.c_double), ("y", C.c_double)]


class Input(C.Structure):
    _fields_ = [
        ("num_images", C.c_size_t),
        ("num_cameras", C.c_size_t),
        ("num_points", C.c_size_t),
        ("num_observations", C.c_size_t),
        ("camera_of_image", C.c_void_p),
        ("num_camera_of_image", C.c_size_t),
        ("cameras", C.c_void_p),
        ("poses", C.c_void_p),
        ("points", C.c_void_p),
        ("observations", C.c_void_p),
        ("fixed_poses", C.c_void_p),
        ("num_fixed_poses", C.c_size_t),
        ("fixed_cameras", C.c_void_p),
        ("num_fixed_cameras", C.c_size_t),
        ("fixed_camera_params", C.c_void_p),
        ("fixed_camera_params_len", C.c_void_p),
        ("num_fixed_camera_params", C.c_size_t),
    ]


class Params(C.Structure):
    _fields_ = [("max_iterations", C.c_uint64), ("initial_lambda", C.c_double),
                ("robust_loss", C.c_int32), ("reserved", C.c_int32), ("loss_scale_px", C.c_double)]


class Output(C.Structure):
    _fields_ = [("poses", C.c_void_p), ("points", C.c_void_p), ("cameras", C.c_void_p),
                ("initial_cost", C.c_double), ("final_cost", C.c_double), ("iterations_run", C.c_uint64)]


class Options(C.Structure):
    _fields_ = [("precision", C.c_int32), ("solver", C.c_int32), ("device", C.c_int32),
                ("profile", C.c_int32), ("refine_steps", C.c_int32), ("pcg_max_iterations", C.c_int32),
                ("pcg_tolerance", C.c_double)]


class Stats(C.Structure):
    _fields_ = [
        ("outer_costs", C.c_void_p), ("outer_costs_capacity", C.c_uint64), ("outer_costs_len", C.c_uint64),
        ("trial_costs", C.c_void_p), ("trial_costs_capacity", C.c_uint64), ("trial_costs_len", C.c_uint64),
        ("dim", C.c_uint64), ("num_linearizations", C.c_uint64), ("num_factorizations", C.c_uint64),
        ("num_cholesky_failures", C.c_uint64), ("num_lu_solves", C.c_uint64),
        ("num_rejected_steps", C.c_uint64), ("num_schur_pairs", C.c_uint64),
        ("time_total_s", C.c_double), ("time_upload_s", C.c_double), ("time_linearize_s", C.c_double),
        ("time_schur_s", C.c_double), ("time_factor_s", C.c_double), ("time_solve_s", C.c_double),
        ("time_update_s", C.c_double), ("count_linearize", C.c_uint64), ("count_schur", C.c_uint64),
        ("count_factor", C.c_uint64), ("time_trailing_update_s", C.c_double),
        ("count_trailing_update", C.c_uint64), ("flop_trailing_update", C.c_double),
        ("num_pcg_iterations", C.c_uint64),
        ("try_lambdas", C.c_void_p), ("try_outcomes", C.c_void_p), ("tries_capacity", C.c_uint64),
        ("tries_len", C.c_uint64), ("num_fp32_tries", C.c_uint64),
        ("time_linearize_fp32_s", C.c_double), ("time_schur_fp32_s", C.c_double),
        ("count_linearize_fp32", C.c_uint64), ("count_schur_fp32", C.c_uint64),
    ]


TRY_ACCEPTED = 0
TRY_REJECTED = 1
TRY_NO_STEP = 2
TRY_LU = 4
TRY_FP32 = 8


class Policy(C.Structure):
    _fields_ = [("intrinsics", C.c_int32), ("robust_loss", C.c_int32), ("max_reprojection_error_px", C.c_double),
                ("camera_caller_fixed", C.c_void_p), ("num_camera_caller_fixed", C.c_size_t)]


class PolicyReport(C.Structure):
    _fields_ = [("chose_free", C.c_int32), ("distortion_plausible", C.c_int32),
                ("fixed_mean_error_px", C.c_double), ("free_mean_error_px", C.c_double),
                ("observations_kept_fixed", C.c_uint64), ("observations_kept_free", C.c_uint64),
                ("filter_rounds_fixed", C.c_uint64), ("filter_rounds_free", C.c_uint64)]


class IntrinsicsParams(C.Structure):
    _fields_ = [("max_iterations", C.c_uint64), ("initial_lambda", C.c_double),
                ("robust_loss", C.c_int32), ("reserved", C.c_int32), ("loss_scale_px", C.c_double)]


class ColmapRefineOptions(C.Structure):
    _fields_ = [("refine_intrinsics", C.c_int32), ("robust_loss", C.c_int32), ("max_iterations", C.c_uint64)]


class ColmapRefineReport(C.Structure):
    _fields_ = [("num_images", C.c_uint64), ("num_points3d", C.c_uint64), ("num_observations", C.c_uint64),
                ("initial_cost", C.c_double), ("final_cost", C.c_double), ("iterations_run", C.c_uint64),
                ("mean_reprojection_error_px_before", C.c_double),
                ("mean_reprojection_error_px_after", C.c_double)]


class Reconstruction(C.Structure):
    _fields_ = [("num_images", C.c_size_t), ("registered", C.c_void_p), ("poses", C.c_void_p),
                ("image_camera_id", C.c_void_p), ("pose_pinned", C.c_void_p), ("keypoint_offsets", C.c_void_p),
                ("keypoints", C.c_void_p), ("num_points", C.c_size_t), ("points", C.c_void_p),
                ("track_offsets", C.c_void_p), ("track_images", C.c_void_p), ("track_keypoints", C.c_void_p),
                ("num_cameras", C.c_size_t), ("camera_ids", C.c_void_p), ("cameras", C.c_void_p),
                ("caller_fixed_camera_ids", C.c_void_p), ("num_caller_fixed_camera_ids", C.c_size_t),
                ("seed_image", C.c_uint64)]


class WindowSummary(C.Structure):
    _fields_ = [("num_images", C.c_uint64), ("num_free_images", C.c_uint64), ("num_cameras", C.c_uint64),
                ("num_points", C.c_uint64), ("num_observations", C.c_uint64), ("initial_cost", C.c_double),
                ("final_cost", C.c_double), ("iterations_run", C.c_uint64)]


MAX_WINDOW_SHAPES = 8


class BatchInfo(C.Structure):
    _fields_ = [("threads_per_window", C.c_int32), ("pairs_in_flight", C.c_int32), ("tuned", C.c_int32),
                ("num_shapes", C.c_int32), ("shape_threads", C.c_int32 * MAX_WINDOW_SHAPES),
                ("shape_pairs", C.c_int32 * MAX_WINDOW_SHAPES), ("shape_seconds", C.c_double * MAX_WINDOW_SHAPES),
                ("create_seconds", C.c_double), ("capture_seconds", C.c_double), ("num_windows", C.c_uint64),
                ("num_large", C.c_uint64)]


SCOPE_GLOBAL = 0
SCOPE_LOCAL = 1
LOCAL_BA_NUM_IMAGES = 6

# Every symbol include/sfmba.h declares, with its ctypes signature.
SIGNATURES = {
    "sfmba_params_default": (None, [C.POINTER(Params)]),
    "sfmba_bundle_adjust": (C.c_int, [C.POINTER(Input), C.POINTER(Params), C.POINTER(Output)]),
    "sfmba_options_default": (None, [C.POINTER(Options)]),
    "sfmba_bundle_adjust_ex": (C.c_int, [C.POINTER(Input), C.POINTER(Params), C.POINTER(Options),
                                         C.POINTER(Output), C.POINTER(Stats)]),
    "sfmba_bundle_adjust_batch": (C.c_int, [C.POINTER(Input), C.c_size_t, C.POINTER(Params),
                                            C.POINTER(Options), C.POINTER(Output), C.POINTER(Stats)]),
    "sfmba_batch_create": (C.c_int, [C.POINTER(Input), C.c_size_t, C.POINTER(Params), C.POINTER(Options), C.c_int32,
                                     C.POINTER(C.c_void_p)]),
    "sfmba_batch_solve": (C.c_int, [C.c_void_p, C.POINTER(Output), C.POINTER(Stats)]),
    "sfmba_batch_get_info": (C.c_int, [C.c_void_p, C.POINTER(BatchInfo)]),
    "sfmba_batch_destroy": (None, [C.c_void_p]),
    "sfmba_problem_create": (C.c_int, [C.POINTER(Input), C.c_int32, C.POINTER(C.c_void_p)]),
    "sfmba_problem_solve": (C.c_int, [C.c_void_p, C.POINTER(Params), C.POINTER(Options), C.POINTER(Output),
                                      C.POINTER(Stats)]),
    "sfmba_problem_destroy": (None, [C.c_void_p]),
    "sfmba_mean_reprojection_error": (C.c_int, [C.POINTER(Input), C.POINTER(C.c_double)]),
    "sfmba_total_cost": (C.c_int, [C.POINTER(Input), C.POINTER(Params), C.POINTER(C.c_double)]),
    "sfmba_total_cost_noalloc": (C.c_int, [C.POINTER(Input), C.POINTER(Params), C.c_void_p,
                                           C.POINTER(C.c_double)]),
    "sfmba_default_fixed_params_mask": (C.c_int, [C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p]),
    "sfmba_intrinsics_params_default": (None, [C.POINTER(IntrinsicsParams)]),
    "sfmba_refine_intrinsics": (C.c_int, [C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_void_p,
                                          C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t,
                                          C.POINTER(IntrinsicsParams)]),
    "sfmba_observation_jacobians": (C.c_int, [C.POINTER(Input), C.c_int32, C.c_int32, C.c_void_p,
                                              C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "sfmba_dense_solve": (C.c_int, [C.c_void_p, C.c_size_t, C.c_void_p, C.c_int32, C.c_int32, C.c_void_p,
                                    C.POINTER(C.c_int32)]),
    "sfmba_filter_and_reoptimize": (C.c_int, [C.POINTER(Input), C.POINTER(Params), C.POINTER(Options), C.c_double,
                                              C.POINTER(Output), C.POINTER(C.c_uint64)]),
    "sfmba_policy_default": (None, [C.POINTER(Policy)]),
    "sfmba_run_bundle_adjustment": (C.c_int, [C.POINTER(Input), C.POINTER(Policy), C.POINTER(Options),
                                              C.POINTER(Output), C.POINTER(PolicyReport)]),
    "sfmba_window_create": (C.c_int, [C.POINTER(Reconstruction), C.c_int32, C.c_uint64, C.POINTER(C.c_void_p)]),
    "sfmba_window_input": (C.c_int, [C.c_void_p, C.POINTER(Input)]),
    "sfmba_window_indices": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_size_t),
                                       C.POINTER(C.c_void_p), C.POINTER(C.c_size_t)]),
    "sfmba_window_scatter": (C.c_int, [C.c_void_p, C.POINTER(Output), C.POINTER(Reconstruction)]),
    "sfmba_window_destroy": (None, [C.c_void_p]),
    "sfmba_reconstruction_bundle_adjust": (C.c_int, [C.POINTER(Reconstruction), C.c_int32, C.c_uint64,
                                                     C.POINTER(Policy), C.POINTER(Options),
                                                     C.POINTER(WindowSummary), C.POINTER(PolicyReport)]),
    "sfmba_reconstruction_bundle_adjust_batch": (C.c_int, [C.POINTER(Reconstruction), C.c_size_t, C.c_void_p,
                                                           C.c_void_p, C.POINTER(Params), C.POINTER(Options),
                                                           C.c_void_p]),
    "sfmba_colmap_refine_options_default": (None, [C.POINTER(ColmapRefineOptions)]),
    "sfmba_colmap_refine": (C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(ColmapRefineOptions), C.POINTER(Options),
                                      C.POINTER(ColmapRefineReport)]),
    "sfmba_colmap_rewrite": (C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(C.c_uint64)]),
    "sfmba_device_count": (C.c_int, []),
    "sfmba_last_error": (C.c_char_p, []),
}

_engine = None


def load_engine() -> C.CDLL:
    """Load the HIP engine.  Raises if it has not been built: there is no
    fallback implementation."""
    global _engine
    if _engine is None:
        if not ENGINE_PATH.exists():
            raise RuntimeError(f"{ENGINE_PATH} is missing; run `python -m sfmtory_amd.build`")
        lib = C.CDLL(str(ENGINE_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _engine = lib
    return _engine


class EngineError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"sfmba error {code}: {message}")
        self.code = code


def check(code: int) -> None:
    if code != SFMBA_OK:
        msg = load_engine().sfmba_last_error().decode(errors="replace")
        if code == SFMBA_ERR_INVALID_ARGUMENT:
            # the Rust API panics on out-of-range indices; surface it as IndexError/ValueError
            if "out of range" in msg:
                raise IndexError(msg)
            raise ValueError(msg)
        raise EngineError(code, msg)


def ptr(a: np.ndarray | None) -> int | None:
    if a is None or a.size == 0:
        return None
    return a.ctypes.data
