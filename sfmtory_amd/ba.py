"""Python mirror of the `sfm-ba` crate's public API over the HIP engine.

Names, argument meaning and error behaviour follow the Rust crate
(crates/sfm-ba/src/lib.rs:56-96, :213-247, :708, :747; intrinsics.rs:49, :66,
:116; sfm-core camera.rs:21-163, pose.rs:15-48):

- ``bundle_adjust(BaInput, BaParams) -> BaOutput`` consumes nothing of the
  caller's (the input arrays are read only) and never raises for numerical
  trouble: a failed solve simply stops LM and returns the last accepted state.
- Out-of-range indices raise ``IndexError`` where the Rust code would panic.
- Every compute call runs on the GPU; without the built HIP library or a
  device the call raises (there is no CPU fallback).

Bulk data lives in numpy arrays with the C ABI's layouts (see
``_native.CAMERA_DTYPE`` etc.), so a 3M-observation problem is passed
without per-element Python objects.
"""
from __future__ import annotations

import ctypes as C
import enum
from dataclasses import dataclass, field
from typing import Iterable, Sequence

import numpy as np

from . import _native as N

# ----------------------------------------------------------------- models
MODEL_NAMES = ["SIMPLE_PINHOLE", "PINHOLE", "SIMPLE_RADIAL", "RADIAL", "OPENCV", "OPENCV_FISHEYE"]
MODEL_PARAM_COUNT = [3, 4, 4, 5, 8, 8]


class RobustLoss(enum.IntEnum):
    """crates/sfm-ba/src/lib.rs:213-218 (``None`` is spelled NONE here)."""

    NONE = 0
    HUBER = 1
    CAUCHY = 2


@dataclass(frozen=True)
class CameraModel:
    """crates/sfm-core/src/camera.rs:21-58: COLMAP model id + params in
    COLMAP order."""

    model_id: int
    params_: tuple

    @staticmethod
    def from_name_and_params(name: str, params: Sequence[float]) -> "CameraModel | None":
        if name not in MODEL_NAMES:
            return None
        mid = MODEL_NAMES.index(name)
        k = MODEL_PARAM_COUNT[mid]
        return CameraModel(mid, tuple(float(v) for v in params[:k]))

    @staticmethod
    def simple_pinhole(f, cx, cy):
        return CameraModel(0, (f, cx, cy))

    @staticmethod
    def pinhole(fx, fy, cx, cy):
        return CameraModel(1, (fx, fy, cx, cy))

    @staticmethod
    def simple_radial(f, cx, cy, k):
        return CameraModel(2, (f, cx, cy, k))

    @staticmethod
    def radial(f, cx, cy, k1, k2):
        return CameraModel(3, (f, cx, cy, k1, k2))

    @staticmethod
    def opencv(fx, fy, cx, cy, k1, k2, p1, p2):
        return CameraModel(4, (fx, fy, cx, cy, k1, k2, p1, p2))

    @staticmethod
    def opencv_fisheye(fx, fy, cx, cy, k1, k2, k3, k4):
        return CameraModel(5, (fx, fy, cx, cy, k1, k2, k3, k4))

    def name(self) -> str:
        return MODEL_NAMES[self.model_id]

    def params(self) -> list:
        return list(self.params_)


def cameras_to_array(cams) -> np.ndarray:
    if isinstance(cams, np.ndarray) and cams.dtype == N.CAMERA_DTYPE:
        return np.ascontiguousarray(cams)
    out = np.zeros(len(cams), dtype=N.CAMERA_DTYPE)
    for i, c in enumerate(cams):
        out[i]["model_id"] = c.model_id
        p = list(c.params_)
        out[i]["params"][: len(p)] = p
    return out


def cameras_from_array(arr: np.ndarray) -> list:
    return [CameraModel(int(r["model_id"]), tuple(float(v) for v in r["params"][: MODEL_PARAM_COUNT[int(r["model_id"])]]))
            for r in arr]


@dataclass
class Pose:
    """crates/sfm-core/src/pose.rs:15-48: world-to-camera, X_cam = R X + t,
    Hamilton quaternion stored (w, x, y, z)."""

    q: np.ndarray = field(default_factory=lambda: np.array([1.0, 0.0, 0.0, 0.0]))
    t: np.ndarray = field(default_factory=lambda: np.zeros(3))

    @staticmethod
    def identity() -> "Pose":
        return Pose()

    @staticmethod
    def from_quaternion_wxyz_translation(q, t) -> "Pose":
        q = np.asarray(q, dtype=np.float64)
        return Pose(q / np.linalg.norm(q), np.asarray(t, dtype=np.float64))

    def as_row(self) -> np.ndarray:
        return np.concatenate([self.q, self.t])


def poses_to_array(poses) -> np.ndarray:
    if isinstance(poses, np.ndarray):
        a = np.ascontiguousarray(poses, dtype=np.float64)
        return a.reshape(-1, 7)
    return np.ascontiguousarray(np.array([p.as_row() for p in poses], dtype=np.float64).reshape(-1, 7))


@dataclass
class Observation:
    """crates/sfm-ba/src/lib.rs:56-62."""

    image_idx: int
    point_idx: int
    x: float
    y: float


def observations_to_array(obs) -> np.ndarray:
    if isinstance(obs, np.ndarray) and obs.dtype == N.OBSERVATION_DTYPE:
        return np.ascontiguousarray(obs)
    out = np.zeros(len(obs), dtype=N.OBSERVATION_DTYPE)
    for i, o in enumerate(obs):
        out[i] = (o.image_idx, o.point_idx, o.x, o.y)
    return out


@dataclass
class BaInput:
    """crates/sfm-ba/src/lib.rs:64-96.  ``fixed_*`` sequences shorter than
    their index range mean "free" for the missing entries."""

    camera_of_image: Sequence[int]
    cameras: object
    poses: object
    points: object
    observations: object
    fixed_poses: Sequence[bool] = ()
    fixed_cameras: Sequence[bool] = ()
    fixed_camera_params: Sequence[Sequence[bool]] = ()


@dataclass
class BaParams:
    """crates/sfm-ba/src/lib.rs:220-238 (defaults)."""

    max_iterations: int = 50
    initial_lambda: float = 1e-3
    robust_loss: RobustLoss = RobustLoss.HUBER
    loss_scale_px: float = 2.0


@dataclass
class IntrinsicsRefineParams:
    """crates/sfm-ba/src/intrinsics.rs:65-82 (defaults)."""

    max_iterations: int = 30
    initial_lambda: float = 1e-3
    robust_loss: RobustLoss = RobustLoss.HUBER
    loss_scale_px: float = 2.0


@dataclass
class BaOutput:
    """crates/sfm-ba/src/lib.rs:240-247, plus the engine's diagnostics."""

    poses: np.ndarray
    points: np.ndarray
    cameras: np.ndarray
    initial_cost: float
    final_cost: float
    iterations_run: int
    stats: dict | None = None


# ---------------------------------------------------------- C conversion
class _CInput:
    """Owns the arrays a C `sfmba_input` points into."""

    def __init__(self, inp: BaInput):
        self.coi = np.ascontiguousarray(np.asarray(inp.camera_of_image, dtype=np.uint64))
        self.cams = cameras_to_array(inp.cameras)
        self.poses = poses_to_array(inp.poses)
        self.points = np.ascontiguousarray(np.asarray(inp.points, dtype=np.float64).reshape(-1, 3))
        self.obs = observations_to_array(inp.observations)
        self.fixed_poses = np.ascontiguousarray(np.asarray(list(inp.fixed_poses), dtype=np.uint8))
        self.fixed_cameras = np.ascontiguousarray(np.asarray(list(inp.fixed_cameras), dtype=np.uint8))
        masks = [np.asarray(list(m), dtype=np.uint8) for m in inp.fixed_camera_params]
        self.mask_len = np.ascontiguousarray(np.array([len(m) for m in masks], dtype=np.uint64))
        self.mask = np.ascontiguousarray(np.concatenate(masks) if masks else np.zeros(0, dtype=np.uint8))
        s = N.Input()
        s.num_images = self.poses.shape[0]
        s.num_cameras = self.cams.shape[0]
        s.num_points = self.points.shape[0]
        s.num_observations = self.obs.shape[0]
        s.camera_of_image = N.ptr(self.coi)
        s.num_camera_of_image = self.coi.shape[0]
        s.cameras = N.ptr(self.cams)
        s.poses = N.ptr(self.poses)
        s.points = N.ptr(self.points)
        s.observations = N.ptr(self.obs)
        s.fixed_poses = N.ptr(self.fixed_poses)
        s.num_fixed_poses = self.fixed_poses.shape[0]
        s.fixed_cameras = N.ptr(self.fixed_cameras)
        s.num_fixed_cameras = self.fixed_cameras.shape[0]
        s.fixed_camera_params = N.ptr(self.mask)
        s.fixed_camera_params_len = N.ptr(self.mask_len)
        s.num_fixed_camera_params = self.mask_len.shape[0]
        self.c = s


def _c_params(p: BaParams) -> N.Params:
    return N.Params(int(p.max_iterations), float(p.initial_lambda), int(p.robust_loss), 0, float(p.loss_scale_px))


PRECISIONS = {"fp64": N.PRECISION_FP64, "mixed": N.PRECISION_MIXED}
SOLVERS = {"cholesky": N.SOLVER_CHOLESKY, "pcg": N.SOLVER_PCG, "pcg_mixed": N.SOLVER_PCG_MIXED}


def _options(precision, solver, device, profile, refine_steps=None, pcg_max_iterations=None,
             pcg_tolerance=None) -> N.Options:
    lib = N.load_engine()
    o = N.Options()
    lib.sfmba_options_default(C.byref(o))
    o.precision = PRECISIONS[precision]
    o.solver = SOLVERS[solver]
    o.device = int(device)
    o.profile = 1 if profile else 0
    if refine_steps is not None:
        o.refine_steps = int(refine_steps)
    if pcg_max_iterations is not None:
        o.pcg_max_iterations = int(pcg_max_iterations)
    if pcg_tolerance is not None:
        o.pcg_tolerance = float(pcg_tolerance)
    return o


class _Outputs:
    def __init__(self, ci: _CInput):
        self.poses = np.zeros_like(ci.poses)
        self.points = np.zeros_like(ci.points)
        self.cams = ci.cams.copy()
        self.c = N.Output(N.ptr(self.poses), N.ptr(self.points), N.ptr(self.cams), 0.0, 0.0, 0)


class _Stats:
    def __init__(self, cap: int = 4096):
        self.outer = np.zeros(cap)
        self.trial = np.zeros(cap)
        self.lambdas = np.zeros(cap)
        self.outcomes = np.zeros(cap, dtype=np.uint8)
        s = N.Stats()
        s.outer_costs = self.outer.ctypes.data
        s.outer_costs_capacity = cap
        s.trial_costs = self.trial.ctypes.data
        s.trial_costs_capacity = cap
        s.try_lambdas = self.lambdas.ctypes.data
        s.try_outcomes = self.outcomes.ctypes.data
        s.tries_capacity = cap
        self.c = s

    _ARRAYS = ("outer_costs", "trial_costs", "outer_costs_capacity", "trial_costs_capacity", "try_lambdas",
               "try_outcomes", "tries_capacity")

    def to_dict(self) -> dict:
        """Counters and times, plus the traces: ``outer_costs``,
        ``trial_costs`` and, per damping try, ``try_lambdas`` and
        ``try_outcomes`` (SFMBA_TRY_* flags)."""
        s = self.c
        d = {name: getattr(s, name) for name, _ in N.Stats._fields_ if name not in self._ARRAYS}
        d["outer_costs"] = self.outer[: min(s.outer_costs_len, len(self.outer))].copy()
        d["trial_costs"] = self.trial[: min(s.trial_costs_len, len(self.trial))].copy()
        n = min(s.tries_len, len(self.lambdas))
        d["try_lambdas"] = self.lambdas[:n].copy()
        d["try_outcomes"] = self.outcomes[:n].copy()
        return d


def bundle_adjust(inp: BaInput, params: BaParams | None = None, *, precision: str = "fp64",
                  solver: str = "cholesky", device: int = 0, profile: bool = False,
                  refine_steps: int | None = None, pcg_max_iterations: int | None = None,
                  pcg_tolerance: float | None = None) -> BaOutput:
    """crates/sfm-ba/src/lib.rs:747 ``pub fn bundle_adjust``.  ``solver="pcg"``
    solves the reduced system by block-Jacobi PCG instead of factoring it."""
    params = params or BaParams()
    lib = N.load_engine()
    ci = _CInput(inp)
    outs = _Outputs(ci)
    st = _Stats()
    opt = _options(precision, solver, device, profile, refine_steps, pcg_max_iterations, pcg_tolerance)
    N.check(lib.sfmba_bundle_adjust_ex(C.byref(ci.c), C.byref(_c_params(params)), C.byref(opt),
                                       C.byref(outs.c), C.byref(st.c)))
    return BaOutput(outs.poses, outs.points, outs.cams, outs.c.initial_cost, outs.c.final_cost,
                    int(outs.c.iterations_run), st.to_dict())


class Problem:
    """A problem made resident on the device once and solved many times
    (``sfmba_problem_create`` / ``_solve``).  Each solve starts from the
    input's initial state with bundle_adjust's semantics."""

    def __init__(self, inp: BaInput, device: int = 0):
        self._lib = N.load_engine()
        self._ci = _CInput(inp)
        h = C.c_void_p()
        N.check(self._lib.sfmba_problem_create(C.byref(self._ci.c), int(device), C.byref(h)))
        self._h = h
        self.device = device

    def solve(self, params: BaParams | None = None, *, precision: str = "fp64", solver: str = "cholesky",
              profile: bool = False, refine_steps: int | None = None, pcg_max_iterations: int | None = None,
              pcg_tolerance: float | None = None) -> BaOutput:
        params = params or BaParams()
        outs = _Outputs(self._ci)
        st = _Stats()
        opt = _options(precision, solver, self.device, profile, refine_steps, pcg_max_iterations, pcg_tolerance)
        N.check(self._lib.sfmba_problem_solve(self._h, C.byref(_c_params(params)), C.byref(opt), C.byref(outs.c),
                                              C.byref(st.c)))
        return BaOutput(outs.poses, outs.points, outs.cams, outs.c.initial_cost, outs.c.final_cost,
                        int(outs.c.iterations_run), st.to_dict())

    def close(self) -> None:
        if self._h:
            self._lib.sfmba_problem_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def bundle_adjust_batch(inputs: Sequence[BaInput], params: BaParams | None = None, *,
                        precision: str = "fp64", device: int = 0, solver: str = "cholesky") -> list:
    """Many independent problems (local-BA windows) in one batched call."""
    params = params or BaParams()
    lib = N.load_engine()
    cis = [_CInput(i) for i in inputs]
    outs = [_Outputs(ci) for ci in cis]
    stats = [_Stats(256) for _ in cis]
    arr_in = (N.Input * len(cis))(*[ci.c for ci in cis])
    arr_out = (N.Output * len(cis))(*[o.c for o in outs])
    arr_st = (N.Stats * len(cis))(*[s.c for s in stats])
    opt = _options(precision, solver, device, False)
    N.check(lib.sfmba_bundle_adjust_batch(arr_in, len(cis), C.byref(_c_params(params)), C.byref(opt),
                                          arr_out, arr_st))
    res = []
    for k, o in enumerate(outs):
        stats[k].c = arr_st[k]
        res.append(BaOutput(o.poses, o.points, o.cams, arr_out[k].initial_cost, arr_out[k].final_cost,
                            int(arr_out[k].iterations_run), stats[k].to_dict()))
    return res


class BatchPlan:
    """A batch of independent problems prepared once and solved by replaying
    a captured HIP graph (sfmba_batch_create / sfmba_batch_solve).  No
    counterpart in the reference, whose callers loop over bundle_adjust
    (crates/sfm-reconstruction/src/lib.rs:708-719).  autotune_iterations > 0
    times every window-kernel launch shape on this batch (iterations capped
    at that value) and keeps the fastest."""

    def __init__(self, inputs: Sequence[BaInput], params: BaParams | None = None, *, precision: str = "fp64",
                 device: int = 0, autotune_iterations: int = 0):
        self._lib = N.load_engine()
        self._h = C.c_void_p()
        params = params or BaParams()
        self._cis = [_CInput(i) for i in inputs]
        arr_in = (N.Input * len(self._cis))(*[ci.c for ci in self._cis])
        opt = _options(precision, "cholesky", device, False)
        N.check(self._lib.sfmba_batch_create(arr_in, len(self._cis), C.byref(_c_params(params)), C.byref(opt),
                                             int(autotune_iterations), C.byref(self._h)))

    def info(self) -> dict:
        inf = N.BatchInfo()
        N.check(self._lib.sfmba_batch_get_info(self._h, C.byref(inf)))
        n = inf.num_shapes
        return {"threads_per_window": inf.threads_per_window, "pairs_in_flight": inf.pairs_in_flight,
                "tuned": bool(inf.tuned),
                "shapes": [{"threads": inf.shape_threads[k], "pairs": inf.shape_pairs[k],
                            "seconds": inf.shape_seconds[k]} for k in range(n)],
                "create_seconds": inf.create_seconds, "capture_seconds": inf.capture_seconds,
                "num_windows": inf.num_windows, "num_large": inf.num_large}

    def solve(self) -> list:
        if not self._h:
            raise ValueError("plan is closed")
        outs = [_Outputs(ci) for ci in self._cis]
        stats = [_Stats(256) for _ in self._cis]
        arr_out = (N.Output * len(self._cis))(*[o.c for o in outs])
        arr_st = (N.Stats * len(self._cis))(*[s.c for s in stats])
        N.check(self._lib.sfmba_batch_solve(self._h, arr_out, arr_st))
        res = []
        for k, o in enumerate(outs):
            stats[k].c = arr_st[k]
            res.append(BaOutput(o.poses, o.points, o.cams, arr_out[k].initial_cost, arr_out[k].final_cost,
                                int(arr_out[k].iterations_run), stats[k].to_dict()))
        return res

    def close(self) -> None:
        if self._h:
            self._lib.sfmba_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


INTRINSICS = {"fixed": N.INTRINSICS_FIXED, "free": N.INTRINSICS_FREE, "free_guarded": N.INTRINSICS_FREE_GUARDED}


def filter_and_reoptimize(inp: BaInput, params: BaParams | None = None, threshold_px: float = 4.0, *,
                          precision: str = "fp64", device: int = 0) -> tuple:
    """crates/sfm-reconstruction/src/lib.rs:1710-1768.  Returns (BaOutput,
    number of observations kept)."""
    params = params or BaParams()
    lib = N.load_engine()
    ci = _CInput(inp)
    outs = _Outputs(ci)
    kept = C.c_uint64()
    opt = _options(precision, "cholesky", device, False)
    N.check(lib.sfmba_filter_and_reoptimize(C.byref(ci.c), C.byref(_c_params(params)), C.byref(opt),
                                            float(threshold_px), C.byref(outs.c), C.byref(kept)))
    return (BaOutput(outs.poses, outs.points, outs.cams, outs.c.initial_cost, outs.c.final_cost,
                     int(outs.c.iterations_run), {}), int(kept.value))


def run_bundle_adjustment(inp: BaInput, intrinsics: str = "free_guarded", *,
                          robust_loss: RobustLoss = RobustLoss.HUBER, max_reprojection_error_px: float = 4.0,
                          camera_caller_fixed=(), precision: str = "fp64", device: int = 0) -> tuple:
    """The solve of crates/sfm-reconstruction/src/lib.rs:1479-1695 on an
    assembled problem (its fixed_cameras / fixed_camera_params are set by the
    policy).  Returns (BaOutput, report dict)."""
    lib = N.load_engine()
    ci = _CInput(inp)
    outs = _Outputs(ci)
    pol = N.Policy()
    lib.sfmba_policy_default(C.byref(pol))
    pol.intrinsics = INTRINSICS[intrinsics]
    pol.robust_loss = int(robust_loss)
    pol.max_reprojection_error_px = float(max_reprojection_error_px)
    cf = np.ascontiguousarray(np.asarray(camera_caller_fixed, dtype=np.uint8))
    pol.camera_caller_fixed = N.ptr(cf)
    pol.num_camera_caller_fixed = cf.size
    rep = N.PolicyReport()
    opt = _options(precision, "cholesky", device, False)
    N.check(lib.sfmba_run_bundle_adjustment(C.byref(ci.c), C.byref(pol), C.byref(opt), C.byref(outs.c),
                                            C.byref(rep)))
    report = {name: getattr(rep, name) for name, _ in N.PolicyReport._fields_}
    return (BaOutput(outs.poses, outs.points, outs.cams, outs.c.initial_cost, outs.c.final_cost,
                     int(outs.c.iterations_run), {}), report)


def mean_reprojection_error(inp: BaInput) -> float:
    """crates/sfm-ba/src/lib.rs:708-725."""
    lib = N.load_engine()
    ci = _CInput(inp)
    out = C.c_double()
    N.check(lib.sfmba_mean_reprojection_error(C.byref(ci.c), C.byref(out)))
    return out.value


def total_cost(inp: BaInput, params: BaParams | None = None) -> float:
    """crates/sfm-ba/src/lib.rs:727-745 (the cost BaOutput reports)."""
    params = params or BaParams()
    lib = N.load_engine()
    ci = _CInput(inp)
    out = C.c_double()
    N.check(lib.sfmba_total_cost(C.byref(ci.c), C.byref(_c_params(params)), C.byref(out)))
    return out.value


def default_fixed_params_mask(cameras) -> list:
    """crates/sfm-ba/src/intrinsics.rs:49-61."""
    lib = N.load_engine()
    arr = cameras_to_array(cameras)
    mask = np.zeros((len(arr), 8), dtype=np.uint8)
    lens = np.zeros(len(arr), dtype=np.uint64)
    N.check(lib.sfmba_default_fixed_params_mask(N.ptr(arr), len(arr), N.ptr(mask), N.ptr(lens)))
    return [[bool(v) for v in mask[i, : int(lens[i])]] for i in range(len(arr))]


def refine_intrinsics(cameras: np.ndarray, camera_of_image, poses, points, observations,
                      params: IntrinsicsRefineParams | None = None) -> np.ndarray:
    """crates/sfm-ba/src/intrinsics.rs:116-146.  ``cameras`` (a CAMERA_DTYPE
    array) is updated in place, as the Rust function mutates its slice; it is
    also returned."""
    params = params or IntrinsicsRefineParams()
    lib = N.load_engine()
    if not (isinstance(cameras, np.ndarray) and cameras.dtype == N.CAMERA_DTYPE and cameras.flags.c_contiguous):
        raise TypeError("cameras must be a contiguous CAMERA_DTYPE array (updated in place)")
    coi = np.ascontiguousarray(np.asarray(camera_of_image, dtype=np.uint64))
    pz = poses_to_array(poses)
    pts = np.ascontiguousarray(np.asarray(points, dtype=np.float64).reshape(-1, 3))
    obs = observations_to_array(observations)
    ip = N.IntrinsicsParams(int(params.max_iterations), float(params.initial_lambda),
                            int(params.robust_loss), 0, float(params.loss_scale_px))
    N.check(lib.sfmba_refine_intrinsics(N.ptr(cameras), len(cameras), N.ptr(coi), pz.shape[0], N.ptr(pz),
                                        N.ptr(pts), pts.shape[0], N.ptr(obs), obs.shape[0], C.byref(ip)))
    return cameras


def observation_jacobians(inp: BaInput, precision: str = "fp64", device: int = 0) -> dict:
    """Residual and analytic Jacobians of every observation, straight from the
    linearization kernel (crates/sfm-ba/src/lib.rs:327-590)."""
    lib = N.load_engine()
    ci = _CInput(inp)
    n = ci.obs.shape[0]
    r = np.zeros((n, 2))
    jp = np.zeros((n, 2, 6))
    jx = np.zeros((n, 2, 3))
    jc = np.zeros((n, 2, 8))
    valid = np.zeros(n, dtype=np.uint8)
    N.check(lib.sfmba_observation_jacobians(C.byref(ci.c), PRECISIONS[precision], device, N.ptr(r), N.ptr(jp),
                                            N.ptr(jx), N.ptr(jc), N.ptr(valid)))
    return {"r": r, "jp": jp, "jx": jx, "jc": jc, "valid": valid.astype(bool)}


def dense_solve(A: np.ndarray, b: np.ndarray, method: str = "cholesky", device: int = 0):
    """Solve the symmetric system A x = b with the engine's dense kernels
    (lower triangle of A read).  Returns (x, ok)."""
    lib = N.load_engine()
    A = np.asfortranarray(np.asarray(A, dtype=np.float64))
    b = np.ascontiguousarray(np.asarray(b, dtype=np.float64))
    n = A.shape[0]
    x = np.zeros(n)
    st = C.c_int32()
    N.check(lib.sfmba_dense_solve(A.ctypes.data, n, b.ctypes.data, {"cholesky": 0, "lu": 1}[method], device,
                                  x.ctypes.data, C.byref(st)))
    return x, st.value == 0


def device_count() -> int:
    return int(N.load_engine().sfmba_device_count())
