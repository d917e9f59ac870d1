"""ORACLE - test infrastructure only (tests/, the smoke check and bench.py's
cpu_baseline leg).  ctypes binding of oracle/sfmba_oracle.cpp, the CPU FP64
restatement of the reference bundle adjuster.  Nothing in sfmtory_amd
imports this package.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

from sfmtory_amd import _native as N
from sfmtory_amd.ba import BaInput, BaParams, IntrinsicsRefineParams, _c_params, _CInput, cameras_to_array

HERE = Path(__file__).resolve().parent
LIB = HERE / "build" / "liboracle.so"
_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def load() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists() or LIB.stat().st_mtime < (HERE / "sfmba_oracle.cpp").stat().st_mtime:
            build()
        lib = C.CDLL(str(LIB))
        P = C.c_void_p
        d = C.c_double
        sigs = {
            "oracle_param_count": (C.c_int, [C.c_int]),
            "oracle_project": (C.c_int, [P, P, P]),
            "oracle_transform_point": (C.c_int, [P, P, P]),
            "oracle_rotation_matrix": (C.c_int, [P, P]),
            "oracle_perturb_pose": (C.c_int, [P, P, P]),
            "oracle_robust_weight": (d, [d, C.c_int, d]),
            "oracle_reprojection_residual": (C.c_int, [P, P, P, d, d, P]),
            "oracle_analytic_jacobians": (C.c_int, [P, P, P, d, d, P, P, P, P]),
            "oracle_numeric_jacobians": (C.c_int, [P, P, P, d, d, d, P, P, P, P]),
            "oracle_total_cost": (C.c_int, [P, P, P]),
            "oracle_mean_reprojection_error": (C.c_int, [P, P]),
            "oracle_residual_norms": (C.c_int, [P, P, P]),
            "oracle_bundle_adjust": (C.c_int, [P, P, P, C.c_int, P, C.c_size_t, P, P, C.c_size_t, P]),
            "oracle_default_fixed_params_mask": (C.c_int, [P, C.c_size_t, P, P]),
            "oracle_refine_intrinsics": (C.c_int, [P, C.c_size_t, P, C.c_size_t, P, P, C.c_size_t, P,
                                                   C.c_size_t, P]),
            "oracle_reduced_system": (C.c_long, [P, P, d, P, P, C.c_size_t]),
            "oracle_cholesky_solve": (C.c_int, [P, C.c_size_t, P]),
            "oracle_lu_solve": (C.c_int, [P, C.c_size_t, P]),
            "oracle_inverse3": (C.c_int, [P, P]),
            "oracle_stage_times": (C.c_int, [P, P, C.c_int, P]),
            "oracle_cholesky_seconds": (d, [C.c_size_t]),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def _pose(row) -> N.Pose:
    r = np.asarray(row, dtype=np.float64)
    return N.Pose(*[float(v) for v in r])


def _cam(model_id: int, params) -> N.Camera:
    c = N.Camera()
    c.model_id = int(model_id)
    p = list(params)
    for i, v in enumerate(p[:8]):
        c.params[i] = float(v)
    return c


def project(model_id, params, pc):
    uv = np.zeros(2)
    pcv = np.ascontiguousarray(pc, dtype=np.float64)
    load().oracle_project(C.byref(_cam(model_id, params)), pcv.ctypes.data, uv.ctypes.data)
    return uv


def transform_point(pose_row, X):
    out = np.zeros(3)
    x = np.ascontiguousarray(X, dtype=np.float64)
    load().oracle_transform_point(C.byref(_pose(pose_row)), x.ctypes.data, out.ctypes.data)
    return out


def rotation_matrix(pose_row):
    R = np.zeros(9)
    load().oracle_rotation_matrix(C.byref(_pose(pose_row)), R.ctypes.data)
    return R.reshape(3, 3)


def perturb_pose(pose_row, delta):
    out = N.Pose()
    d = np.ascontiguousarray(delta, dtype=np.float64)
    load().oracle_perturb_pose(C.byref(_pose(pose_row)), d.ctypes.data, C.byref(out))
    return np.array([out.qw, out.qx, out.qy, out.qz, out.tx, out.ty, out.tz])


def robust_weight(n, loss, scale):
    return load().oracle_robust_weight(float(n), int(loss), float(scale))


def residual(pose_row, model_id, params, X, obs):
    r = np.zeros(2)
    x = np.ascontiguousarray(X, dtype=np.float64)
    ok = load().oracle_reprojection_residual(C.byref(_pose(pose_row)), C.byref(_cam(model_id, params)),
                                             x.ctypes.data, float(obs[0]), float(obs[1]), r.ctypes.data)
    return r if ok else None


def analytic_jacobians(pose_row, model_id, params, X, obs):
    r = np.zeros(2); jp = np.zeros(12); jx = np.zeros(6); jc = np.zeros(16)
    x = np.ascontiguousarray(X, dtype=np.float64)
    ok = load().oracle_analytic_jacobians(C.byref(_pose(pose_row)), C.byref(_cam(model_id, params)), x.ctypes.data,
                                          float(obs[0]), float(obs[1]), r.ctypes.data, jp.ctypes.data,
                                          jx.ctypes.data, jc.ctypes.data)
    if not ok:
        return None
    return r, jp.reshape(2, 6), jx.reshape(2, 3), jc.reshape(2, 8)


def numeric_jacobians(pose_row, model_id, params, X, obs, eps=1e-6):
    r = np.zeros(2); jp = np.zeros(12); jx = np.zeros(6); jc = np.zeros(16)
    x = np.ascontiguousarray(X, dtype=np.float64)
    ok = load().oracle_numeric_jacobians(C.byref(_pose(pose_row)), C.byref(_cam(model_id, params)), x.ctypes.data,
                                         float(obs[0]), float(obs[1]), float(eps), r.ctypes.data, jp.ctypes.data,
                                         jx.ctypes.data, jc.ctypes.data)
    if not ok:
        return None
    return r, jp.reshape(2, 6), jx.reshape(2, 3), jc.reshape(2, 8)


def total_cost(inp: BaInput, params: BaParams | None = None) -> float:
    ci = _CInput(inp)
    out = C.c_double()
    cp = _c_params(params or BaParams())
    rc = load().oracle_total_cost(C.addressof(ci.c), C.addressof(cp), C.addressof(out))
    if rc:
        raise ValueError("oracle: invalid input")
    return out.value


def mean_reprojection_error(inp: BaInput) -> float:
    ci = _CInput(inp)
    out = C.c_double()
    if load().oracle_mean_reprojection_error(C.addressof(ci.c), C.addressof(out)):
        raise ValueError("oracle: invalid input")
    return out.value


def residual_norms(inp: BaInput):
    """|reprojection residual| per observation and its in-front flag
    (lib.rs:249-261)."""
    ci = _CInput(inp)
    n = len(inp.observations)
    norm = np.zeros(n)
    front = np.zeros(n, dtype=np.uint8)
    if load().oracle_residual_norms(C.addressof(ci.c), norm.ctypes.data, front.ctypes.data):
        raise ValueError("oracle: invalid input")
    return norm, front.astype(bool)


def rms_reprojection_error(inp: BaInput) -> float:
    """Plain RMS of |r| over the observations in front of their camera."""
    norm, front = residual_norms(inp)
    return float(np.sqrt(np.mean(norm[front] ** 2))) if front.any() else 0.0


def bundle_adjust(inp: BaInput, params: BaParams | None = None, nthreads: int = 1) -> dict:
    """Returns poses (n,7), points (n,3), cameras (CAMERA_DTYPE), costs and
    the per-outer-iteration and per-try cost traces."""
    params = params or BaParams()
    ci = _CInput(inp)
    poses = np.zeros_like(ci.poses)
    points = np.zeros_like(ci.points)
    cams = ci.cams.copy()
    out = N.Output(N.ptr(poses), N.ptr(points), N.ptr(cams), 0.0, 0.0, 0)
    cap = 4096
    outer = np.zeros(cap); trial = np.zeros(cap)
    olen = C.c_size_t(); tlen = C.c_size_t()
    cp = _c_params(params)
    rc = load().oracle_bundle_adjust(C.addressof(ci.c), C.addressof(cp), C.addressof(out), int(nthreads),
                                     outer.ctypes.data, cap, C.addressof(olen), trial.ctypes.data, cap,
                                     C.addressof(tlen))
    if rc:
        raise ValueError("oracle: invalid input")
    return dict(poses=poses, points=points, cameras=cams, initial_cost=out.initial_cost,
                final_cost=out.final_cost, iterations_run=int(out.iterations_run),
                outer_costs=outer[: min(olen.value, cap)].copy(), trial_costs=trial[: min(tlen.value, cap)].copy())


def default_fixed_params_mask(cameras) -> list:
    arr = cameras_to_array(cameras)
    mask = np.zeros((len(arr), 8), dtype=np.uint8)
    lens = np.zeros(len(arr), dtype=np.uint64)
    if load().oracle_default_fixed_params_mask(N.ptr(arr), len(arr), N.ptr(mask), N.ptr(lens)):
        raise ValueError("oracle: unknown model")
    return [[bool(v) for v in mask[i, : int(lens[i])]] for i in range(len(arr))]


def refine_intrinsics(cameras: np.ndarray, camera_of_image, poses, points, observations,
                      params: IntrinsicsRefineParams | None = None) -> np.ndarray:
    from sfmtory_amd.ba import observations_to_array, poses_to_array
    params = params or IntrinsicsRefineParams()
    cams = np.ascontiguousarray(cameras.copy())
    coi = np.ascontiguousarray(np.asarray(camera_of_image, dtype=np.uint64))
    pz = poses_to_array(poses)
    pts = np.ascontiguousarray(np.asarray(points, dtype=np.float64).reshape(-1, 3))
    obs = observations_to_array(observations)
    ip = N.IntrinsicsParams(int(params.max_iterations), float(params.initial_lambda), int(params.robust_loss), 0,
                            float(params.loss_scale_px))
    if load().oracle_refine_intrinsics(N.ptr(cams), len(cams), N.ptr(coi), pz.shape[0], N.ptr(pz), N.ptr(pts),
                                       pts.shape[0], N.ptr(obs), obs.shape[0], C.addressof(ip)):
        raise ValueError("oracle: invalid input")
    return cams


def reduced_system(inp: BaInput, params: BaParams, lam: float):
    """(S, rhs) of the first linearization at damping lam, reference layout."""
    ci = _CInput(inp)
    cp = _c_params(params)
    D = load().oracle_reduced_system(C.addressof(ci.c), C.addressof(cp), float(lam), None, None, 0)
    if D < 0:
        raise ValueError("oracle: invalid input")
    S = np.zeros(D * D)
    rhs = np.zeros(D)
    load().oracle_reduced_system(C.addressof(ci.c), C.addressof(cp), float(lam), S.ctypes.data, rhs.ctypes.data, D * D)
    return S.reshape(D, D).T.copy(), rhs  # column-major -> numpy


def cholesky_solve(A, b):
    A = np.asfortranarray(np.array(A, dtype=np.float64))
    x = np.array(b, dtype=np.float64)
    ok = load().oracle_cholesky_solve(A.ctypes.data, A.shape[0], x.ctypes.data)
    return x if ok else None


def lu_solve(A, b):
    A = np.asfortranarray(np.array(A, dtype=np.float64))
    x = np.array(b, dtype=np.float64)
    ok = load().oracle_lu_solve(A.ctypes.data, A.shape[0], x.ctypes.data)
    return x if ok else None


def inverse3(m):
    a = np.ascontiguousarray(m, dtype=np.float64).reshape(9)
    o = np.zeros(9)
    ok = load().oracle_inverse3(a.ctypes.data, o.ctypes.data)
    return o.reshape(3, 3) if ok else None


def stage_times(inp: BaInput, params: BaParams | None = None, nthreads: int = 1) -> dict:
    """Seconds of the reference's stages for one outer iteration: initial
    cost, Jacobians (parallel), accumulation (serial), reduced-system
    assembly (Schur, parallel above 1500 points); plus D of the reference
    layout."""
    ci = _CInput(inp)
    cp = _c_params(params or BaParams())
    t = np.zeros(5)
    if load().oracle_stage_times(C.addressof(ci.c), C.addressof(cp), int(nthreads), t.ctypes.data):
        raise ValueError("oracle: invalid input")
    return dict(cost=t[0], jacobians=t[1], accumulate=t[2], schur=t[3], dim=int(t[4]))


def cholesky_seconds(n: int) -> float:
    return float(load().oracle_cholesky_seconds(int(n)))
