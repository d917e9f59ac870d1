"""Bundle adjustment of a growing reconstruction in global or local scope.

Python face of the sfmba_window_* / sfmba_reconstruction_bundle_adjust*
entry points, which replace the problem assembly of the reference's
run_bundle_adjustment (crates/sfm-reconstruction/src/lib.rs:1339-1474,
BaScope :1287-1299, write-back :1697-1707).  A Reconstruction holds the
pipeline's working state as flat arrays; the in/out arrays (poses, points,
cameras) are updated in place.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _native as N
from .ba import INTRINSICS, BaInput, BaParams, RobustLoss, _c_params, _options

SCOPES = {"global": N.SCOPE_GLOBAL, "local": N.SCOPE_LOCAL}


@dataclass
class Reconstruction:
    """Working state of incremental reconstruction: per image its
    registered flag, world-to-camera pose row (qw qx qy qz tx ty tz),
    physical camera id and keypoints (single precision, image-major with
    offsets); per point its position and track of (image, keypoint index);
    the camera table by id; the gauge (seed) image; the caller's pinned
    poses and camera ids."""

    registered: np.ndarray
    poses: np.ndarray
    image_camera_id: np.ndarray
    keypoint_offsets: np.ndarray
    keypoints: np.ndarray
    points: np.ndarray
    track_offsets: np.ndarray
    track_images: np.ndarray
    track_keypoints: np.ndarray
    camera_ids: np.ndarray
    cameras: np.ndarray
    seed_image: int = 0
    pose_pinned: np.ndarray | None = None
    caller_fixed_camera_ids: np.ndarray = field(default_factory=lambda: np.zeros(0, dtype=np.uint32))

    def __post_init__(self):
        def arr(a, dt, shape=None):
            a = np.ascontiguousarray(a, dtype=dt)
            return a.reshape(shape) if shape is not None else a

        self.registered = arr(self.registered, np.uint8)
        self.poses = arr(self.poses, np.float64, (-1, 7))
        self.image_camera_id = arr(self.image_camera_id, np.uint32)
        self.keypoint_offsets = arr(self.keypoint_offsets, np.uint64)
        self.keypoints = arr(self.keypoints, np.float32, (-1, 2))
        self.points = arr(self.points, np.float64, (-1, 3))
        self.track_offsets = arr(self.track_offsets, np.uint64)
        self.track_images = arr(self.track_images, np.uint32)
        self.track_keypoints = arr(self.track_keypoints, np.uint32)
        self.camera_ids = arr(self.camera_ids, np.uint32)
        self.cameras = np.ascontiguousarray(self.cameras, dtype=N.CAMERA_DTYPE)
        if self.pose_pinned is not None:
            self.pose_pinned = arr(self.pose_pinned, np.uint8)
        self.caller_fixed_camera_ids = arr(self.caller_fixed_camera_ids, np.uint32)
        n = len(self.registered)
        if not (len(self.poses) == len(self.image_camera_id) == n and len(self.keypoint_offsets) == n + 1):
            raise ValueError("image arrays disagree in length")
        if len(self.track_offsets) != len(self.points) + 1:
            raise ValueError("track_offsets must have num_points + 1 entries")
        if len(self.camera_ids) != len(self.cameras):
            raise ValueError("camera_ids and cameras disagree in length")
        if self.pose_pinned is not None and len(self.pose_pinned) != n:
            raise ValueError("pose_pinned must have one entry per image")
        if int(self.keypoint_offsets[-1]) != len(self.keypoints):
            raise ValueError("keypoint_offsets[-1] must equal the number of keypoints")
        if int(self.track_offsets[-1]) != len(self.track_images) or len(self.track_images) != len(self.track_keypoints):
            raise ValueError("track arrays disagree in length")

    def c(self) -> N.Reconstruction:
        r = N.Reconstruction()
        r.num_images = len(self.registered)
        r.registered = N.ptr(self.registered)
        r.poses = N.ptr(self.poses)
        r.image_camera_id = N.ptr(self.image_camera_id)
        r.pose_pinned = N.ptr(self.pose_pinned)
        r.keypoint_offsets = self.keypoint_offsets.ctypes.data
        r.keypoints = N.ptr(self.keypoints)
        r.num_points = len(self.points)
        r.points = N.ptr(self.points)
        r.track_offsets = self.track_offsets.ctypes.data
        r.track_images = N.ptr(self.track_images)
        r.track_keypoints = N.ptr(self.track_keypoints)
        r.num_cameras = len(self.cameras)
        r.camera_ids = N.ptr(self.camera_ids)
        r.cameras = N.ptr(self.cameras)
        r.caller_fixed_camera_ids = N.ptr(self.caller_fixed_camera_ids)
        r.num_caller_fixed_camera_ids = len(self.caller_fixed_camera_ids)
        r.seed_image = int(self.seed_image)
        return r


def _summary(s: N.WindowSummary) -> dict:
    return {name: getattr(s, name) for name, _ in N.WindowSummary._fields_}


class Window:
    """One assembled problem (sfmba_window_create).  ``input`` is a BaInput
    copy of the compact problem with intrinsics held (the in-loop local
    bundle); ``scatter`` writes a solve of it back into the reconstruction."""

    def __init__(self, rec: Reconstruction, scope: str = "local", center: int = 0):
        self._lib = N.load_engine()
        self._rec = rec
        c = rec.c()
        h = C.c_void_p()
        N.check(self._lib.sfmba_window_create(C.byref(c), SCOPES[scope], int(center), C.byref(h)))
        self._h = h
        inp = N.Input()
        N.check(self._lib.sfmba_window_input(self._h, C.byref(inp)))
        ip, npp, pp, npt = C.c_void_p(), C.c_size_t(), C.c_void_p(), C.c_size_t()
        N.check(self._lib.sfmba_window_indices(self._h, C.byref(ip), C.byref(npp), C.byref(pp), C.byref(npt)))

        def take(ptr, n, dtype, shape=None):
            if n == 0:
                return np.zeros((0,) + (shape or ()), dtype=dtype)
            nbytes = n * np.dtype(dtype).itemsize * (int(np.prod(shape)) if shape else 1)
            a = np.frombuffer((C.c_char * nbytes).from_address(ptr), dtype=dtype).copy()
            return a.reshape((n,) + shape) if shape else a

        self.image_indices = take(ip.value, npp.value, np.uint64)
        self.point_indices = take(pp.value, npt.value, np.uint64)
        ni, nc, npts, no = inp.num_images, inp.num_cameras, inp.num_points, inp.num_observations
        lens = take(inp.fixed_camera_params_len, inp.num_fixed_camera_params, np.uintp)
        flat = take(inp.fixed_camera_params, int(lens.sum()), np.uint8)
        masks, k = [], 0
        for n in lens:
            masks.append([bool(v) for v in flat[k:k + int(n)]])
            k += int(n)
        self.input = BaInput(take(inp.camera_of_image, ni, np.uint64), take(inp.cameras, nc, N.CAMERA_DTYPE),
                             take(inp.poses, ni, np.float64, (7,)), take(inp.points, npts, np.float64, (3,)),
                             take(inp.observations, no, N.OBSERVATION_DTYPE),
                             fixed_poses=take(inp.fixed_poses, ni, np.uint8).astype(bool),
                             fixed_cameras=take(inp.fixed_cameras, nc, np.uint8).astype(bool),
                             fixed_camera_params=masks)

    @property
    def empty(self) -> bool:
        return len(self.image_indices) == 0

    def scatter(self, out) -> None:
        poses = np.ascontiguousarray(out.poses, dtype=np.float64)
        pts = np.ascontiguousarray(out.points, dtype=np.float64)
        cams = np.ascontiguousarray(out.cameras, dtype=N.CAMERA_DTYPE)
        o = N.Output(N.ptr(poses), N.ptr(pts), N.ptr(cams), 0.0, 0.0, 0)
        c = self._rec.c()  # the reconstruction's arrays as they are now
        N.check(self._lib.sfmba_window_scatter(self._h, C.byref(o), C.byref(c)))

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.sfmba_window_destroy(self._h)
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


def bundle_adjust_scope(rec: Reconstruction, scope: str = "local", center: int = 0, *, intrinsics: str = "fixed",
                        robust_loss: RobustLoss = RobustLoss.HUBER, max_reprojection_error_px: float = 4.0,
                        precision: str = "fp64", device: int = 0) -> tuple:
    """run_bundle_adjustment (crates/sfm-reconstruction/src/lib.rs:1323-1708)
    on ``rec`` in place.  Returns (window summary, policy report)."""
    lib = N.load_engine()
    c = rec.c()
    pol = N.Policy()
    lib.sfmba_policy_default(C.byref(pol))
    pol.intrinsics = INTRINSICS[intrinsics]
    pol.robust_loss = int(robust_loss)
    pol.max_reprojection_error_px = float(max_reprojection_error_px)
    opt = _options(precision, "cholesky", device, False)
    s, rep = N.WindowSummary(), N.PolicyReport()
    N.check(lib.sfmba_reconstruction_bundle_adjust(C.byref(c), SCOPES[scope], int(center), C.byref(pol),
                                                   C.byref(opt), C.byref(s), C.byref(rep)))
    return _summary(s), {name: getattr(rep, name) for name, _ in N.PolicyReport._fields_}


def bundle_adjust_scope_batch(recs, scopes, centers, params: BaParams | None = None, *, precision: str = "fp64",
                              device: int = 0) -> list:
    """One window per independent reconstruction, intrinsics held, solved in
    one batched launch and written back in place.  Returns the summaries."""
    lib = N.load_engine()
    recs = list(recs)
    arr = (N.Reconstruction * max(len(recs), 1))(*[r.c() for r in recs])
    sc = np.ascontiguousarray([SCOPES[s] for s in scopes], dtype=np.int32)
    ce = np.ascontiguousarray(centers, dtype=np.uint64)
    if len(sc) != len(recs) or len(ce) != len(recs):
        raise ValueError("one scope and one center per reconstruction")
    sums = (N.WindowSummary * max(len(recs), 1))()
    opt = _options(precision, "cholesky", device, False)
    N.check(lib.sfmba_reconstruction_bundle_adjust_batch(arr, len(recs), N.ptr(sc), N.ptr(ce),
                                                         C.byref(_c_params(params or BaParams())), C.byref(opt),
                                                         C.cast(sums, C.c_void_p)))
    out = [_summary(sums[k]) for k in range(len(recs))]
    # write-back went through the C structs' pointers into the arrays
    return out
