"""Seeded synthetic bundle-adjustment scenes (generator in csrc/scenegen.cpp).

Presets reproduce the benchmark workloads:

  ``c0``  16 SIMPLE_RADIAL (1024x768, f=800, k=-0.05), 2,000 points, track 2..8
  ``c1``  200 OPENCV (1920x1080, f~1500+-5%, k1,k2,p1,p2 ~ N(0, 0.01)), 100k points,
          track 2 + Geom(1/3)
  ``c2``  1,000 PINHOLE, 500k points, track 2 + Geom(1/5), 1% outliers U(+-50 px)
  ``c3w`` one local-BA window: 10..40 SIMPLE_RADIAL, 2k..20k points, track mean 4
  ``c4``  5,000 RADIAL, 2M points, track mean ~6 with a Pareto tail capped at 200

Every scene uses one camera per image, fixes pose 0 as the gauge, leaves all
cameras free with the principal point pinned (the default fixed-parameter
mask), and perturbs the start (rotation, translation, points, focal).
"""
from __future__ import annotations

import ctypes as C
import hashlib
from dataclasses import dataclass

import numpy as np

from . import _native as N
from .ba import BaInput


class SceneSpec(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64), ("model_id", C.c_int32), ("num_images", C.c_int32), ("num_points", C.c_int64),
        ("width", C.c_double), ("height", C.c_double), ("focal", C.c_double), ("focal_jitter", C.c_double),
        ("k_fixed", C.c_double), ("dist_sigma", C.c_double), ("radius", C.c_double), ("cube_half", C.c_double),
        ("track_kind", C.c_int32), ("track_min", C.c_int32), ("track_max", C.c_int32), ("shared_camera", C.c_int32),
        ("track_p", C.c_double), ("track_alpha", C.c_double), ("pixel_sigma", C.c_double),
        ("outlier_frac", C.c_double), ("outlier_px", C.c_double), ("sigma_rot", C.c_double),
        ("sigma_t", C.c_double), ("sigma_x", C.c_double), ("sigma_f", C.c_double),
    ]


class SceneData(C.Structure):
    _fields_ = [
        ("num_images", C.c_uint64), ("num_cameras", C.c_uint64), ("num_points", C.c_uint64),
        ("num_observations", C.c_uint64), ("camera_of_image", C.c_void_p), ("cameras_true", C.c_void_p),
        ("cameras_init", C.c_void_p), ("poses_true", C.c_void_p), ("poses_init", C.c_void_p),
        ("points_true", C.c_void_p), ("points_init", C.c_void_p), ("observations", C.c_void_p),
    ]


_lib = None


def _load():
    global _lib
    if _lib is None:
        if not N.SCENE_PATH.exists():
            raise RuntimeError(f"{N.SCENE_PATH} is missing; run `python -m sfmtory_amd.build`")
        _lib = C.CDLL(str(N.SCENE_PATH))
        _lib.scene_generate.restype = C.c_int
        _lib.scene_generate.argtypes = [C.POINTER(SceneSpec), C.POINTER(SceneData)]
        _lib.scene_free.restype = None
        _lib.scene_free.argtypes = [C.POINTER(SceneData)]
    return _lib


BASE_SEED = 20261203


def spec(**kw) -> SceneSpec:
    d = dict(seed=BASE_SEED, model_id=2, num_images=16, num_points=2000, width=1024, height=768, focal=800.0,
             focal_jitter=0.0, k_fixed=-0.05, dist_sigma=0.0, radius=10.0, cube_half=2.0, track_kind=0,
             track_min=2, track_max=8, shared_camera=0, track_p=0.5, track_alpha=1.5, pixel_sigma=0.5,
             outlier_frac=0.0, outlier_px=0.0, sigma_rot=0.01, sigma_t=0.01, sigma_x=0.02, sigma_f=0.03)
    d.update(kw)
    return SceneSpec(**d)


PRESETS = {
    "c0": dict(seed=BASE_SEED + 0),
    "c1": dict(seed=BASE_SEED + 1, model_id=4, num_images=200, num_points=100_000, width=1920, height=1080,
               focal=1500.0, focal_jitter=0.05, dist_sigma=0.01, track_kind=1, track_min=2, track_max=200,
               track_p=1.0 / 3.0),
    "c2": dict(seed=BASE_SEED + 2, model_id=1, num_images=1000, num_points=500_000, width=1920, height=1080,
               focal=1500.0, focal_jitter=0.05, track_kind=1, track_min=2, track_max=1000, track_p=0.2,
               outlier_frac=0.01, outlier_px=50.0),
    "c3w": dict(seed=BASE_SEED + 3, model_id=2, width=1024, height=768, focal=800.0, track_kind=1,
                track_min=2, track_max=40, track_p=1.0 / 3.0, pixel_sigma=1.0),
    "c4": dict(seed=BASE_SEED + 4, model_id=3, num_images=5000, num_points=2_000_000, width=1920, height=1080,
               focal=1500.0, focal_jitter=0.05, dist_sigma=0.01, k_fixed=0.0, track_kind=2, track_min=4,
               track_max=200, track_alpha=3.0),
}


@dataclass
class Scene:
    inp: BaInput                 # perturbed starting estimate, pose 0 fixed
    poses_true: np.ndarray
    points_true: np.ndarray
    cameras_true: np.ndarray

    def sha256(self) -> str:
        h = hashlib.sha256()
        for a in (np.asarray(self.inp.camera_of_image), self.inp.cameras, self.inp.poses, self.inp.points,
                  self.inp.observations):
            h.update(np.ascontiguousarray(a).tobytes())
        return h.hexdigest()


def generate(preset: str | None = None, **overrides) -> Scene:
    kw = dict(PRESETS[preset]) if preset else {}
    kw.update(overrides)
    sp = spec(**kw)
    lib = _load()
    sd = SceneData()
    if lib.scene_generate(C.byref(sp), C.byref(sd)) != 0:
        raise ValueError("invalid scene spec")
    try:
        def arr(p, dtype, n, shape=None):
            if n == 0:
                return np.zeros((0,) if shape is None else (0,) + shape, dtype=dtype)
            nbytes = n * np.dtype(dtype).itemsize * (int(np.prod(shape)) if shape else 1)
            buf = (C.c_char * nbytes).from_address(p)
            a = np.frombuffer(buf, dtype=dtype).copy()
            return a.reshape((n,) + shape) if shape else a

        ni, nc, npt, no = sd.num_images, sd.num_cameras, sd.num_points, sd.num_observations
        coi = arr(sd.camera_of_image, np.uint64, ni)
        cams_true = arr(sd.cameras_true, N.CAMERA_DTYPE, nc)
        cams_init = arr(sd.cameras_init, N.CAMERA_DTYPE, nc)
        poses_true = arr(sd.poses_true, np.float64, ni, (7,))
        poses_init = arr(sd.poses_init, np.float64, ni, (7,))
        pts_true = arr(sd.points_true, np.float64, npt, (3,))
        pts_init = arr(sd.points_init, np.float64, npt, (3,))
        obs = arr(sd.observations, N.OBSERVATION_DTYPE, no)
    finally:
        lib.scene_free(C.byref(sd))
    mask = default_mask_for_models(cams_init["model_id"])
    fixed_poses = np.zeros(int(sp.num_images), dtype=bool)
    fixed_poses[0] = True
    inp = BaInput(camera_of_image=coi, cameras=cams_init, poses=poses_init, points=pts_init, observations=obs,
                  fixed_poses=fixed_poses, fixed_cameras=np.zeros(len(cams_init), dtype=bool),
                  fixed_camera_params=mask)
    return Scene(inp, poses_true, pts_true, cams_true)


def default_mask_for_models(model_ids) -> list:
    """The principal-point mask of default_fixed_params_mask, built on the host
    for scene construction (crates/sfm-ba/src/intrinsics.rs:33-61)."""
    out = []
    for m in model_ids:
        m = int(m)
        k = [3, 4, 4, 5, 8, 8][m]
        mask = [False] * k
        idx = (1, 2) if m in (0, 2, 3) else (2, 3)
        for i in idx:
            mask[i] = True
        out.append(mask)
    return out


def window_scenes(count: int = 256, seed: int = BASE_SEED + 3) -> list:
    """Independent local-BA windows: 10..40 SIMPLE_RADIAL images, 2k..20k
    points, track mean 4, 1 px noise."""
    rng = np.random.default_rng(seed)
    out = []
    for w in range(count):
        n_img = int(rng.integers(10, 41))
        n_pt = int(rng.integers(2000, 20001))
        out.append(generate("c3w", seed=seed * 1000 + w, num_images=n_img, num_points=n_pt))
    return out
