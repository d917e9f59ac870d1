"""COLMAP plain-text sparse models (cameras.txt / images.txt / points3D.txt)
and the standalone refine pass over one.

Python face of sfmba_colmap_refine / sfmba_colmap_rewrite, which replace
the reference's refine command (crates/sfm-cli/src/main.rs:946-1130) and
its model reader/writer (crates/sfm-io/src/colmap.rs:25-300).  Parsing,
writing and the compaction run in the native library; the solve and the
per-point residuals run on the GPU.
"""
from __future__ import annotations

import ctypes as C
import os

from . import _native as N
from .ba import RobustLoss, _options


def _path(p) -> bytes:
    return os.fsencode(os.fspath(p))


def refine_model(model_dir, output_dir=None, *, refine_intrinsics: bool = False,
                 robust_loss: RobustLoss = RobustLoss.HUBER, max_iterations: int = 200,
                 precision: str = "fp64", solver: str = "cholesky", device: int = 0) -> dict:
    """Refine the model in ``model_dir`` and write it to ``output_dir``
    (default: in place, as the reference does).  Returns the report the
    reference logs: image / point counts, initial and final cost, LM
    iterations, and the track-weighted mean reprojection error before and
    after."""
    lib = N.load_engine()
    o = N.ColmapRefineOptions()
    lib.sfmba_colmap_refine_options_default(C.byref(o))
    o.refine_intrinsics = 1 if refine_intrinsics else 0
    o.robust_loss = int(robust_loss)
    o.max_iterations = int(max_iterations)
    opt = _options(precision, solver, device, False)
    rep = N.ColmapRefineReport()
    out = model_dir if output_dir is None else output_dir
    N.check(lib.sfmba_colmap_refine(_path(model_dir), _path(out), C.byref(o), C.byref(opt), C.byref(rep)))
    return {name: getattr(rep, name) for name, _ in N.ColmapRefineReport._fields_}


def rewrite_model(model_dir, output_dir) -> tuple[int, int, int]:
    """Read the model in ``model_dir`` and write it to ``output_dir`` through
    the same reader and writer (host only).  Returns (cameras, images,
    points3d) counts."""
    lib = N.load_engine()
    counts = (C.c_uint64 * 3)()
    N.check(lib.sfmba_colmap_rewrite(_path(model_dir), _path(output_dir), counts))
    return int(counts[0]), int(counts[1]), int(counts[2])
