"""The C ABI library loads and exports every symbol include/sfmba.h
declares; struct layouts agree; argument errors are reported without a GPU."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from sfmtory_amd import _native as N
from sfmtory_amd import ba, scenes

HEADER = Path(__file__).resolve().parent.parent / "include" / "sfmba.h"


def declared_functions():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(sfmba_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = N.load_engine()
    names = declared_functions()
    assert len(names) >= 14
    for name in names:
        assert hasattr(lib, name), name
        assert name in N.SIGNATURES, name


def test_struct_sizes():
    assert C.sizeof(N.Camera) == 72 and N.CAMERA_DTYPE.itemsize == 72
    assert C.sizeof(N.Pose) == 56
    assert C.sizeof(N.Observation) == 32 and N.OBSERVATION_DTYPE.itemsize == 32
    assert C.sizeof(N.Params) == 32
    assert C.sizeof(N.Options) == 32
    assert C.sizeof(N.Input) == 17 * 8


def test_header_constants_mirrored():
    text = HEADER.read_text()
    lam = re.search(r"#define SFMBA_MIXED_MIN_LAMBDA (\S+)", text).group(1)
    assert float(lam) == N.MIXED_MIN_LAMBDA
    lim = re.search(r"#define SFMBA_NOALLOC_MAX_OBSERVATIONS \(1ull << (\d+)\)", text).group(1)
    assert 1 << int(lim) == N.NOALLOC_MAX_OBSERVATIONS
    for name in ("CHOLESKY", "PCG", "PCG_MIXED"):
        v = re.search(rf"#define SFMBA_SOLVER_{name} (\d+)", text).group(1)
        assert int(v) == getattr(N, f"SOLVER_{name}")


def test_defaults_match_reference():
    lib = N.load_engine()
    p = N.Params()
    lib.sfmba_params_default(C.byref(p))
    assert (p.max_iterations, p.initial_lambda, p.robust_loss, p.loss_scale_px) == (50, 1e-3, 1, 2.0)
    ip = N.IntrinsicsParams()
    lib.sfmba_intrinsics_params_default(C.byref(ip))
    assert (ip.max_iterations, ip.initial_lambda, ip.robust_loss, ip.loss_scale_px) == (30, 1e-3, 1, 2.0)


def test_default_fixed_params_mask_matches_oracle():
    import oracle
    cams = [ba.CameraModel(m, tuple([1.0] * k)) for m, k in enumerate(ba.MODEL_PARAM_COUNT)]
    got = ba.default_fixed_params_mask(cams)
    assert got == oracle.default_fixed_params_mask(cams)
    assert got[0] == [False, True, True] and got[4] == [False, False, True, True] + [False] * 4


def test_out_of_range_index_raises_index_error():
    sc = scenes.generate("c0", num_points=50)
    obs = sc.inp.observations.copy()
    obs[3]["point_idx"] = 10_000
    bad = ba.BaInput(sc.inp.camera_of_image, sc.inp.cameras, sc.inp.poses, sc.inp.points, obs)
    with pytest.raises(IndexError):
        ba.bundle_adjust(bad)
    with pytest.raises(IndexError):
        ba.mean_reprojection_error(bad)


def test_unknown_model_is_invalid_argument():
    sc = scenes.generate("c0", num_points=50)
    cams = sc.inp.cameras.copy()
    cams[0]["model_id"] = 9
    bad = ba.BaInput(sc.inp.camera_of_image, cams, sc.inp.poses, sc.inp.points, sc.inp.observations)
    with pytest.raises(ValueError):
        ba.bundle_adjust(bad)


def test_no_device_fails_loudly():
    if ba.device_count() > 0:
        pytest.skip("a device is present")
    sc = scenes.generate("c0", num_points=50)
    with pytest.raises(N.EngineError) as e:
        ba.bundle_adjust(sc.inp)
    assert e.value.code == N.SFMBA_ERR_DEVICE
    with pytest.raises(N.EngineError) as e:
        ba.BatchPlan([sc.inp], ba.BaParams(max_iterations=2), autotune_iterations=1)
    assert e.value.code == N.SFMBA_ERR_DEVICE


def test_batch_plan_argument_errors_need_no_device():
    lib = N.load_engine()
    h = C.c_void_p()
    p = N.Params()
    lib.sfmba_params_default(C.byref(p))
    # NULL output handle, negative tuning iterations, NULL params
    assert lib.sfmba_batch_create(None, 0, C.byref(p), None, 0, None) == N.SFMBA_ERR_INVALID_ARGUMENT
    assert lib.sfmba_batch_create(None, 0, C.byref(p), None, -1, C.byref(h)) == N.SFMBA_ERR_INVALID_ARGUMENT
    assert lib.sfmba_batch_create(None, 0, None, None, 0, C.byref(h)) == N.SFMBA_ERR_INVALID_ARGUMENT
    assert not h
    assert lib.sfmba_batch_solve(None, None, None) == N.SFMBA_ERR_INVALID_ARGUMENT
    assert lib.sfmba_batch_get_info(None, None) == N.SFMBA_ERR_INVALID_ARGUMENT
    lib.sfmba_batch_destroy(None)


# ctypes mirror -> C struct of include/sfmba.h
MIRRORS = {
    "Camera": "sfmba_camera", "Pose": "sfmba_pose", "Observation": "sfmba_observation", "Input": "sfmba_input",
    "Params": "sfmba_params", "Output": "sfmba_output", "Options": "sfmba_options", "Stats": "sfmba_stats",
    "Policy": "sfmba_policy", "PolicyReport": "sfmba_policy_report", "IntrinsicsParams": "sfmba_intrinsics_params",
    "ColmapRefineOptions": "sfmba_colmap_refine_options", "ColmapRefineReport": "sfmba_colmap_refine_report",
    "Reconstruction": "sfmba_reconstruction", "WindowSummary": "sfmba_window_summary",
    "BatchInfo": "sfmba_batch_info",
}


def test_struct_layouts_match_the_header(tmp_path):
    """Every ctypes mirror has the size, and every field the offset, that a
    C compiler gives the header's struct (field names as in the header)."""
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    lines = ["#include <stdio.h>", "#include <stddef.h>", f'#include "{HEADER}"', "int main(void) {"]
    for py, c in MIRRORS.items():
        lines.append(f'  printf("{py} size %zu\\n", sizeof({c}));')
        for name, *_ in getattr(N, py)._fields_:
            lines.append(f'  printf("{py} {name} %zu\\n", offsetof({c}, {name}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-std=c11", "-o", str(exe), str(src)], check=True, capture_output=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {}
    for line in out:
        if line:
            py, field, val = line.split()
            got[(py, field)] = int(val)
    for py in MIRRORS:
        st = getattr(N, py)
        assert got[(py, "size")] == C.sizeof(st), py
        for name, *_ in st._fields_:
            assert got[(py, name)] == getattr(st, name).offset, (py, name)
