"""Pins the oracle's analytic Jacobians (crates/sfm-ba/src/lib.rs:327-590):
against the reference's own central-difference procedure and tolerances
(lib.rs:1252-1363) and against sympy-derived goldens
(tests/golden/make_jacobian_goldens.py)."""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle
from conftest import pose_row_from_euler

GOLD = json.loads((Path(__file__).parent / "golden" / "jacobians.json").read_text())
CAMS = [(c["name"], c["model_id"], c["params"]) for c in GOLD["cameras"]]
POSE = pose_row_from_euler(0.15, -0.3, 0.08, (0.6, -0.2, 1.3))
POINT = np.array([0.35, -0.18, 2.7])
OBS = (410.0, 260.0)


def test_golden_pose_matches_euler_convention():
    assert np.allclose(GOLD["pose_wxyz"], POSE[:4], rtol=0, atol=1e-15)


@pytest.mark.parametrize("name,model,params", CAMS)
def test_analytic_matches_numerical_reference_tolerances(name, model, params):
    # assert_analytic_matches_numerical, lib.rs:1252-1295
    rn, jpn, jxn, jcn = oracle.numeric_jacobians(POSE, model, params, POINT, OBS, 1e-6)
    ra, jpa, jxa, jca = oracle.analytic_jacobians(POSE, model, params, POINT, OBS)
    assert np.linalg.norm(rn - ra) < 1e-9
    assert np.linalg.norm(jpn - jpa) < 1e-5
    assert np.linalg.norm(jxn - jxa) < 1e-5
    k = len(params)
    assert np.linalg.norm(jcn[:, :k] - jca[:, :k]) < 1e-4


@pytest.mark.parametrize("idx", range(len(GOLD["cameras"])))
def test_analytic_matches_sympy_golden(idx):
    g = GOLD["cameras"][idx]
    r, jp, jx, jc = oracle.analytic_jacobians(POSE, g["model_id"], g["params"], POINT, OBS)
    k = len(g["params"])
    for got, want in ((r, g["r"]), (jp, g["jp"]), (jx, g["jx"]), (jc[:, :k], g["jc"])):
        want = np.asarray(want)
        scale = max(1.0, np.abs(want).max())
        assert np.abs(got - want).max() <= 1e-11 * scale, (g["name"], got, want)
    # unused camera columns stay zero
    assert np.all(jc[:, k:] == 0.0)


def test_behind_camera_is_none():
    # lib.rs:256-258, 339-341: z <= 1e-9 -> no residual
    pose = np.array([1.0, 0, 0, 0, 0, 0, 0])
    assert oracle.residual(pose, 0, [500, 320, 240], [0.1, 0.2, -1.0], (0, 0)) is None
    assert oracle.residual(pose, 0, [500, 320, 240], [0.1, 0.2, 1e-9], (0, 0)) is None
    assert oracle.analytic_jacobians(pose, 0, [500, 320, 240], [0.1, 0.2, 0.0], (0, 0)) is None
    assert oracle.residual(pose, 0, [500, 320, 240], [0.1, 0.2, 2e-9], (0, 0)) is not None


def test_residual_closed_form_pinhole():
    pose = np.array([1.0, 0, 0, 0, 0.1, -0.2, 0.5])
    r = oracle.residual(pose, 1, [1000, 900, 320, 240], [0.3, 0.4, 2.5], (300.0, 200.0))
    x, y, z = 0.4, 0.2, 3.0
    assert np.allclose(r, [1000 * x / z + 320 - 300, 900 * y / z + 240 - 200], rtol=1e-15)
