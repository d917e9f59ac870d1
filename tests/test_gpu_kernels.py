"""Linearization kernel and dense solver on the GPU, against the oracle,
central differences, the sympy goldens and numpy/LAPACK."""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle
from conftest import pose_row_from_euler
from sfmtory_amd import ba, scenes

pytestmark = pytest.mark.gpu
GOLD = json.loads((Path(__file__).parent / "golden" / "jacobians.json").read_text())
POSE = pose_row_from_euler(0.15, -0.3, 0.08, (0.6, -0.2, 1.3))


def single_obs_input(model, params):
    cams = ba.cameras_to_array([ba.CameraModel(model, tuple(params))])
    obs = np.zeros(1, dtype=[("image_idx", "<u8"), ("point_idx", "<u8"), ("x", "<f8"), ("y", "<f8")])
    obs[0] = (0, 0, 410.0, 260.0)
    return ba.BaInput([0], cams, POSE[None, :], np.array([[0.35, -0.18, 2.7]]), obs)


@pytest.mark.parametrize("idx", range(6))
def test_kernel_jacobians_match_sympy_golden(idx):
    g = GOLD["cameras"][idx]
    J = ba.observation_jacobians(single_obs_input(g["model_id"], g["params"]), "fp64")
    k = len(g["params"])
    assert J["valid"][0]
    for got, want in ((J["r"][0], g["r"]), (J["jp"][0], g["jp"]), (J["jx"][0], g["jx"]), (J["jc"][0][:, :k], g["jc"])):
        want = np.asarray(want)
        assert np.abs(got - want).max() <= 1e-10 * max(1.0, np.abs(want).max())
    J32 = ba.observation_jacobians(single_obs_input(g["model_id"], g["params"]), "mixed")
    assert np.abs(J32["r"][0] - g["r"]).max() <= 1e-9 * max(1.0, np.abs(g["r"]).max())  # residual stays FP64
    for got, want in ((J32["jp"][0], g["jp"]), (J32["jx"][0], g["jx"]), (J32["jc"][0][:, :k], g["jc"])):
        want = np.asarray(want)
        assert np.abs(got - want).max() <= 2e-5 * max(1.0, np.abs(want).max())


@pytest.mark.parametrize("model", range(6))
def test_kernel_jacobians_match_oracle_and_fd_on_scene(model):
    sc = scenes.generate("c0", model_id=model, num_images=8, num_points=150, seed=41 + model,
                         dist_sigma=0.01 if model >= 3 else 0.0)
    inp = sc.inp
    J = ba.observation_jacobians(inp, "fp64")
    for i, o in enumerate(inp.observations[:60]):
        img, pt = int(o["image_idx"]), int(o["point_idx"])
        cam = inp.cameras[int(inp.camera_of_image[img])]
        r, jp, jx, jc = oracle.analytic_jacobians(inp.poses[img], cam["model_id"], cam["params"], inp.points[pt],
                                                  (o["x"], o["y"]))
        for got, want in ((J["r"][i], r), (J["jp"][i], jp), (J["jx"][i], jx), (J["jc"][i], jc)):
            assert np.abs(got - want).max() <= 1e-12 * max(1.0, np.abs(want).max())
        rn, jpn, jxn, jcn = oracle.numeric_jacobians(inp.poses[img], cam["model_id"], cam["params"],
                                                     inp.points[pt], (o["x"], o["y"]), 1e-6)
        # the reference's tolerances (lib.rs:1266-1294)
        assert np.linalg.norm(J["jp"][i] - jpn) < 1e-5 * max(1.0, np.abs(jpn).max() / 100)
        assert np.linalg.norm(J["jx"][i] - jxn) < 1e-5 * max(1.0, np.abs(jxn).max() / 100)


def test_kernel_behind_camera_flag():
    inp = single_obs_input(0, [500.0, 320.0, 240.0])
    inp.points = np.array([[0.1, 0.1, -3.0]])
    J = ba.observation_jacobians(inp, "fp64")
    assert not J["valid"][0]


@pytest.mark.parametrize("n", [1, 7, 128, 129, 200, 256, 300, 517, 1100])
def test_dense_cholesky_and_lu_match_lapack(n):
    rng = np.random.default_rng(n)
    M = rng.standard_normal((n, n))
    A = M @ M.T + n * np.eye(n)
    b = rng.standard_normal(n)
    want = np.linalg.solve(A, b)
    x, ok = ba.dense_solve(A, b, "cholesky")
    assert ok and np.abs(x - want).max() <= 1e-11 * np.abs(want).max()
    x, ok = ba.dense_solve(A, b, "lu")
    assert ok and np.abs(x - want).max() <= 1e-11 * np.abs(want).max()


@pytest.mark.parametrize("n", [3100, 8500])
def test_dense_cholesky_large_matches_lapack(n):
    # 3100: more trailing tiles than CUs in the first steps, so each
    # persistent workgroup walks several tiles; 8500: 67 blocks, so the first
    # steps go in pairs (the two-panel K = 256 update) before single steps.
    rng = np.random.default_rng(n)
    R = rng.uniform(-1.0, 1.0, (n, n))
    A = 0.5 * (R + R.T) + 2.0 * n * np.eye(n)  # diagonally dominant: SPD
    b = rng.standard_normal(n)
    want = np.linalg.solve(A, b)
    x, ok = ba.dense_solve(A, b, "cholesky")
    assert ok
    assert np.abs(x - want).max() <= 1e-12 * np.abs(want).max()
    assert np.abs(A @ x - b).max() <= 1e-12 * np.abs(b).max() * n


def test_dense_cholesky_reports_indefinite_and_lu_solves_it():
    A = np.diag([1.0, -2.0, 3.0] * 50)
    b = np.ones(150)
    _, ok = ba.dense_solve(A, b, "cholesky")
    assert not ok
    x, ok = ba.dense_solve(A, b, "lu")
    assert ok and np.allclose(x, np.linalg.solve(A, b), rtol=1e-14)
    _, ok = ba.dense_solve(np.zeros((5, 5)), np.ones(5), "lu")
    assert not ok


def test_cost_and_mean_error_match_oracle(c0_scene):
    inp = c0_scene.inp
    for loss in (ba.RobustLoss.NONE, ba.RobustLoss.HUBER, ba.RobustLoss.CAUCHY):
        p = ba.BaParams(robust_loss=loss, loss_scale_px=1.5)
        assert ba.total_cost(inp, p) == pytest.approx(oracle.total_cost(inp, p), rel=1e-13)
    assert ba.mean_reprojection_error(inp) == pytest.approx(oracle.mean_reprojection_error(inp), rel=1e-13)
