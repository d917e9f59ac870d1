"""bundle_adjust on the GPU (FP64) against the oracle restatement:
per-iteration cost within 1e-9 relative, same iteration count, same final
state; every camera model, shared and fixed cameras, masks, losses, edge
cases, the reference's own scenarios, and the LU fallback."""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle
import ref_scenarios as R
from sfmtory_amd import ba, scenes

pytestmark = pytest.mark.gpu
GOLD = json.loads((Path(__file__).parent / "golden" / "c0_trajectory.json").read_text())


def cost_tol(cost, n_obs, rtol):
    """1e-9 relative, plus the floor set by FP64 rounding of pixel
    coordinates (a few ulp, ~1e-13 px at 2000 px): with cost = sum r^2 a
    rounding error e in every residual moves the cost by ~2 sqrt(N cost) e.
    The floor only matters on noiseless scenes whose cost falls to ~1e-20."""
    return rtol * cost + 2.0 * np.sqrt(n_obs * np.maximum(cost, 0.0)) * 1e-13


def assert_parity(inp, params, rtol=1e-9, state_atol=1e-6):
    out = ba.bundle_adjust(inp, params, precision="fp64")
    ref = oracle.bundle_adjust(inp, params)
    gtr, rtr = np.asarray(out.stats["outer_costs"]), ref["outer_costs"]
    n = max(1, len(inp.observations))
    assert out.iterations_run == ref["iterations_run"], (gtr, rtr)
    assert gtr.shape == rtr.shape, (gtr, rtr)
    if len(rtr):
        assert np.all(np.abs(gtr - rtr) <= cost_tol(rtr, n, rtol)), (gtr, rtr)
    assert out.initial_cost == pytest.approx(ref["initial_cost"], rel=1e-12)
    assert abs(out.final_cost - ref["final_cost"]) <= cost_tol(ref["final_cost"], n, rtol)
    assert np.allclose(out.points, ref["points"], rtol=0, atol=state_atol)
    assert np.allclose(out.poses, ref["poses"], rtol=0, atol=state_atol)
    k = max(1, state_atol * 1e3)
    assert np.allclose(out.cameras["params"], ref["cameras"]["params"], rtol=1e-8, atol=k)
    return out, ref


def test_c0_matches_golden_trajectory(c0_scene):
    out = ba.bundle_adjust(c0_scene.inp, ba.BaParams(max_iterations=30))
    assert out.iterations_run == GOLD["iterations_run"]
    assert np.allclose(out.stats["outer_costs"], GOLD["outer_costs"], rtol=1e-9, atol=0)
    assert np.allclose(out.stats["trial_costs"], GOLD["trial_costs"], rtol=1e-9, atol=0)


@pytest.mark.parametrize("model", range(6))
def test_every_camera_model(model):
    sc = scenes.generate("c0", model_id=model, num_images=12, num_points=400, seed=51 + model,
                         dist_sigma=0.01 if model >= 3 else 0.0)
    assert_parity(sc.inp, ba.BaParams(max_iterations=20))


@pytest.mark.parametrize("n_img", [20, 40])
def test_multi_block_reduced_system(n_img):
    # D > 128: several factorization blocks
    sc = scenes.generate("c0", num_images=n_img, num_points=800, seed=61 + n_img)
    assert_parity(sc.inp, ba.BaParams(max_iterations=15))


@pytest.mark.parametrize("model,fixed", [
    (2, (1, 2, 3)), (2, (2, 3)), (2, (3,)), (2, ()),  # SIMPLE_RADIAL: blocks of 7 to 10 unknowns
    (4, (2, 3)), (4, (3,)), (4, ()),                  # OPENCV: blocks of 12, 13 and 14
])
def test_parameter_block_sizes_on_both_schur_forms(model, fixed):
    # A pose block (6) and its own camera's free parameters form one block.
    # The Schur assembly reads incidence records of 8, 12 or 16 rows (the
    # largest block rounded up): blocks of at most 8 take its four-pairs-per-
    # MFMA form, larger ones its two-pairs form.
    n_params = 4 if model == 2 else 8
    n_free = n_params - len(fixed)
    sc = scenes.generate("c0", model_id=model, num_points=400, seed=13 + 7 * model + n_free,
                         dist_sigma=0.01 if model >= 3 else 0.0)
    inp = sc.inp
    inp.fixed_camera_params = [[k in fixed for k in range(n_params)] for _ in range(len(inp.cameras))]
    out, _ = assert_parity(inp, ba.BaParams(max_iterations=12 if model == 2 else 8))
    # pose 0 is fixed, so its camera is a block of its own
    assert out.stats["dim"] == 15 * (6 + n_free) + n_free


def test_opencv_c1_shaped_small():
    sc = scenes.generate("c1", num_images=60, num_points=4000)
    assert_parity(sc.inp, ba.BaParams(max_iterations=8))


def test_shared_free_camera():
    sc = scenes.generate("c0", shared_camera=1, num_points=500, seed=5)
    assert_parity(sc.inp, ba.BaParams(max_iterations=20))


def test_fixed_cameras_and_no_masks():
    sc = scenes.generate("c0", num_points=300, seed=7)
    inp = sc.inp
    inp.fixed_cameras = np.ones(16, dtype=bool)
    assert_parity(inp, ba.BaParams(max_iterations=10))
    sc = scenes.generate("c0", num_points=300, seed=8)
    inp = sc.inp
    inp.fixed_camera_params = []
    inp.fixed_cameras = [True, False]  # shorter than cameras: the rest are free
    assert_parity(inp, ba.BaParams(max_iterations=10))


@pytest.mark.parametrize("loss", [ba.RobustLoss.NONE, ba.RobustLoss.CAUCHY])
def test_losses(loss):
    sc = scenes.generate("c0", num_points=400, seed=9, outlier_frac=0.05, outlier_px=30.0)
    assert_parity(sc.inp, ba.BaParams(max_iterations=10, robust_loss=loss, loss_scale_px=1.0))


def test_behind_camera_and_unobserved_points():
    sc = scenes.generate("c0", num_points=300, seed=10)
    inp = sc.inp
    pts = np.asarray(inp.points).copy()
    pts[5] = [50.0, 0.0, 0.0]     # outside the camera circle: behind some cameras
    pts = np.vstack([pts, [[0.1, 0.2, 0.3]]])  # a point nobody observes
    inp.points = pts
    assert_parity(inp, ba.BaParams(max_iterations=10))


def test_long_tracks_across_linearization_spans():
    # Tracks of 60-139 observations, whose observations (in point order)
    # cross one or two boundaries of the linearization's 64-observation
    # spans, so their blocks are summed by the separate accumulation pass;
    # 30 points of three observations after them, most of which lie in one
    # span (summed inside the linearization); and an unobserved point.
    sc = scenes.generate("c0", num_images=140, num_points=100, seed=77, track_kind=0, track_min=60,
                         track_max=139)
    inp = sc.inp
    obs, pts = inp.observations, np.asarray(inp.points)
    extra = []
    for k in range(30):
        o = obs[obs["point_idx"] == k][:3].copy()
        o["point_idx"] = 100 + k
        extra.append(o)
    inp.observations = np.concatenate([obs] + extra)
    inp.points = np.vstack([pts, pts[:30], [[0.1, 0.2, 0.3]]])
    counts = np.bincount(inp.observations["point_idx"].astype(int), minlength=len(inp.points))
    end = np.cumsum(counts)
    begin = end - counts
    whole = (counts > 0) & (begin // 64 == (end - 1) // 64)
    assert (counts > 128).sum() > 0 and whole.sum() > 20 and counts[-1] == 0
    assert_parity(inp, ba.BaParams(max_iterations=5))


def test_duplicate_observations():
    sc = scenes.generate("c0", num_points=200, seed=12)
    inp = sc.inp
    inp.observations = np.concatenate([inp.observations, inp.observations[:40]])
    assert_parity(inp, ba.BaParams(max_iterations=10))


def test_all_fixed_and_empty():
    sc = scenes.generate("c0", num_points=100, seed=13)
    inp = sc.inp
    inp.fixed_poses = np.ones(16, dtype=bool)
    inp.fixed_cameras = np.ones(16, dtype=bool)
    out, ref = assert_parity(inp, ba.BaParams())
    assert out.iterations_run == 0
    inp2 = scenes.generate("c0", num_points=100, seed=13).inp
    inp2.observations = inp2.observations[:0]
    out, ref = assert_parity(inp2, ba.BaParams())
    assert out.iterations_run == 1 and out.final_cost == 0.0


def test_ungauged_problem_uses_lu_fallback():
    # No fixed pose, tiny lambda: S is (near) singular, Cholesky fails and the
    # LU path (lib.rs:1083-1093) or the lambda bump takes over.
    sc = scenes.generate("c0", num_images=6, num_points=120, seed=14, pixel_sigma=0.0)
    inp = sc.inp
    inp.fixed_poses = []
    inp.fixed_cameras = np.ones(6, dtype=bool)
    out = ba.bundle_adjust(inp, ba.BaParams(max_iterations=6, initial_lambda=1e-14))
    ref = oracle.bundle_adjust(inp, ba.BaParams(max_iterations=6, initial_lambda=1e-14))
    assert out.final_cost <= out.initial_cost and ref["final_cost"] <= ref["initial_cost"]
    assert out.final_cost == pytest.approx(ref["final_cost"], rel=1e-3, abs=1e-6)


def test_recovers_ground_truth_reference_scenario():
    inp, true_poses, pts = R.recover_scene()
    out, _ = assert_parity(inp, ba.BaParams(max_iterations=30))
    assert out.final_cost < out.initial_cost * 0.01
    t_rec, t_true = out.poses[1, 4:], true_poses[1, 4:]
    assert np.dot(t_rec / np.linalg.norm(t_rec), t_true / np.linalg.norm(t_true)) > 0.999
    assert abs(np.linalg.norm(t_rec) / np.linalg.norm(t_true) - 1.0) < 0.1


def test_joint_focal_reference_scenario():
    pre, joint = R.focal_scenes()
    p1 = ba.bundle_adjust(pre, ba.BaParams(max_iterations=50))
    out = ba.bundle_adjust(joint(p1.poses, p1.points), ba.BaParams(max_iterations=50))
    assert abs(out.cameras[0]["params"][0] - 1500.0) / 1500.0 < 0.02


def test_input_not_modified(c0_scene):
    before = [np.asarray(a).copy() for a in (c0_scene.inp.poses, c0_scene.inp.points, c0_scene.inp.cameras)]
    ba.bundle_adjust(c0_scene.inp, ba.BaParams(max_iterations=3))
    for b, a in zip(before, (c0_scene.inp.poses, c0_scene.inp.points, c0_scene.inp.cameras)):
        assert np.array_equal(b, np.asarray(a))


def test_failed_factorization_does_not_poison_later_solves():
    # Regression: a Cholesky failure (NaN factor) must not leave NaN in the
    # padding of the reduced system; a later solve on the same resident
    # problem has to reproduce a fresh solve exactly.
    # A negative damping makes the first pivot B00 (1 + lambda) < 0: the
    # factorization fails with NaN (sqrt of a negative) on every try.
    sc = scenes.generate("c0", num_images=20, num_points=300, seed=14)
    inp = sc.inp
    prob = ba.Problem(inp)
    bad = prob.solve(ba.BaParams(max_iterations=1, initial_lambda=-2.0))
    assert bad.stats["num_cholesky_failures"] >= 1
    again = prob.solve(ba.BaParams(max_iterations=10))
    fresh = ba.bundle_adjust(inp, ba.BaParams(max_iterations=10))
    assert np.array_equal(again.stats["outer_costs"], fresh.stats["outer_costs"])
    assert np.array_equal(again.points, fresh.points)
    for precision in ("mixed", "fp64"):
        a = prob.solve(ba.BaParams(max_iterations=5), precision=precision)
        b = prob.solve(ba.BaParams(max_iterations=5), precision=precision)
        assert np.array_equal(a.stats["outer_costs"], b.stats["outer_costs"])
