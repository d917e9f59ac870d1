"""Batched local-BA windows (one persistent launch, one workgroup per
window) against the oracle: per window, the FP64 per-iteration costs within
1e-9 relative and the same iteration count as the reference algorithm;
camera models, losses, fixed cameras, the LU fallback inside the kernel, and
problems too large for one workgroup routed to the single-problem path."""
from collections import Counter

import numpy as np
import pytest

import oracle
from sfmtory_amd import ba, scenes
from test_gpu_ba_parity import cost_tol

pytestmark = pytest.mark.gpu


def check(inp, out, params, rtol=1e-9, state_atol=1e-6):
    ref = oracle.bundle_adjust(inp, params)
    g, r = np.asarray(out.stats["outer_costs"]), ref["outer_costs"]
    n = max(1, len(inp.observations))
    assert out.iterations_run == ref["iterations_run"], (g, r)
    assert g.shape == r.shape, (g, r)
    if len(r):
        assert np.all(np.abs(g - r) <= cost_tol(r, n, rtol)), (g, r)
    assert out.initial_cost == pytest.approx(ref["initial_cost"], rel=1e-12)
    assert abs(out.final_cost - ref["final_cost"]) <= cost_tol(ref["final_cost"], n, rtol)
    assert np.allclose(out.points, ref["points"], rtol=0, atol=state_atol)
    assert np.allclose(out.poses, ref["poses"], rtol=0, atol=state_atol)
    assert np.allclose(out.cameras["params"], ref["cameras"]["params"], rtol=1e-8, atol=1e-3)


def small_windows(count, seed, **kw):
    rng = np.random.default_rng(seed)
    return [scenes.generate("c3w", seed=seed * 100 + w, num_images=int(rng.integers(4, 13)),
                            num_points=int(rng.integers(150, 700)), **kw) for w in range(count)]


def test_fp64_windows_match_oracle():
    wins = small_windows(6, 11)
    p = ba.BaParams(max_iterations=15)
    outs = ba.bundle_adjust_batch([w.inp for w in wins], p)
    for w, o in zip(wins, outs):
        check(w.inp, o, p)


@pytest.mark.parametrize("model", [0, 1, 3, 4, 5])
def test_fp64_windows_every_model(model):
    wins = small_windows(3, 20 + model, model_id=model, dist_sigma=0.005, k_fixed=0.0)
    p = ba.BaParams(max_iterations=8)
    outs = ba.bundle_adjust_batch([w.inp for w in wins], p)
    for w, o in zip(wins, outs):
        check(w.inp, o, p)


@pytest.mark.parametrize("loss", [ba.RobustLoss.NONE, ba.RobustLoss.CAUCHY])
def test_fp64_windows_losses_and_fixed_cameras(loss):
    wins = small_windows(3, 31, outlier_frac=0.02, outlier_px=30.0)
    wins[1].inp.fixed_cameras = np.ones(len(wins[1].inp.cameras), dtype=bool)
    wins[2].inp.fixed_camera_params = []
    p = ba.BaParams(max_iterations=10, robust_loss=loss)
    outs = ba.bundle_adjust_batch([w.inp for w in wins], p)
    for w, o in zip(wins, outs):
        check(w.inp, o, p)


def test_lu_fallback_inside_the_window_kernel():
    # A negative damping makes the damped reduced system indefinite: the
    # window kernel's Cholesky fails and its LU solve (lib.rs:1083-1093)
    # takes the step, as the reference's chain does.
    sc = scenes.generate("c0", num_images=8, num_points=200, seed=14)
    p = ba.BaParams(max_iterations=4, initial_lambda=-2.0)
    out = ba.bundle_adjust_batch([sc.inp], p)[0]
    assert out.stats["num_cholesky_failures"] >= 1 and out.stats["num_lu_solves"] >= 1
    check(sc.inp, out, p)
    single = ba.bundle_adjust(sc.inp, p)
    assert single.stats["num_lu_solves"] == out.stats["num_lu_solves"]


def test_mixed_windows_close_to_fp64():
    wins = small_windows(4, 41)
    p = ba.BaParams(max_iterations=20)
    f64 = ba.bundle_adjust_batch([w.inp for w in wins], p, precision="fp64")
    mix = ba.bundle_adjust_batch([w.inp for w in wins], p, precision="mixed")
    for a, b, w in zip(f64, mix, wins):
        assert b.final_cost == pytest.approx(a.final_cost, rel=1e-5)
        # plain RMS reprojection error from the residuals, by the oracle
        ra = oracle.rms_reprojection_error(ba.BaInput(w.inp.camera_of_image, a.cameras, a.poses, a.points,
                                                      w.inp.observations))
        rb = oracle.rms_reprojection_error(ba.BaInput(w.inp.camera_of_image, b.cameras, b.poses, b.points,
                                                      w.inp.observations))
        assert abs(ra - rb) < 1e-4


def test_long_destination_chunks():
    # Four images and 700 points: every image pair shares 400-560 points, so
    # the window solver's one chunk per destination block spans several
    # 128-pair blocks of the matrix-core assembly (blocks of 8 unknowns).
    wins = [scenes.generate("c3w", seed=4200 + k, num_images=4, num_points=700) for k in range(2)]
    for w in wins:
        obs = w.inp.observations
        seen = {}
        for img, pt in zip(obs["image_idx"], obs["point_idx"]):
            seen.setdefault(int(pt), set()).add(int(img))
        shared = Counter((a, b) for s in seen.values() for a in s for b in s if a > b)
        assert min(shared.values()) > 256
    p = ba.BaParams(max_iterations=10)
    f64 = ba.bundle_adjust_batch([w.inp for w in wins], p)
    for w, o in zip(wins, f64):
        check(w.inp, o, p)
    mix = ba.bundle_adjust_batch([w.inp for w in wins], p, precision="mixed")
    for a, b in zip(f64, mix):
        assert b.final_cost == pytest.approx(a.final_cost, rel=1e-5)


def test_large_problem_in_a_batch_takes_the_single_problem_path():
    big = scenes.generate("c0", num_images=140, num_points=3000, seed=5)   # D > 1024
    small = small_windows(2, 51)
    p = ba.BaParams(max_iterations=5)
    outs = ba.bundle_adjust_batch([small[0].inp, big.inp, small[1].inp], p)
    assert outs[1].stats["dim"] > 1024
    single = ba.bundle_adjust(big.inp, p)
    assert outs[1].iterations_run == single.iterations_run
    assert outs[1].final_cost == single.final_cost
    for w, o in zip(small, (outs[0], outs[2])):
        check(w.inp, o, p)


def test_empty_and_all_fixed_windows():
    sc = small_windows(1, 61)[0]
    fixed = scenes.generate("c3w", seed=99, num_images=5, num_points=200)
    fixed.inp.fixed_poses = np.ones(5, dtype=bool)
    fixed.inp.fixed_cameras = np.ones(len(fixed.inp.cameras), dtype=bool)
    p = ba.BaParams(max_iterations=5)
    outs = ba.bundle_adjust_batch([sc.inp, fixed.inp], p)
    check(sc.inp, outs[0], p)
    assert outs[1].iterations_run == 0 and outs[1].final_cost == outs[1].initial_cost
    assert ba.bundle_adjust_batch([], p) == []
