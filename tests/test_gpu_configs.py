"""The benchmark workloads' generator presets on the GPU.

Reduced sizes against the oracle, per iteration at 1e-9 (FP64): the
1,000-PINHOLE-camera preset (1% outliers at U(+-50 px), Huber 2 px) and the
5,000-RADIAL-camera preset (heavy-tailed tracks up to 200 cameras) at a
few dozen cameras.  Full sizes through properties the oracle cannot run in
seconds: the mixed solve of the 1,000-camera scene against the FP64 solve
(same iteration count, plain RMS reprojection error within 1e-4 px), and
PCG against Cholesky on the 5,000-camera scene (final cost within 1e-6).
Also the mixed form's FP64 iterative refinement and the single-problem LU
fallback (crates/sfm-ba/src/lib.rs:1076-1094)."""
import numpy as np
import pytest

import oracle
from sfmtory_amd import _native as N
from sfmtory_amd import ba, scenes
from test_gpu_ba_parity import assert_parity

pytestmark = pytest.mark.gpu


def plain_rms(inp, out):
    """RMS of |r| at a solve's final state, evaluated by the oracle."""
    st = ba.BaInput(inp.camera_of_image, out.cameras, out.poses, out.points, inp.observations)
    return oracle.rms_reprojection_error(st)


def test_c2_preset_reduced_matches_oracle():
    # 40 PINHOLE cameras, 1% outliers under Huber (the 1,000-camera preset's generator)
    sc = scenes.generate("c2", num_images=40, num_points=3000, seed=scenes.BASE_SEED + 200)
    assert len(sc.inp.observations) > 15000
    out, ref = assert_parity(sc.inp, ba.BaParams(max_iterations=50))
    assert np.allclose(out.stats["trial_costs"], ref["trial_costs"], rtol=1e-9, atol=0)
    assert out.iterations_run >= 5


def test_c4_preset_reduced_matches_oracle():
    # 60 RADIAL cameras, heavy-tailed tracks (the 5,000-camera preset's generator)
    sc = scenes.generate("c4", num_images=60, num_points=2500, seed=scenes.BASE_SEED + 400)
    tracks = np.bincount(sc.inp.observations["point_idx"].astype(np.int64))
    assert tracks.max() >= 15  # the tail reaches well past the mean
    out, ref = assert_parity(sc.inp, ba.BaParams(max_iterations=30))
    assert np.allclose(out.stats["trial_costs"], ref["trial_costs"], rtol=1e-9, atol=0)


def test_c2_full_mixed_converges_to_the_fp64_rms():
    sc = scenes.generate("c2")
    prob = ba.Problem(sc.inp)
    p = ba.BaParams(max_iterations=50)
    f64 = prob.solve(p, precision="fp64")
    mix = prob.solve(p, precision="mixed")
    prob.close()
    # both stop by the reference's rule (lib.rs:1215-1221), mixed no later
    assert f64.iterations_run < 50
    assert mix.iterations_run <= f64.iterations_run
    assert mix.stats["num_fp32_tries"] > 0
    assert abs(plain_rms(sc.inp, mix) - plain_rms(sc.inp, f64)) < 1e-4
    assert mix.final_cost == pytest.approx(f64.final_cost, rel=1e-6)


def test_c4_full_pcg_matches_cholesky():
    sc = scenes.generate("c4")
    prob = ba.Problem(sc.inp)
    p = ba.BaParams(max_iterations=3)
    chol = prob.solve(p, precision="fp64", solver="cholesky")
    pcg = prob.solve(p, precision="fp64", solver="pcg")
    m32 = prob.solve(p, precision="fp64", solver="pcg_mixed")
    prob.close()
    assert chol.stats["dim"] > 40000
    assert pcg.iterations_run == chol.iterations_run
    assert pcg.final_cost == pytest.approx(chol.final_cost, rel=1e-6)
    assert pcg.final_cost < 0.5 * pcg.initial_cost
    assert m32.iterations_run == chol.iterations_run
    assert m32.final_cost == pytest.approx(chol.final_cost, rel=1e-6)


def test_mixed_trace_switches_to_fp64_once(c0_scene):
    out = ba.bundle_adjust(c0_scene.inp, ba.BaParams(max_iterations=30), precision="mixed")
    oc = out.stats["try_outcomes"]
    lam = out.stats["try_lambdas"]
    fp32 = (oc & N.TRY_FP32) != 0
    assert len(oc) == out.stats["num_factorizations"] and fp32.sum() == out.stats["num_fp32_tries"]
    # FP32 tries come first, at damping >= SFMBA_MIXED_MIN_LAMBDA; every later try is FP64
    k = int(fp32.sum())
    assert fp32[:k].all() and not fp32[k:].any()
    assert np.all(lam[:k] >= N.MIXED_MIN_LAMBDA)
    f64 = ba.bundle_adjust(c0_scene.inp, ba.BaParams(max_iterations=30), precision="fp64")
    assert out.iterations_run == f64.iterations_run
    assert abs(plain_rms(c0_scene.inp, out) - plain_rms(c0_scene.inp, f64)) < 1e-4


def test_mixed_refinement_steps(c0_scene):
    p = ba.BaParams(max_iterations=30)
    base = ba.bundle_adjust(c0_scene.inp, p, precision="mixed")
    ref1 = ba.bundle_adjust(c0_scene.inp, p, precision="mixed", refine_steps=2)
    f64 = ba.bundle_adjust(c0_scene.inp, p, precision="fp64")
    assert ref1.stats["num_fp32_tries"] > 0
    assert ref1.iterations_run == base.iterations_run == f64.iterations_run
    # refining against the assembled system changes the FP32 steps only at
    # rounding level
    assert np.allclose(ref1.stats["outer_costs"], base.stats["outer_costs"], rtol=1e-6, atol=0)
    assert ref1.final_cost == pytest.approx(f64.final_cost, rel=1e-6)


def test_single_problem_lu_fallback_matches_oracle():
    # A negative damping makes every damped reduced system indefinite: the
    # Cholesky fails and the LU solve (lib.rs:1083-1093) gives the step.
    sc = scenes.generate("c0", num_images=20, num_points=300, seed=14)
    out, ref = assert_parity(sc.inp, ba.BaParams(max_iterations=6, initial_lambda=-2.0))
    assert out.stats["num_lu_solves"] >= 1
    assert out.stats["num_cholesky_failures"] >= 1
    assert np.allclose(out.stats["trial_costs"], ref["trial_costs"], rtol=1e-9, atol=0)
    lu = (out.stats["try_outcomes"] & N.TRY_LU) != 0
    assert lu.sum() == out.stats["num_lu_solves"]
