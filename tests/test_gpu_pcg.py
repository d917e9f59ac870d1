"""Block-Jacobi PCG on the reduced system (solver="pcg", and "pcg_mixed",
whose iterations read an FP32 copy of the system with FP64 residual
replacement) against the dense Cholesky path and the oracle: with a tight
tolerance the LM trajectory is the factorization's; with the default
tolerance the solve converges to the same cost; a breakdown counts as a
failed try (damping bump, no LU)."""
import numpy as np
import pytest

import oracle
from sfmtory_amd import _native as N
from sfmtory_amd import ba, scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("solver", ["pcg", "pcg_mixed"])
def test_tight_pcg_follows_the_oracle_trajectory(c0_scene, solver):
    p = ba.BaParams(max_iterations=15)
    out = ba.bundle_adjust(c0_scene.inp, p, solver=solver, pcg_tolerance=1e-13, pcg_max_iterations=2000)
    ref = oracle.bundle_adjust(c0_scene.inp, p)
    assert out.iterations_run == ref["iterations_run"]
    g, r = np.asarray(out.stats["outer_costs"]), ref["outer_costs"]
    assert g.shape == r.shape
    assert np.allclose(g, r, rtol=1e-8, atol=0)
    assert out.stats["num_pcg_iterations"] > 0


@pytest.mark.parametrize("model", [2, 4])
def test_pcg_converges_to_the_cholesky_cost_multi_block(model):
    sc = scenes.generate("c0", num_images=60, num_points=3000, seed=71, model_id=model, dist_sigma=0.005,
                         k_fixed=0.0)
    p = ba.BaParams(max_iterations=25)
    chol = ba.bundle_adjust(sc.inp, p)
    it = ba.bundle_adjust(sc.inp, p, solver="pcg")
    assert chol.stats["dim"] > 256
    assert it.final_cost == pytest.approx(chol.final_cost, rel=1e-6)
    assert it.final_cost < 0.2 * it.initial_cost
    mixed = ba.bundle_adjust(sc.inp, p, solver="pcg", precision="mixed")
    assert mixed.final_cost == pytest.approx(chol.final_cost, rel=1e-5)
    stored32 = ba.bundle_adjust(sc.inp, p, solver="pcg_mixed")
    assert stored32.final_cost == pytest.approx(chol.final_cost, rel=1e-6)
    assert stored32.stats["num_pcg_iterations"] > 0


def test_pcg_mixed_meets_the_tolerance_on_the_true_residual():
    # every PCG solve of the FP32-stored form ends on an FP64 residual, so a
    # loose iteration cap still gives the FP64 PCG's accepted steps
    sc = scenes.generate("c0", num_images=40, num_points=2000, seed=5)
    p = ba.BaParams(max_iterations=10)
    f64 = ba.bundle_adjust(sc.inp, p, solver="pcg", pcg_tolerance=1e-9)
    m32 = ba.bundle_adjust(sc.inp, p, solver="pcg_mixed", pcg_tolerance=1e-9)
    assert m32.iterations_run == f64.iterations_run
    assert np.allclose(m32.stats["outer_costs"], f64.stats["outer_costs"], rtol=1e-7, atol=0)


def test_pcg_breakdown_bumps_damping_without_lu():
    sc = scenes.generate("c0", num_images=20, num_points=300, seed=14)
    p = ba.BaParams(max_iterations=4, initial_lambda=-2.0)
    out = ba.bundle_adjust(sc.inp, p, solver="pcg")
    assert out.stats["num_cholesky_failures"] >= 1
    assert out.stats["num_lu_solves"] == 0
    assert out.final_cost <= out.initial_cost


def test_pcg_options_are_validated(c0_scene):
    with pytest.raises(ValueError):
        ba.bundle_adjust(c0_scene.inp, solver="pcg", pcg_max_iterations=0)
    with pytest.raises(ValueError):
        ba.bundle_adjust(c0_scene.inp, solver="pcg_mixed", pcg_tolerance=-1.0)
    for solver in ("pcg", "pcg_mixed"):
        with pytest.raises(N.EngineError) as e:
            ba.bundle_adjust_batch([c0_scene.inp], solver=solver)
        assert e.value.code == N.SFMBA_ERR_UNSUPPORTED
