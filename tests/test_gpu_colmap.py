"""The refine pass over a COLMAP text model (sfmba_colmap_refine) against the
oracle restatement of cmd_refine (crates/sfm-cli/src/main.rs:946-1130):
the same compact problem, the same LM trajectory, the refined model written
back with recomputed per-point errors, and the before/after mean
reprojection error the reference logs."""
import numpy as np
import pytest

import colmap_models as M
from oracle import colmap as OC
from sfmtory_amd import ba, colmap, scenes

pytestmark = pytest.mark.gpu


def compare_models(a_dir, b_dir):
    a, b = OC.read_model(a_dir), OC.read_model(b_dir)
    assert a["cameras"].keys() == b["cameras"].keys()
    assert a["images"].keys() == b["images"].keys()
    assert a["points"].keys() == b["points"].keys()
    for cid in a["cameras"]:
        assert np.allclose(a["cameras"][cid]["params"], b["cameras"][cid]["params"], rtol=1e-9, atol=2e-10)
    for iid in a["images"]:
        ia, ib = a["images"][iid], b["images"][iid]
        assert np.allclose(ia["q"] + ia["t"], ib["q"] + ib["t"], rtol=0, atol=1e-8)
        assert ia["kx"] == ib["kx"] and ia["ky"] == ib["ky"] and ia["pid"] == ib["pid"]
    for pid in a["points"]:
        pa, pb = a["points"][pid], b["points"][pid]
        assert np.allclose(pa["xyz"], pb["xyz"], rtol=1e-9, atol=1e-8)
        assert abs(pa["error"] - pb["error"]) <= 2e-6
        assert pa["track"] == pb["track"] and pa["rgb"] == pb["rgb"]


def check_report(rep, ref):
    for k in ("num_images", "num_points3d", "num_observations", "iterations_run"):
        assert rep[k] == ref[k], k
    assert rep["initial_cost"] == pytest.approx(ref["initial_cost"], rel=1e-11)
    assert rep["final_cost"] == pytest.approx(ref["final_cost"], rel=1e-9)
    assert rep["mean_reprojection_error_px_before"] == ref["mean_reprojection_error_px_before"]
    assert rep["mean_reprojection_error_px_after"] == pytest.approx(ref["mean_reprojection_error_px_after"],
                                                                     rel=1e-7)


@pytest.mark.parametrize("refine_intrinsics,loss,model_id,shared", [
    (False, ba.RobustLoss.HUBER, 2, 0),
    (True, ba.RobustLoss.HUBER, 2, 1),
    (True, ba.RobustLoss.CAUCHY, 4, 1),
    (False, ba.RobustLoss.CAUCHY, 5, 0),
])
def test_refine_matches_oracle(tmp_path, refine_intrinsics, loss, model_id, shared):
    sc = scenes.generate("c0", num_images=10, num_points=400, seed=90 + model_id, model_id=model_id,
                         shared_camera=shared)
    OC.write_model(M.scene_model(sc.inp, seed=model_id), tmp_path / "in")
    rep = colmap.refine_model(tmp_path / "in", tmp_path / "gpu", refine_intrinsics=refine_intrinsics,
                              robust_loss=loss)
    ref = OC.refine(tmp_path / "in", tmp_path / "cpu", refine_intrinsics=refine_intrinsics, robust_loss=loss)
    check_report(rep, ref)
    assert rep["mean_reprojection_error_px_after"] < rep["mean_reprojection_error_px_before"]
    compare_models(tmp_path / "gpu", tmp_path / "cpu")
    if not refine_intrinsics:
        assert (tmp_path / "gpu" / "cameras.txt").read_bytes() == (tmp_path / "in" / "cameras.txt").read_bytes()


def test_refine_in_place_and_again(tmp_path):
    # Output defaults to the input directory, as the reference writes back to
    # sparse/0; a second pass starts from the refined model and converges
    # to a cost no higher than the first pass's.
    sc = scenes.generate("c0", num_images=8, num_points=300, seed=97)
    d = tmp_path / "model"
    OC.write_model(M.scene_model(sc.inp), d)
    first = colmap.refine_model(d)
    second = colmap.refine_model(d)
    assert second["initial_cost"] <= first["final_cost"] * (1 + 1e-6)
    assert second["final_cost"] <= second["initial_cost"]
    assert second["mean_reprojection_error_px_before"] == pytest.approx(first["mean_reprojection_error_px_after"],
                                                                        abs=1e-6)


def test_refine_mixed_precision_reaches_the_fp64_cost(tmp_path):
    sc = scenes.generate("c0", num_images=12, num_points=600, seed=98)
    OC.write_model(M.scene_model(sc.inp), tmp_path / "in")
    fp64 = colmap.refine_model(tmp_path / "in", tmp_path / "a")
    mixed = colmap.refine_model(tmp_path / "in", tmp_path / "b", precision="mixed")
    assert mixed["final_cost"] == pytest.approx(fp64["final_cost"], rel=1e-5)
    assert mixed["mean_reprojection_error_px_after"] == pytest.approx(fp64["mean_reprojection_error_px_after"],
                                                                      rel=1e-4)
