"""Batch plans (sfmba_batch_create / sfmba_batch_solve): a batch of windows
prepared once and solved by replaying a captured HIP graph, with the window
kernel's launch shape optionally chosen by the autotuner.

Every replay is checked against the eager batched call
(sfmba_bundle_adjust_batch) on the same inputs and launch shape, with eager
GPU work and host synchronizations between replays (a graph must give the
same result on every replay, not only the first), and against the oracle."""
import numpy as np
import pytest

from sfmtory_amd import ba, scenes
from test_gpu_windows import check, small_windows

pytestmark = pytest.mark.gpu


def same(a, b):
    assert a.iterations_run == b.iterations_run
    assert a.initial_cost == b.initial_cost
    assert a.final_cost == b.final_cost
    assert np.array_equal(a.points, b.points)
    assert np.array_equal(a.poses, b.poses)
    assert np.array_equal(a.cameras["params"], b.cameras["params"])
    assert np.array_equal(a.stats["outer_costs"], b.stats["outer_costs"])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_graph_replays_match_the_eager_batch(precision):
    wins = small_windows(7, 71)
    big = scenes.generate("c0", num_images=140, num_points=3000, seed=9)   # D > 1024: single-problem path
    inputs = [w.inp for w in wins[:3]] + [big.inp] + [w.inp for w in wins[3:]]
    p = ba.BaParams(max_iterations=12)
    eager = ba.bundle_adjust_batch(inputs, p, precision=precision)
    with ba.BatchPlan(inputs, p, precision=precision) as plan:
        info = plan.info()
        assert info["num_windows"] == 7 and info["num_large"] == 1 and not info["tuned"]
        for replay in range(3):
            outs = plan.solve()
            assert len(outs) == len(inputs)
            for e, o in zip(eager, outs):
                same(e, o)
            # eager work between replays
            ba.bundle_adjust_batch(inputs[:2], p, precision=precision)
            ba.total_cost(inputs[0], p)
    if precision == "fp64":
        for k in (0, 4):
            check(inputs[k], outs[k], p)
        assert outs[3].stats["dim"] > 1024


def test_autotuner_times_every_shape_and_keeps_the_fastest():
    wins = small_windows(10, 81)
    inputs = [w.inp for w in wins]
    p = ba.BaParams(max_iterations=10)
    with ba.BatchPlan(inputs, p, autotune_iterations=2) as plan:
        info = plan.info()
        assert info["tuned"]
        secs = [s["seconds"] for s in info["shapes"]]
        assert len(secs) == 6 and all(t > 0 for t in secs)
        best = info["shapes"][int(np.argmin(secs))]
        assert (info["threads_per_window"], info["pairs_in_flight"]) == (best["threads"], best["pairs"])
        outs = plan.solve()
        again = plan.solve()
    for w, o, a in zip(wins, outs, again):
        check(w.inp, o, p)
        same(o, a)
    # The eager call on a batch of the same size class uses the tuned shape:
    # same shape, same inputs, same results.
    for e, o in zip(ba.bundle_adjust_batch(inputs, p), outs):
        same(e, o)


def test_empty_plan_and_argument_errors():
    p = ba.BaParams(max_iterations=3)
    with ba.BatchPlan([], p) as plan:
        assert plan.solve() == []
        assert plan.info()["num_windows"] == 0
    with pytest.raises(ValueError):
        ba.BatchPlan([small_windows(1, 91)[0].inp], p, autotune_iterations=-1)
