"""Write tests/golden/c0_trajectory.json: the 16-camera SIMPLE_RADIAL scene's
SHA-256 and the oracle's per-iteration and per-try cost trajectory under
BaParams{max_iterations: 30} (the benchmark's first configuration).

Run: python tests/golden/make_c0_trajectory.py
"""
import json
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import oracle  # noqa: E402
from sfmtory_amd import scenes  # noqa: E402
from sfmtory_amd.ba import BaParams  # noqa: E402

sc = scenes.generate("c0")
out = oracle.bundle_adjust(sc.inp, BaParams(max_iterations=30))
data = {"scene": "c0", "sha256": sc.sha256(), "num_observations": int(len(sc.inp.observations)),
        "initial_cost": out["initial_cost"], "final_cost": out["final_cost"],
        "iterations_run": out["iterations_run"], "outer_costs": out["outer_costs"].tolist(),
        "trial_costs": out["trial_costs"].tolist()}
Path(__file__).with_name("c0_trajectory.json").write_text(json.dumps(data, indent=1))
print(data)
