"""bench.py on the GPU with tiny settings: the output is one line of strict
JSON (no NaN) carrying the contract's fields.  The untimed clock warm-up
leaves the state post-collision, so even a one-step timed region is one
fused pass whose duration the roofline uses."""
import json
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
pytestmark = pytest.mark.gpu


def _no_constants(name):
    raise ValueError(f"non-JSON constant {name}")


@pytest.mark.parametrize("steps", [1, 3])
def test_bench_prints_one_strict_json_line(steps):
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--n", "64", "--steps", str(steps), "--warmup", "0",
                        "--no-cpu", "--no-copy", "--serve-requests", "0", "--two-level-steps", "2",
                        "--three-level-steps", "0", "--walls-steps", "3"], capture_output=True, text=True,
                       timeout=300, cwd=ROOT)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = [line for line in r.stdout.splitlines() if line.strip()]
    assert len(lines) == 1
    out = json.loads(lines[0], parse_constant=_no_constants)
    for key in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better",
                "scaling", "vs_baseline", "dtype", "data", "config", "roofline"):
        assert key in out, key
    assert out["n_gpus"] == 1 and out["steps"] == steps and out["value"] > 0
    rf = out["roofline"]
    assert rf["bound"] == "hbm" and rf["peak"] == 8000.0 and rf["unit"] == "GB/s"
    assert rf["achieved"] > 0 and 0 < rf["frac"] < 1
    assert out["clock_warmup_ms"] > 0
    assert out["two_level"]["mlups"] > 0
    tr = out["two_level"]["roofline"]
    assert tr["bound"] == "hbm" and 0 < tr["algorithmic_frac"] < 1
    assert tr["measured_gbs"] > tr["algorithmic_gbs"]  # the step moves more than 240 B per update
    w = out["walls"]
    assert w["periodic"] == [1, 1, 0] and w["mlups"] > 0 and 0 < w["frac_of_hbm_peak"] < 1
    assert w["mass_drift_rel"] < 1e-12
