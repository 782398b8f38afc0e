"""FP32 shifted storage against FP64 on the 512^3 configuration.

512^3 periodic in boxes of 64^3, rho = 1 + 1e-4 N(0,1), u_a ~ U(-0.005,
0.005) (seed 512), tau 0.55, 2000 steps, both precisions from the same
initial state.  Reported: max and RMS of |rho32 - rho64| / rho64, of
|u32 - u64| / max|u64|, and the FP64 relative mass drift.  Bar: max relative
density error <= 1e-6.
"""
import json
import os

import numpy as np
import pytest

from lambrex_amd import lbx, workloads

pytestmark = pytest.mark.gpu


def run(ctx, dims, boxes, rho, u, precision, steps):
    sim = lbx.Sim(ctx, dims, precision=precision, rho0=1.0)
    try:
        sim.define_level(0, boxes)
        sim.set_tau(0, 0.55)
        sim.init_from_fields(rho, u)
        c0 = sim.conservation()
        sim.iterate(steps)
        c1 = sim.conservation()
        r, v = sim.get_hydro()
    finally:
        sim.close()
    return r, v, c0, c1


@pytest.mark.timeout(600)
def test_fp32_shifted_vs_fp64_512cubed(ctx):
    n, steps = 512, 2000
    dims = (n, n, n)
    boxes = workloads.chop(dims, 64)
    assert len(boxes) == 512
    rho, u = workloads.random_state(dims, seed=512, rho_sigma=1e-4, u_amp=0.005)
    r64, u64, c0, c1 = run(ctx, dims, boxes, rho, u, lbx.LBX_FP64, steps)
    r32, u32, _, _ = run(ctx, dims, boxes, rho, u, lbx.LBX_FP32_SHIFTED, steps)
    drho = np.abs(r32 - r64) / r64
    umax = np.abs(u64).max()
    du = np.abs(u32 - u64) / umax
    report = {
        "config": "512^3 periodic, 512 boxes of 64^3, tau 0.55, 2000 steps, seed 512",
        "rho_rel_err_max": float(drho.max()), "rho_rel_err_rms": float(np.sqrt((drho ** 2).mean())),
        "u_err_over_umax_max": float(du.max()), "u_err_over_umax_rms": float(np.sqrt((du ** 2).mean())),
        "u64_max": float(umax),
        "fp64_mass_drift_rel": float(abs(c1[0] - c0[0]) / c0[0]),
    }
    print(json.dumps(report))
    out = os.environ.get("LBX_REPORT_DIR")
    if out:
        with open(os.path.join(out, "precision_512cubed.json"), "w") as fh:
            json.dump(report, fh, indent=1)
    assert report["rho_rel_err_max"] <= 1e-6
    assert report["fp64_mass_drift_rel"] <= 1e-12
