"""Write tests/golden/pulse_profile.json from the reference's pulse golden data.

The reference test ``tests/catch2RegressionTests.cpp:6-93`` compares density
and velocity of a 10x10x50 periodic pulse (tau = 0.5) at t = 0, 100 and 200
against the arrays in ``tests/pulseRegression.h:6-22`` (RHO_t*: 5000 values,
k slowest / i fastest; VEL_t*: 15000 values, component n fastest).  Every
array is exactly uniform over (i, j) for each k, so the 519 KB header reduces
to 50 z-values x {rho, ux, uy, uz} x {t0, t100, t200}.  This script checks the
uniformity string by string and keeps the printed strings verbatim (they are
what Catch2's ``Approx`` compares against after parsing).

Run from the repository root (reads the reference checkout read-only):
    python tests/golden/make_pulse_fixture.py /root/reference/tests/pulseRegression.h
"""
import json
import re
import sys
from pathlib import Path

NX, NY, NZ = 10, 10, 50


def parse_arrays(text):
    out = {}
    pattern = re.compile(r"std::array<double,\s*(\d+)>\s+(\w+)\s*=\s*\{(.*?)\};", re.S)
    for size, name, body in pattern.findall(text):
        toks = [t.strip() for t in body.split(",")]
        if len(toks) != int(size):
            raise ValueError(f"{name}: {len(toks)} entries, header says {size}")
        out[name] = toks
    return out


def z_profile_scalar(toks):
    prof = []
    for k in range(NZ):
        plane = toks[k * NX * NY:(k + 1) * NX * NY]
        if any(t != plane[0] for t in plane):
            raise ValueError(f"density not xy-uniform at k={k}")
        prof.append(plane[0])
    return prof


def z_profile_vector(toks):
    prof = [[], [], []]
    for k in range(NZ):
        for n in range(3):
            vals = [toks[(k * NX * NY + ij) * 3 + n] for ij in range(NX * NY)]
            if any(v != vals[0] for v in vals):
                raise ValueError(f"velocity component {n} not xy-uniform at k={k}")
            prof[n].append(vals[0])
    return prof


def main(header_path):
    arrays = parse_arrays(Path(header_path).read_text())
    fixture = {
        "source": "tests/pulseRegression.h:6-22 (xy-uniform z-profiles, strings verbatim)",
        "case": {"nx": NX, "ny": NY, "nz": NZ, "tau": 0.5, "amplitude": 0.01,
                 "pulse_plane_k": NZ // 2 - 1, "times": [0, 100, 200]},
        "rho": {}, "u": {},
    }
    for t in (0, 100, 200):
        fixture["rho"][str(t)] = z_profile_scalar(arrays[f"RHO_t{t}"])
        fixture["u"][str(t)] = z_profile_vector(arrays[f"VEL_t{t}"])
    out = Path(__file__).with_name("pulse_profile.json")
    out.write_text(json.dumps(fixture, indent=1) + "\n")
    print(f"wrote {out}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/tests/pulseRegression.h")
