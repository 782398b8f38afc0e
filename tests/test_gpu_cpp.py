"""The C++ AmrSim facade on the GPU, driven the way the reference's own tests
and example drive AmrSim (tests/cpp/*.cpp, built by build())."""
import json
import re
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

HERE = Path(__file__).resolve().parent
BIN = HERE / "cpp" / "_bin"
GOLDEN = HERE / "golden" / "pulse_profile.json"


def run(name, timeout=120):
    exe = BIN / name
    assert exe.exists(), f"{exe} not built (run __graft_entry__.build())"
    r = subprocess.run([str(exe)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=timeout)
    return r.returncode, r.stdout


def test_init_and_single_level_cases():
    rc, out = run("init_tests")
    print(out)
    assert rc == 0, out
    assert "0 failed" in out


def test_two_level_amr_cases():
    rc, out = run("amr_tests")
    print(out)
    assert rc == 0, out
    assert "0 failed" in out


def test_amr_pulse_prints_the_reference_goldens():
    """examples/amr_pulse.cpp re-run through the facade: every printed value
    (precision 6, the format the reference's golden header was made with)
    equals the golden string, and matches under Catch2 Approx."""
    rc, out = run("amr_pulse")
    assert rc == 0, out
    blocks = re.findall(r"\{ (.*?) \};", out, re.S)
    assert len(blocks) == 6, out[:2000]
    fx = json.loads(GOLDEN.read_text())
    nx = ny = 10
    nz = 50
    eps = 100 * 1.1920928955078125e-07
    for bi, t in enumerate((0, 100, 200)):
        rho = [s.strip() for s in blocks[2 * bi].split(",")]
        vel = [s.strip() for s in blocks[2 * bi + 1].split(",")]
        assert len(rho) == nx * ny * nz and len(vel) == 3 * nx * ny * nz
        mismatched = 0
        for k in range(nz):
            exp = fx["rho"][str(t)][k]
            for ij in range(nx * ny):
                got = rho[k * nx * ny + ij]
                assert abs(float(got) - float(exp)) <= eps * abs(float(exp)), (t, k, got, exp)
                mismatched += got != exp
            for n in range(3):
                exp = fx["u"][str(t)][n][k]
                for ij in range(nx * ny):
                    got = vel[(k * nx * ny + ij) * 3 + n]
                    assert abs(float(got) - float(exp)) <= eps * abs(float(exp)), (t, k, n, got, exp)
                    mismatched += got != exp
        assert mismatched == 0, f"t={t}: {mismatched} printed values differ from the golden strings"
