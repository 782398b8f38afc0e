"""The C++ AmrSim facade on the GPU, driven the way the reference's own tests
and example drive AmrSim (tests/cpp/*.cpp, built by build())."""
import json
import re
import subprocess
from pathlib import Path

import numpy as np
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


def test_checkpoint_restart_is_bitwise():
    """WriteCheckpoint after Iterate(a), a fresh AmrSim's ReadCheckpoint and
    Iterate(b) equal Iterate(a + b) bit for bit (one, two and three levels,
    both masks, walls, FP32-shifted)."""
    rc, out = run("checkpoint_restart", timeout=300)
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


def read_plotfile(d):
    """Header and FABs of an AMReX-layout plotfile written by
    AmrSim::WritePlotfile: {"header": ..., "levels": [[(lo, hi, data[4][nz][ny][nx]), ...], ...]}."""
    it = iter((d / "Header").read_text().split("\n"))
    h = {"version": next(it)}
    ncomp = int(next(it))
    h["names"] = [next(it) for _ in range(ncomp)]
    h["dim"] = int(next(it))
    h["time"] = float(next(it))
    h["finest"] = int(next(it))
    h["prob_lo"] = next(it).split()
    h["prob_hi"] = next(it).split()
    h["ref"] = next(it).split()
    h["domains"] = re.findall(r"\(\((\d+),(\d+),(\d+)\) \((\d+),(\d+),(\d+)\) \(0,0,0\)\)", next(it))
    h["steps"] = [int(v) for v in next(it).split()]
    h["dx"] = [[float(v) for v in next(it).split()] for _ in range(h["finest"] + 1)]
    h["coord"] = int(next(it))
    h["bwidth"] = int(next(it))
    h["level_boxes"] = []
    for lev in range(h["finest"] + 1):
        l, nb, t = next(it).split()
        assert int(l) == lev and float(t) == h["time"]
        step = int(next(it))
        phys = [[[float(v) for v in next(it).split()] for _ in range(3)] for _ in range(int(nb))]
        path = next(it)
        h["level_boxes"].append((step, phys, path))
    levels = []
    box_re = r"\(\((-?\d+),(-?\d+),(-?\d+)\) \((-?\d+),(-?\d+),(-?\d+)\) \(0,0,0\)\)"
    for lev in range(h["finest"] + 1):
        ch = (d / f"Level_{lev}" / "Cell_H").read_text()
        lines = ch.split("\n")
        assert lines[:4] == ["1", "1", "4", "0"]
        boxes = [tuple(map(int, m)) for m in re.findall(box_re, ch)]
        offsets = [int(v) for v in re.findall(r"FabOnDisk: Cell_D_00000 (\d+)", ch)]
        stats = re.findall(r"^([-0-9.e+,]+),$", ch, re.M)
        raw = (d / f"Level_{lev}" / "Cell_D_00000").read_bytes()
        fabs = []
        for box, off in zip(boxes, offsets):
            end = raw.index(b"\n", off)
            head = raw[off:end].decode()
            assert head.startswith("FAB ((8, (64 11 52 0 1 12 0 1023)),(8, (8 7 6 5 4 3 2 1)))"), head
            fb = tuple(map(int, re.search(box_re, head).groups()))
            assert fb == box and head.endswith(" 4")
            n = [fb[3 + a] - fb[a] + 1 for a in range(3)]
            data = np.frombuffer(raw, dtype="<f8", count=4 * n[0] * n[1] * n[2], offset=end + 1)
            fabs.append((fb[:3], fb[3:], data.reshape(4, n[2], n[1], n[0])))
        levels.append({"fabs": fabs, "stats": stats, "nboxes": len(boxes)})
    return h, levels


def test_plotfile_round_trip(tmp_path):
    """AmrSim::WritePlotfile (AMReX plotfile layout) holds the same density
    and velocity the facade's GetDensity/GetVelocity return, with consistent
    Header, Cell_H and FAB headers."""
    exe = BIN / "plotfile_dump"
    assert exe.exists()
    out_dir = tmp_path / "plt00003"
    r = subprocess.run([str(exe), str(out_dir)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout
    h, levels = read_plotfile(out_dir)
    assert h["version"] == "HyperCLaw-V1.1" and h["dim"] == 3 and h["finest"] == 1
    assert h["names"] == ["density", "x_velocity", "y_velocity", "z_velocity"]
    assert h["ref"] == ["2"] and h["coord"] == 0 and h["steps"] == [3, 6]
    assert [tuple(map(int, dm)) for dm in h["domains"]] == [(0, 0, 0, 23, 15, 11), (0, 0, 0, 47, 31, 23)]
    assert h["dx"][0] == [1 / 24, 1 / 16, 1 / 12] and h["dx"][1] == [1 / 48, 1 / 32, 1 / 24]
    for lev, (step, phys, path) in enumerate(h["level_boxes"]):
        assert path == f"Level_{lev}/Cell" and len(phys) == levels[lev]["nboxes"]
        ext = [24 * 2 ** lev, 16 * 2 ** lev, 12 * 2 ** lev]
        for (lo, hi, _), p in zip(levels[lev]["fabs"], phys):
            for a in range(3):
                assert p[a] == [lo[a] / ext[a], (hi[a] + 1) / ext[a]]
    for lev in range(2):
        fabs = levels[lev]["fabs"]
        stats = [[float(v) for v in s.split(",")] for s in levels[lev]["stats"]]
        nb = len(fabs)
        assert len(stats) == 2 * nb
        for b, (_, _, data) in enumerate(fabs):
            assert np.isfinite(data).all()
            assert stats[b] == [float(data[c].min()) for c in range(4)]
            assert stats[nb + b] == [float(data[c].max()) for c in range(4)]
    samples = [line.split()[1:] for line in r.stdout.splitlines() if line.startswith("SAMPLE")]
    assert len(samples) == 6
    for s in samples:
        lev, i, j, k = map(int, s[:4])
        want = [float(v) for v in s[4:]]
        hit = [data[:, k - lo[2], j - lo[1], i - lo[0]] for lo, hi, data in levels[lev]["fabs"]
               if all(lo[a] <= p <= hi[a] for a, p in enumerate((i, j, k)))]
        assert len(hit) == 1, (lev, i, j, k)
        assert [float(v) for v in hit[0]] == want


def test_three_level_facade_matches_oracle(tmp_path):
    """A three-level run driven only through the AmrSim facade
    (SetStaticRefinement on levels 0 and 1, the reference's own fine masks,
    tau_s != tau_b, two level-0 steps) equals the oracle run on the boxes the
    facade made: density and velocity of every level, bitwise."""
    from oracle import oracle as orc

    exe = BIN / "three_level_dump"
    assert exe.exists()
    out_dir = tmp_path / "plt3"
    r = subprocess.run([str(exe), str(out_dir)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout
    assert "FINEST 2" in r.stdout
    h, levels = read_plotfile(out_dir)
    assert h["finest"] == 2 and h["steps"] == [2, 4, 8]
    NX, NY, NZ = 16, 12, 8
    rho = np.empty((NX, NY, NZ))
    u = np.empty((NX, NY, NZ, 3))
    for i in range(NX):
        for j in range(NY):
            for k in range(NZ):
                rho[i, j, k] = 1.0 + 0.01 * (i == 7) + 0.001 * ((j + 2 * k) % 5)
                u[i, j, k] = (0.002 * ((i + k) % 3), -0.001 * (j % 2), 0.0005 * ((i * j) % 4))
    boxes = [[(list(lo), list(hi)) for lo, hi, _ in lev["fabs"]] for lev in levels]
    o = orc.Sim(NX, NY, NZ, 0.7, 0.9, boxes=boxes[0])
    o.init(rho.reshape(-1), u.reshape(-1))
    for lev in (1, 2):
        o.add_fine(boxes[lev], boxes[lev - 1])  # MakeFineMask passes the parent's own BoxArray
    o.iterate(2)
    for lev in range(3):
        orho, ou = o.hydro(lev)
        bb = o.bbox(lev)
        for lo, hi, data in levels[lev]["fabs"]:
            sl = tuple(slice(lo[a] - bb[a], hi[a] - bb[a] + 1) for a in (2, 1, 0))
            want = np.concatenate([orho[sl][None], ou[(slice(None),) + sl]])
            assert (np.ascontiguousarray(want).view(np.uint64) == np.ascontiguousarray(data).view(np.uint64)).all(), \
                (lev, lo, hi)
    o.close()


def test_ml_pulse_reference_case():
    """The reference's "ml_pulse Regression" (tests/catch2RegressionTests.cpp:
    95-196) through the facade.  Its t = 0 assertions pass: both levels,
    level 1 sampled at 2 (i, j, k), equal the single-level goldens under
    Catch2's Approx, since a new level is the piecewise-constant copy of its
    parent.  Its t = 100 assertion on level 0 cannot pass and does not here:
    the Rohde cycle is not the single-level step (and the reference's own
    MakeFineMask zeroes the coarse low octant each step), so level 0 leaves
    the goldens -- the reference test fails the same way, as the oracle,
    which this build equals bitwise, shows."""
    rc, out = run("ml_pulse_dump")
    assert rc == 0, out[-2000:]
    fx = json.loads(GOLDEN.read_text())
    eps = 100 * 1.1920928955078125e-07
    pts = [line.split()[1:] for line in out.splitlines() if line.startswith("P ")]
    assert len(pts) == 3 * 10 * 10 * 50

    def approx(got, exp):
        return abs(got - exp) <= eps * abs(exp)

    worst100 = 0.0
    for t, level, k, r, u0, u1, u2 in pts:
        t, level, k = int(t), int(level), int(k)
        exp_r = float(fx["rho"][str(t)][k])
        exp_u = [float(fx["u"][str(t)][n][k]) for n in range(3)]
        got_u = [float(u0), float(u1), float(u2)]
        if t == 0:
            assert approx(float(r), exp_r), (level, k, r, exp_r)
            for n in range(3):
                assert approx(got_u[n], exp_u[n]), (level, k, n, got_u[n], exp_u[n])
        else:
            worst100 = max(worst100, abs(float(r) - exp_r) / abs(exp_r))
    print(f"[ml_pulse] t=0 both levels within Approx; t=100 level 0 max rel rho deviation {worst100:.3e}")
    assert worst100 > eps  # level 0 left the single-level goldens by t = 100

