"""bench.py's command line: more than one GPU is refused with a message
before any GPU call (this build measures one GPU)."""
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=300, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_hierarchy_workloads_are_nested_and_counted():
    """The two- and three-level bench workloads: lattice updates per level-0
    step match the boxes (level l steps 2^l times), boxes do not overlap, and
    every level-2 cell of each box grown by 2 has its parent in level 1
    (periodic), as lbx_level_define requires."""
    import numpy as np
    sys.path.insert(0, str(ROOT))
    import bench

    def cells(boxes):
        return sum(int(np.prod([hi[d] - lo[d] + 1 for d in range(3)])) for lo, hi in boxes)

    dims, cboxes, fboxes, mask, rho, u, lu2 = bench.two_level_geometry()
    assert lu2 == cells(cboxes) + 2 * cells(fboxes)
    dims, levels, masks, rho, u, lu3 = bench.three_level_geometry()
    assert lu3 == sum(2 ** l * cells(b) for l, b in enumerate(levels))
    assert masks == [levels[1], levels[2]]
    n1 = [2 * d for d in dims]
    cover = np.zeros(n1[::-1], dtype=np.int8)
    for lo, hi in levels[1]:
        cover[lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] += 1
    assert cover.max() == 1 and cover.sum() == cells(levels[1])
    for lo, hi in levels[2]:
        zs = np.arange(lo[2] - 2, hi[2] + 3) // 2 % n1[2]
        ys = np.arange(lo[1] - 2, hi[1] + 3) // 2 % n1[1]
        xs = np.arange(lo[0] - 2, hi[0] + 3) // 2 % n1[0]
        assert cover[np.ix_(zs, ys, xs)].all(), (lo, hi)


def test_hierarchy_traffic_entries_match_the_workloads():
    """The measured per-step traffic that bench.py reports beside the two- and
    three-level lines is on file, and its algorithmic bytes are 240 B per
    lattice update of the same workloads."""
    sys.path.insert(0, str(ROOT))
    import bench
    lu2 = bench.two_level_geometry()[-1]
    lu3 = bench.three_level_geometry()[-1]
    for key, workload, lu in (("two_level_step", "128^3 + 64x256x256", lu2),
                              ("three_level_step", "128^3 + 64x256x256 + 32x512x512", lu3)):
        t = bench.step_traffic(key, workload)
        assert t is not None, key
        assert t["algorithmic_bytes_per_step"] == 240 * lu
        assert 1.0 < t["over_algorithmic"] < 5.0
        assert abs(t["bytes_per_step"] / t["algorithmic_bytes_per_step"] - t["over_algorithmic"]) < 1e-3
        assert (ROOT / t["source"]).is_file()
