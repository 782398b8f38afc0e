"""Every kernel of the HIP library runs without scratch (private) memory.

A kernel whose per-lane state spills to scratch writes and re-reads it
through the memory hierarchy: the fill kernel of a patched parent once kept
an 80-byte copy of its patch descriptor there, about 27 MB of extra writes
per group of 8 patches.  This compiles each kernel source for gfx950 with the
library's own flags (device code only, to assembly, no GPU needed) and reads
the private segment size of every kernel."""
import re
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
CSRC = ROOT / "lambrex_amd" / "csrc"
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off", "-Xclang", "-target-feature",
         "-Xclang", "-packed-fp32-ops", "--cuda-device-only", "-S"]


def kernel_private_sizes(src, out):
    r = subprocess.run([HIPCC, *FLAGS, "-o", str(out), str(src)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    text = out.read_text()
    names = re.findall(r"^\s+\.name:\s+(_Z\S+)", text, re.M)
    sizes = [int(v) for v in re.findall(r"^\s+\.private_segment_fixed_size:\s+(\d+)", text, re.M)]
    assert names and len(names) == len(sizes)
    return dict(zip(names, sizes))


@pytest.mark.skipif(not Path(HIPCC).exists(), reason="hipcc not installed")
def test_no_kernel_uses_scratch(tmp_path):
    sources = sorted(CSRC.glob("*.hip"))
    assert {s.name for s in sources} >= {"lbx_lattice.hip", "lbx_twolevel.hip", "lbx_serve.hip"}
    with ThreadPoolExecutor(len(sources)) as pool:
        results = list(pool.map(lambda s: kernel_private_sizes(s, tmp_path / (s.stem + ".s")), sources))
    kernels = {k: v for r in results for k, v in r.items()}
    assert len(kernels) > 50
    spilled = {k: v for k, v in kernels.items() if v}
    assert not spilled, spilled
