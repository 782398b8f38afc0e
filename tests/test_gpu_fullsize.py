"""Bitwise parity at the benchmark's own full sizes: the 256^3
single-level workload (64 boxes of 64^3) for three steps, so that the fused
stream-collide kernel runs at its benchmark launch shape (Iterate(3) from an
initialised state is one collide-only pass and two fused passes, and the
readers pull the last stream), the two-level workload (coarse 128^3 + fine
64x256x256) with both the fine-BoxArray mask and the reference's own mask, and
the three-level one (+ 32x512x512), GPU against the oracle on the same seeded
inputs.  The oracle takes a few seconds per step at these sizes on the GPU
host; the smaller parity tests cover many steps."""
import sys
from pathlib import Path

import numpy as np
import pytest

from lambrex_amd import lbx, workloads
from oracle import oracle as orc

ROOT = Path(__file__).resolve().parents[1]
pytestmark = pytest.mark.gpu


def c_order(rho_x, u_x):
    return (np.ascontiguousarray(rho_x.transpose(2, 1, 0)).reshape(-1),
            np.ascontiguousarray(u_x.transpose(3, 2, 1, 0)).reshape(-1))


def lists(boxes):
    return [(list(a), list(b)) for a, b in boxes]


def same_bits(a, b):
    return bool((np.ascontiguousarray(a).view(np.uint64) == np.ascontiguousarray(b).view(np.uint64)).all())


def run_256(ctx, steps, precision=lbx.LBX_FP64):
    dims = (256, 256, 256)
    boxes = workloads.chop(dims, 64)
    rho, u = workloads.random_state(dims, seed=2024, rho_sigma=1e-3, u_amp=0.01)
    sim = lbx.Sim(ctx, dims, precision=precision, rho0=1.0)
    try:
        sim.define_level(0, boxes)
        sim.set_tau(0, 0.6)
        sim.init_from_fields(rho, u)
        sim.iterate(steps)
        return dims, boxes, rho, u, sim.get_f(0), sim.get_hydro(0)
    finally:
        sim.close()


def test_256_cubed_three_steps_bitwise(ctx):
    """The headline workload through the fused kernel at its benchmark launch
    shape (512-lane workgroups over whole x-rows, one resident per CU):
    every population after 3 steps equals the oracle's, bit for bit."""
    dims, boxes, rho, u, fg, _ = run_256(ctx, 3)
    o = orc.Sim(*dims, 0.6, boxes=lists(boxes))
    o.init(*c_order(rho, u))
    del rho, u
    o.iterate(3)
    fo = o.get_f(0)
    o.close()
    assert same_bits(fo, fg)


def test_256_cubed_fp32_shifted_three_steps_vs_fp64(ctx):
    """FP32-shifted storage on the same 256^3 state and launch path: density
    within 1e-6 relative of FP64 after 3 steps (the precision target)."""
    _, _, _, _, _, (rho64, u64) = run_256(ctx, 3)
    _, _, _, _, _, (rho32, u32) = run_256(ctx, 3, precision=lbx.LBX_FP32_SHIFTED)
    rel = float(np.max(np.abs(rho32 - rho64) / rho64))
    du = float(np.max(np.abs(u32 - u64)) / np.max(np.abs(u64)))
    print(f"[fp32 256^3 x3] max rel rho err {rel:.3e}, max |du|/max|u| {du:.3e}", flush=True)
    assert rel <= 1e-6
    assert du <= 1e-4


@pytest.mark.parametrize("mask_kind", ["fine", "reference"])
def test_two_level_workload_two_coarse_steps_bitwise(ctx, mask_kind):
    """mask_kind "fine": the benchmark's mask (coarse cells under the fine
    level); "reference": MakeFineMask's own, the coarse BoxArray passed as the
    fine one (src/AmrSim.cpp:419-428), which masks coarsen(coarse boxes, 2).
    Two coarse steps in one call: the fine level's 16 patch groups run in
    storage order in the first and in reverse order in the second, each
    step from its own captured graph."""
    sys.path.insert(0, str(ROOT))
    import bench
    dims, cboxes, fboxes, mask, rho, u, _ = bench.two_level_geometry()
    if mask_kind == "reference":
        mask = list(cboxes)
    sim = lbx.Sim(ctx, dims)
    try:
        sim.define_level(0, cboxes)
        sim.set_tau(0, 0.8)
        sim.init_from_fields(rho, u)
        sim.define_level(1, fboxes)
        sim.set_tau(1, 2 * (0.8 - 0.5) + 0.5)
        sim.fill_from_coarse(1)
        sim.set_fine_mask(mask)
        sim.iterate(2)
        fg0 = sim.get_f(0)
        fg1, _ = sim.get_f_bbox(1)
    finally:
        sim.close()
    o = orc.Sim(*dims, 0.8, boxes=lists(cboxes))
    o.init(*c_order(rho, u))
    o.add_fine(lists(fboxes), lists(mask))
    o.iterate(2)
    fo0, fo1 = o.get_f(0), o.get_f(1)
    o.close()
    assert same_bits(fo0, fg0)
    assert same_bits(fo1, fg1)


def test_three_level_workload_one_step_bitwise(ctx):
    sys.path.insert(0, str(ROOT))
    import bench
    dims, levels, masks, rho, u, _ = bench.three_level_geometry()
    sim = lbx.Sim(ctx, dims)
    tau = [0.8]
    for _ in (1, 2):
        tau.append(2 * (tau[-1] - 0.5) + 0.5)  # ComputeDt, src/AmrSim.cpp:313-314 (1.7000000000000002 on level 2)
    try:
        sim.define_level(0, levels[0])
        sim.set_tau(0, tau[0])
        sim.init_from_fields(rho, u)
        for l in (1, 2):
            sim.define_level(l, levels[l])
            sim.set_tau(l, tau[l])
            sim.fill_from_coarse(l)
            sim.set_level_mask(l - 1, masks[l - 1])
        sim.iterate(1)
        fg = [sim.get_f(0)] + [sim.get_f_bbox(l)[0] for l in (1, 2)]
    finally:
        sim.close()
    o = orc.Sim(*dims, tau[0], boxes=lists(levels[0]))
    o.init(*c_order(rho, u))
    for l in (1, 2):
        o.add_fine(lists(levels[l]), lists(masks[l - 1]))
        assert o.tau(l) == tau[l]
    o.iterate(1)
    for l in range(3):
        assert same_bits(o.get_f(l), fg[l]), f"level {l}"
    o.close()


def test_serving_largest_request_size_bitwise(ctx):
    """A 128^3 FP64 serving run (the largest size of the serving trace, 8,192
    work items per pass), batched beside two small runs, equals the oracle."""
    dims, steps, tau = (128, 128, 128), 3, 0.65
    rho, u = workloads.random_state(dims, seed=47, rho_sigma=1e-3, u_amp=0.01)
    srv = lbx.Server(ctx, precision=lbx.LBX_FP64, arena_bytes=1 << 30)
    _, o = srv.submit(dims, steps, tau, rho=rho, u=u, want_f=True)
    srv.submit((32, 32, 32), 7, 0.9, amplitude=1e-3)
    srv.submit((20, 12, 16), 2, 1.2, amplitude=1e-2)
    srv.drain()
    srv.poll()
    srv.close()
    s = orc.Sim(*dims, tau, boxes=lists(workloads.chop(dims, 64)))
    s.init(*c_order(rho, u))
    s.iterate(steps)
    assert same_bits(o["f"], s.get_f())
    s.close()


def walls_step(o, f, periodic):
    """One collide + halfway bounce-back stream step of the whole lattice.
    The oracle's periodic step gives g_q(x) = post_q(x - c_q) (wrapped), so
    g already holds every population that does not cross a wall; one that
    would enter x through a wall takes post_opp(q)(x) = g_opp(x - c_q)
    (wrapped) instead -- index shifts only, so the result stays exact."""
    _, _, C = orc.tables()
    opp = [int(np.flatnonzero((C == -C[q]).all(axis=1))[0]) for q in range(15)]
    o.set_f(f)
    o.iterate(1)
    g = o.get_f(0)
    out = g.copy()
    n = g.shape[1:][::-1]  # nx, ny, nz
    for q in range(15):
        cross = np.zeros(g.shape[1:], dtype=bool)
        for d in range(3):  # axis d of the [z][y][x] array is 2 - d
            if periodic[d] or C[q][d] == 0:
                continue
            face = [slice(None)] * 3
            face[2 - d] = 0 if C[q][d] > 0 else n[d] - 1
            cross[tuple(face)] = True
        bounced = np.roll(g[opp[q]], shift=(int(C[q][2]), int(C[q][1]), int(C[q][0])), axis=(0, 1, 2))
        out[q][cross] = bounced[cross]
    return out


@pytest.mark.parametrize("dims,mgs,periodic,steps", [((256, 256, 256), 64, (1, 1, 0), 2),
                                                     ((96, 80, 72), 32, (0, 0, 0), 3)])
def test_walls_bitwise(ctx, dims, mgs, periodic, steps):
    """Bounce-back walls at the walls measurement's own size (256^3, walls on
    both z faces) and a closed ragged box, through the fused kernel (Iterate
    with more than one step), against the oracle's collision and the
    bounce-back stream above, bitwise."""
    boxes = workloads.chop(dims, mgs)
    rho, u = workloads.random_state(dims, seed=71, rho_sigma=1e-3, u_amp=0.01)
    sim = lbx.Sim(ctx, dims, periodic=periodic)
    try:
        sim.define_level(0, boxes)
        sim.set_tau(0, 0.6)
        sim.init_from_fields(rho, u)
        f = sim.get_f(0)
        sim.iterate(steps)
        fg = sim.get_f(0)
    finally:
        sim.close()
    del rho, u
    o = orc.Sim(*dims, 0.6, boxes=lists(boxes))
    for _ in range(steps):
        f = walls_step(o, f, periodic)
    o.close()
    assert same_bits(f, fg)
