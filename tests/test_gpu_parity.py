"""GPU parity of the FP64 HIP path against the CPU oracle (bit-exact).

Each case builds the same initial state on both sides (inputs through the
same C-order vectors the reference's SetInitialDensity/Velocity take), runs N
steps and compares every population of every cell bitwise.  The oracle keeps
the reference's structure (ghosted FABs, FillPatch / collide / FillBoundary /
stream, src/AmrSim.cpp:124-135, 109-122); the GPU runs one fused pull +
collide pass per step, so bitwise agreement pins the re-association
S (C S)^(n-1) C as well as the arithmetic.
"""
import json
from pathlib import Path

import numpy as np
import pytest

from lambrex_amd import lbx, workloads
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).with_name("golden") / "pulse_profile.json"
APPROX_EPS = 100 * float(np.finfo(np.float32).eps)  # Catch2 Approx default


def omega(tau):
    return 1.0 / (tau + 0.5)


def c_order_fields(rho_x, u_x):
    """x-fastest rho[nz][ny][nx], u[3][..] -> reference C order vectors."""
    rho_c = np.ascontiguousarray(rho_x.transpose(2, 1, 0)).reshape(-1)
    u_c = np.ascontiguousarray(u_x.transpose(3, 2, 1, 0)).reshape(-1)
    return rho_c, u_c


def run_gpu(ctx, dims, boxes, tau, rho_x, u_x, steps, precision=lbx.LBX_FP64, rho0=1.0):
    sim = lbx.Sim(ctx, dims, precision=precision, rho0=rho0)
    try:
        sim.define_level(0, boxes)
        sim.set_tau(0, tau)
        sim.init_from_fields(rho_x, u_x)
        if steps:
            sim.iterate(steps)
        return sim.get_f(), sim.get_hydro(), sim.conservation()
    finally:
        sim.close()


def run_oracle(dims, boxes, tau, rho_x, u_x, steps):
    s = orc.Sim(*dims, tau, boxes=[(list(lo), list(hi)) for lo, hi in boxes])
    rho_c, u_c = c_order_fields(rho_x, u_x)
    s.init(rho_c, u_c)
    if steps:
        s.iterate(steps)
    out = s.get_f(), s.hydro()
    s.close()
    return out


def assert_bitwise(a, b, what):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    assert a.shape == b.shape, what
    same = a.view(np.uint64) == b.view(np.uint64)
    if not same.all():
        idx = np.argwhere(~same)[0]
        pytest.fail(f"{what}: {int((~same).sum())} values differ; first at {tuple(idx)}: "
                    f"gpu {a[tuple(idx)]!r} oracle {b[tuple(idx)]!r}")


def test_c1_pulse_32cubed_100_steps_bitwise(ctx):
    """32^3, one box, tau 0.8, rho 1.01 on plane x=16, 100 steps."""
    dims = (32, 32, 32)
    boxes = workloads.chop(dims, 32)
    assert len(boxes) == 1
    rho, u = workloads.plane_pulse(dims, 0.01, axis=0, plane=16)
    f_gpu, (rho_gpu, u_gpu), _ = run_gpu(ctx, dims, boxes, 0.8, rho, u, 100)
    f_orc, (rho_orc, u_orc) = run_oracle(dims, boxes, 0.8, rho, u, 100)
    assert_bitwise(f_gpu, f_orc, "f after 100 steps")
    assert_bitwise(rho_gpu, rho_orc, "rho")
    assert_bitwise(u_gpu, u_orc, "u")
    # the pulse has moved: not a trivial fixed point
    assert np.abs(rho_gpu - 1.0).max() > 1e-4


@pytest.mark.parametrize("steps", [0, 1, 2, 7])
def test_random_uneven_boxes_bitwise(ctx, steps):
    """Random state on a non-cubic domain cut into uneven boxes: exercises
    periodic wrap on every face and cross-box pulls."""
    dims = (24, 20, 18)
    boxes = workloads.chop(dims, 8)
    rho, u = workloads.random_state(dims, seed=11, rho_sigma=1e-2, u_amp=0.05)
    f_gpu, (rho_gpu, u_gpu), _ = run_gpu(ctx, dims, boxes, 0.6, rho, u, steps)
    f_orc, (rho_orc, u_orc) = run_oracle(dims, boxes, 0.6, rho, u, steps)
    assert_bitwise(f_gpu, f_orc, f"f after {steps} steps")
    assert_bitwise(rho_gpu, rho_orc, "rho")
    assert_bitwise(u_gpu, u_orc, "u")


def test_tau_s_differs_from_tau_b(ctx):
    dims = (16, 12, 10)
    boxes = workloads.chop(dims, 8)
    rho, u = workloads.random_state(dims, seed=5, rho_sigma=1e-2, u_amp=0.05)
    sim = lbx.Sim(ctx, dims)
    sim.define_level(0, boxes)
    sim.set_tau(0, 0.7, 1.3)
    sim.init_from_fields(rho, u)
    sim.iterate(5)
    f_gpu = sim.get_f()
    sim.close()
    s = orc.Sim(*dims, 0.7, 1.3, boxes=[(list(a), list(b)) for a, b in boxes])
    s.init(*c_order_fields(rho, u))
    s.iterate(5)
    assert_bitwise(f_gpu, s.get_f(), "f with tau_s != tau_b")


def test_reference_pulse_goldens_on_gpu(ctx):
    """tests/catch2RegressionTests.cpp:6-93 on the GPU: 10x10x50, tau 0.5,
    two boxes (max_grid_size 32), rho and u at t = 0, 100, 200 against the
    reference's golden arrays under Catch2 Approx rules."""
    dims, rho_c, u_c = workloads.reference_pulse_c_order()
    rho_x, u_x = lbx.c_order_to_xfast(rho_c, u_c, dims)
    fx = json.loads(GOLDEN.read_text())
    boxes = workloads.chop(dims, 32)
    assert len(boxes) == 2
    sim = lbx.Sim(ctx, dims)
    sim.define_level(0, boxes)
    sim.set_tau(0, 0.5)
    sim.init_from_fields(rho_x, u_x)
    done = 0
    for t in (0, 100, 200):
        sim.iterate(t - done)
        done = t
        rho, u = sim.get_hydro()
        exp_rho = np.array([float(v) for v in fx["rho"][str(t)]])[:, None, None]
        exp_u = np.array([[float(v) for v in fx["u"][str(t)][n]] for n in range(3)])[:, :, None, None]
        ok_rho = np.abs(rho - exp_rho) <= APPROX_EPS * np.abs(exp_rho)
        ok_u = np.abs(u - exp_u) <= APPROX_EPS * np.abs(exp_u)
        assert ok_rho.all(), f"t={t}: {int((~ok_rho).sum())} density values outside Approx"
        assert ok_u.all(), f"t={t}: {int((~ok_u).sum())} velocity values outside Approx"
    sim.close()


def test_decomposition_invariance_bitwise(ctx):
    """One box vs many uneven boxes: identical populations after 9 steps."""
    dims = (40, 24, 32)
    rho, u = workloads.random_state(dims, seed=3, rho_sigma=1e-3, u_amp=0.01)
    f1, _, _ = run_gpu(ctx, dims, workloads.chop(dims, 64), 0.6, rho, u, 9)
    f2, _, _ = run_gpu(ctx, dims, workloads.chop(dims, 7), 0.6, rho, u, 9)
    assert_bitwise(f1, f2, "1 box vs many boxes")


def test_iterate_in_chunks_equals_one_call(ctx):
    """Iterate(3)+Iterate(4) == Iterate(7): the collide prologue / stream
    epilogue per call must not change results."""
    dims = (16, 16, 16)
    boxes = workloads.chop(dims, 8)
    rho, u = workloads.random_state(dims, seed=9, rho_sigma=1e-2, u_amp=0.02)
    sim = lbx.Sim(ctx, dims)
    sim.define_level(0, boxes)
    sim.set_tau(0, 0.9)
    sim.init_from_fields(rho, u)
    sim.iterate(3)
    sim.iterate(4)
    fa = sim.get_f()
    sim.close()
    fb, _, _ = run_gpu(ctx, dims, boxes, 0.9, rho, u, 7)
    assert_bitwise(fa, fb, "3+4 vs 7 steps")


def test_upload_download_roundtrip(ctx):
    dims = (12, 10, 8)
    boxes = workloads.chop(dims, 5)
    rng = np.random.default_rng(0)
    f = rng.random((15,) + dims[::-1])
    sim = lbx.Sim(ctx, dims)
    sim.define_level(0, boxes)
    sim.set_f(f)
    assert_bitwise(sim.get_f(), f, "upload/download")
    sim.close()


def test_tiny_odd_domain_one_step(ctx):
    """Odd extents smaller than a wavefront row, boxes of 1-4 cells: one step
    equals the oracle's collide + stream."""
    dims = (9, 7, 5)
    boxes = workloads.chop(dims, 4)
    rho, u = workloads.random_state(dims, seed=21, rho_sigma=5e-2, u_amp=0.05)
    f_gpu, _, _ = run_gpu(ctx, dims, boxes, 1.1, rho, u, 1)
    f_orc, _ = run_oracle(dims, boxes, 1.1, rho, u, 1)
    assert_bitwise(f_gpu, f_orc, "one step on a tiny odd domain")


def test_conservation_256_cubed_64_boxes(ctx):
    """Config: 256^3 as 64 boxes of 64^3, tau 0.6, rho = 1 + 1e-3 N(0,1),
    u ~ U(-0.01, 0.01).  Mass and momentum are conserved to round-off."""
    dims = (256, 256, 256)
    boxes = workloads.chop(dims, 64)
    assert len(boxes) == 64
    rho, u = workloads.random_state(dims, seed=2024, rho_sigma=1e-3, u_amp=0.01)
    sim = lbx.Sim(ctx, dims)
    sim.define_level(0, boxes)
    sim.set_tau(0, 0.6)
    sim.init_from_fields(rho, u)
    c0 = sim.conservation()
    sim.iterate(50)
    c1 = sim.conservation()
    sim.close()
    assert abs(c1[0] - c0[0]) <= 1e-12 * abs(c0[0])
    # momentum: absolute drift relative to the mass scale
    assert np.all(np.abs(c1[1:] - c0[1:]) <= 1e-12 * abs(c0[0]))


def test_fp32_shifted_close_to_fp64(ctx):
    """FP32 storage of f - w rho0 with FP32 collision stays within 1e-6
    relative density of FP64 on the same initial state."""
    dims = (32, 32, 32)
    boxes = workloads.chop(dims, 16)
    rho, u = workloads.random_state(dims, seed=4, rho_sigma=1e-4, u_amp=0.005)
    _, (r64, u64), c64 = run_gpu(ctx, dims, boxes, 0.55, rho, u, 200)
    _, (r32, u32), c32 = run_gpu(ctx, dims, boxes, 0.55, rho, u, 200, precision=lbx.LBX_FP32_SHIFTED)
    rel = np.abs(r32 - r64) / np.abs(r64)
    assert rel.max() <= 1e-6, rel.max()
    # velocity error relative to the largest FP64 speed
    assert np.abs(u32 - u64).max() <= 1e-4 * np.abs(u64).max()


def test_error_paths(ctx):
    lib = ctx._lib
    walled = lbx.Sim(ctx, (8, 8, 8), periodic=(1, 0, 1))  # bounce-back walls in y: single level only
    walled.define_level(0, workloads.chop((8, 8, 8), 8))
    with pytest.raises(lbx.LbxError) as e:
        walled.define_level(1, [((0, 0, 0), (3, 3, 3))])
    assert e.value.code == lbx.LBX_E_UNSUPPORTED
    walled.close()
    sim = lbx.Sim(ctx, (8, 8, 8))
    with pytest.raises(lbx.LbxError) as e:
        sim.define_level(0, [((0, 0, 0), (7, 7, 6))])  # does not tile
    assert e.value.code == lbx.LBX_E_ARG
    with pytest.raises(lbx.LbxError) as e:
        sim.iterate(1)  # nothing defined
    assert e.value.code == lbx.LBX_E_STATE
    sim.define_level(0, [((0, 0, 0), (7, 7, 7))])
    with pytest.raises(lbx.LbxError) as e:
        sim.iterate(1)  # relaxation not set
    assert e.value.code == lbx.LBX_E_STATE
    rho = np.zeros(512)
    u = np.zeros(3 * 512)
    rc = lib.lbx_hydro(sim.h, 0, 3, rho.ctypes.data_as(lbx._dp), u.ctypes.data_as(lbx._dp))
    assert rc == lbx.LBX_E_ARG  # box 3 does not exist
    assert b"box 3" in lib.lbx_last_error(ctx.h)
    assert lib.lbx_iterate(sim.h, -1) == lbx.LBX_E_ARG
    assert lib.lbx_set_relax(sim.h, 0, float("nan"), 1.0) == lbx.LBX_E_ARG
    sim.close()
    for dims in ((1 << 30, 1, 1), (1 << 20, 1 << 20, 1 << 20), (0, 4, 4)):  # no overflow, no allocation
        with pytest.raises(lbx.LbxError) as e:
            lbx.Sim(ctx, dims)
        assert e.value.code == lbx.LBX_E_ARG


def test_autotuned_shapes_do_not_change_results(ctx):
    """With autotuning on, the fused kernel runs with the fastest measured
    workgroup shape; populations stay bit-identical to the default shape."""
    dims = (48, 40, 24)
    boxes = workloads.chop(dims, 16)
    rho, u = workloads.random_state(dims, seed=77, rho_sigma=1e-2, u_amp=0.02)
    (dbx, dby, docc), dms = ctx.tuned_shape(48, 40)  # never tuned: the default shape
    assert (dbx, dby, dms) == (48, 10, 0.0) and docc in (0, 1)  # FP64: 512 lanes, one per CU
    f_default, _, _ = run_gpu(ctx, dims, boxes, 0.7, rho, u, 6)
    ctx.set_autotune(True)
    try:
        f_tuned, _, _ = run_gpu(ctx, dims, boxes, 0.7, rho, u, 6)
        (bx, by, occ), ms = ctx.tuned_shape(48, 40)
    finally:
        ctx.set_autotune(False)
    assert_bitwise(f_tuned, f_default, "autotuned vs default shape")
    assert bx * by <= 512 and bx in (48, 32) and occ in (0, 1, 2, 3, 4, 6) and ms > 0


def test_lazy_state_between_calls_bitwise(ctx):
    """Between Iterate calls the device keeps the post-collision state and
    readers pull the last stream: reads after every call, a box upload while
    the state is post-collision (which first streams it in place) and the
    walls-free pull of the conservation sums must all agree with the oracle,
    which keeps the reference's pre-collision state."""
    dims = (20, 18, 16)
    boxes = workloads.chop(dims, 8)
    rho, u = workloads.random_state(dims, seed=31, rho_sigma=1e-2, u_amp=0.03)
    sim = lbx.Sim(ctx, dims)
    sim.define_level(0, boxes)
    sim.set_tau(0, 0.75)
    sim.init_from_fields(rho, u)
    o = orc.Sim(*dims, 0.75, boxes=[(list(a), list(b)) for a, b in boxes])
    o.init(*c_order_fields(rho, u))
    rng = np.random.default_rng(5)
    for chunk in (2, 1, 3):
        sim.iterate(chunk)
        o.iterate(chunk)
        fo = o.get_f()
        ro, uo = o.hydro()
        rg, ug = sim.get_hydro()
        assert_bitwise(sim.get_f(), fo, f"f after a chunk of {chunk}")
        assert_bitwise(rg, ro, "rho")
        assert_bitwise(ug, uo, "u")
        # the sums of the pulled populations equal those of the oracle's f
        mass = sim.conservation()[0]
        assert abs(mass - ro.sum()) <= 1e-12 * abs(mass)
        # overwrite box 1 while the device state is post-collision
        (lo, hi) = boxes[1]
        sub = fo[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1]
        new = sub * (1.0 + 1e-3 * rng.standard_normal(sub.shape))
        sim.upload_f(0, 1, new)
        fo[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] = new
        o.set_f(fo)
    sim.iterate(0)
    sim.iterate(2)
    o.iterate(2)
    assert_bitwise(sim.get_f(), o.get_f(), "f after the uploads and 2 more steps")
    sim.close()
    o.close()
