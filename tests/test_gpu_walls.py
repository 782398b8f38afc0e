"""Halfway bounce-back walls on non-periodic directions (the reference aborts
there, src/AmrSim.cpp:788-797).  No reference result exists, so the path is
pinned by a numpy restatement of the bounce-back stream around the oracle's
per-cell collision (bitwise), exact mass conservation in a closed box, and
the closed-form decay of the slowest shear mode between two walls."""
import itertools

import numpy as np
import pytest

from lambrex_amd import lbx, workloads
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

C = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
              [1, 1, 1], [1, 1, -1], [1, -1, 1], [1, -1, -1],
              [-1, 1, 1], [-1, 1, -1], [-1, -1, 1], [-1, -1, -1]])
OPP = [int(np.flatnonzero((C == -C[q]).all(axis=1))[0]) for q in range(15)]


def bb_stream(post, periodic):
    """Pull stream with halfway bounce-back: f_q(x) = post_q(x - c_q), or
    post_opp(q)(x) where x - c_q leaves the domain through a wall."""
    _, nz, ny, nx = post.shape
    out = np.empty_like(post)
    dims = (nx, ny, nz)
    for q in range(15):
        for z, y, x in itertools.product(range(nz), range(ny), range(nx)):
            up = [x - C[q][0], y - C[q][1], z - C[q][2]]
            wall = any(not periodic[d] and not 0 <= up[d] < dims[d] for d in range(3))
            if wall:
                out[q, z, y, x] = post[OPP[q], z, y, x]
            else:
                ux, uy, uz = (up[d] % dims[d] for d in range(3))
                out[q, z, y, x] = post[q, uz, uy, ux]
    return out


def collide_all(f, omega):
    out = np.empty_like(f)
    _, nz, ny, nx = f.shape
    for z, y, x in itertools.product(range(nz), range(ny), range(nx)):
        out[:, z, y, x] = orc.collide_cell(f[:, z, y, x], omega, omega)
    return out


@pytest.mark.parametrize("periodic", [(1, 0, 1), (0, 0, 0), (1, 1, 0), (0, 1, 1)])
def test_bounce_back_matches_numpy_restatement(ctx, periodic):
    dims = (9, 7, 6)
    tau = 0.7
    omega = 1.0 / (tau + 0.5)
    rho, u = workloads.random_state(dims, seed=83, rho_sigma=1e-2, u_amp=0.03)
    sim = lbx.Sim(ctx, dims, periodic=periodic)
    try:
        sim.define_level(0, workloads.chop(dims, 4))  # uneven boxes
        sim.set_tau(0, tau)
        sim.init_from_fields(rho, u)
        f = sim.get_f(0)
        sim.iterate(4)
        g = sim.get_f(0)
    finally:
        sim.close()
    for _ in range(4):
        f = bb_stream(collide_all(f, omega), periodic)
    assert (f.view(np.uint64) == g.view(np.uint64)).all()


def test_closed_box_conserves_mass(ctx):
    dims = (24, 20, 16)
    rho, u = workloads.random_state(dims, seed=89, rho_sigma=1e-2, u_amp=0.03)
    sim = lbx.Sim(ctx, dims, periodic=(0, 0, 0))
    try:
        sim.define_level(0, workloads.chop(dims, 8))
        sim.set_tau(0, 0.6)
        sim.init_from_fields(rho, u)
        m0 = sim.conservation()[0]
        sim.iterate(300)
        m1 = sim.conservation()[0]
        r, v = sim.get_hydro()
    finally:
        sim.close()
    assert abs(m1 - m0) <= 1e-12 * m0
    assert np.isfinite(r).all() and np.isfinite(v).all()


def test_shear_mode_between_walls_decays_at_the_viscous_rate(ctx):
    """u_x(y) = A sin(pi (y + 1/2) / H) between halfway bounce-back walls at
    y = -1/2 and y = H - 1/2 decays as exp(-nu (pi/H)^2 t), nu = tau/3."""
    nx, ny, nz = 4, 48, 4
    tau = 0.8
    amp = 1e-4
    y = np.arange(ny)
    ux = amp * np.sin(np.pi * (y + 0.5) / ny)
    rho = np.ones((nz, ny, nx))
    u = np.zeros((3, nz, ny, nx))
    u[0] = ux[None, :, None]
    sim = lbx.Sim(ctx, (nx, ny, nz), periodic=(1, 0, 1))
    try:
        sim.define_level(0, workloads.chop((nx, ny, nz), 64))
        sim.set_tau(0, tau)
        sim.init_from_fields(rho, u)
        sim.iterate(200)  # let the start-up transient pass
        a1 = sim.get_hydro()[1][0, 2, :, 2]
        sim.iterate(800)
        a2 = sim.get_hydro()[1][0, 2, :, 2]
    finally:
        sim.close()
    mode = np.sin(np.pi * (y + 0.5) / ny)
    p1, p2 = a1 @ mode, a2 @ mode
    nu = tau / 3.0
    expected = np.exp(-nu * (np.pi / ny) ** 2 * 800)
    print(f"[walls] decay ratio {p2 / p1:.6f}, closed form {expected:.6f}, rel {(p2 / p1 - expected) / expected:.2e}")
    assert abs(p2 / p1 - expected) <= 2e-3 * expected


@pytest.mark.parametrize("periodic", [(1, 0, 1), (0, 0, 0)])
def test_fp32_shifted_walls_track_fp64(ctx, periodic):
    """The FP32-shifted stream-collide with walls (its own kernel instances)
    against the FP64 one from the same state, 200 steps through the fused
    kernel: density within 1e-6 relative, velocity within 1e-3 of the largest
    FP64 speed, and the FP64 mass sum of the FP32 run conserved to 1e-9 (the
    walls reflect populations, they do not remove them)."""
    dims = (40, 36, 32)
    rho, u = workloads.random_state(dims, seed=97, rho_sigma=1e-3, u_amp=0.01)
    out = {}
    for prec in (lbx.LBX_FP64, lbx.LBX_FP32_SHIFTED):
        sim = lbx.Sim(ctx, dims, precision=prec, rho0=1.0, periodic=periodic)
        try:
            sim.define_level(0, workloads.chop(dims, 16))
            sim.set_tau(0, 0.6)
            sim.init_from_fields(rho, u)
            m0 = sim.conservation()[0]
            sim.iterate(200)
            m1 = sim.conservation()[0]
            out[prec] = (sim.get_hydro(), m0, m1)
        finally:
            sim.close()
    (r64, u64), _, _ = out[lbx.LBX_FP64]
    (r32, u32), m0, m1 = out[lbx.LBX_FP32_SHIFTED]
    e_rho = float(np.max(np.abs(r32 - r64) / r64))
    e_u = float(np.max(np.abs(u32 - u64)) / np.max(np.abs(u64)))
    print(f"[walls fp32] periodic {periodic}: max rel rho err {e_rho:.2e}, max |du|/max|u| {e_u:.2e}, "
          f"mass drift {abs(m1 - m0) / m0:.2e}")
    assert e_rho < 1e-6
    assert e_u < 1e-3
    assert abs(m1 - m0) <= 1e-9 * m0
