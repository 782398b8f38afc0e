"""Pins of the CPU oracle (oracle/lbx_oracle.c) that need no GPU.

What pins what:
  - tables: exact rational inverse and identities (independent Fraction
    Gauss-Jordan here vs the oracle's own C derivation);
  - the whole single-level path: the reference's pulse goldens
    (tests/pulseRegression.h) under Catch2 Approx, noise-level velocities
    included -- plus a check that a restatement with another summation
    order fails them, so the goldens do pin operation order;
  - collision physics: rest state fixed point, exact-rational evaluation of
    the mode-space relaxation written as matrix algebra, conservation, and the
    shear-wave decay rate nu = tau/3 (reference tau, omega = 1/(tau+1/2));
  - streaming: a permutation equal to numpy.roll per population;
  - equilibrium: zeroth and first moments reproduce rho and rho u;
  - init layout: C order (i slowest, component fastest);
  - decomposition invariance: one box vs many boxes, bitwise.
"""
import json
from fractions import Fraction
from pathlib import Path

import numpy as np
import pytest

from lambrex_amd import workloads
from oracle import oracle as orc

GOLDEN = Path(__file__).with_name("golden") / "pulse_profile.json"
EPS = 100 * float(np.finfo(np.float32).eps)

C = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
              [1, 1, 1], [1, 1, -1], [1, -1, 1], [1, -1, -1],
              [-1, 1, 1], [-1, 1, -1], [-1, -1, 1], [-1, -1, -1]])
W = [Fraction(2, 9)] + [Fraction(1, 9)] * 6 + [Fraction(1, 72)] * 8


def exact_mode_matrix():
    """Mode basis from its definition (density, momentum, second moments
    c_a c_b - delta_ab/3 in xx,xy,xz,yy,yz,zz order, ghosts)."""
    M = [[Fraction(0)] * 15 for _ in range(15)]
    pairs = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    for p in range(15):
        c = C[p]
        c2 = int((c * c).sum())
        g = 1 if c2 == 1 else -2
        M[0][p] = Fraction(1)
        for a in range(3):
            M[1 + a][p] = Fraction(int(c[a]))
        for r, (a, b) in enumerate(pairs):
            M[4 + r][p] = Fraction(int(c[a] * c[b])) - (Fraction(1, 3) if a == b else 0)
        M[10][p] = Fraction(g)
        for a in range(3):
            M[11 + a][p] = Fraction(int(c[a]) * g)
        M[14][p] = Fraction(int(c[0] * c[1] * c[2]))
    return M


def exact_inverse(A):
    n = len(A)
    aug = [list(A[i]) + [Fraction(int(i == j)) for j in range(n)] for i in range(n)]
    for col in range(n):
        piv = next(r for r in range(col, n) if aug[r][col] != 0)
        aug[col], aug[piv] = aug[piv], aug[col]
        inv = aug[col][col]
        aug[col] = [v / inv for v in aug[col]]
        for r in range(n):
            if r != col and aug[r][col] != 0:
                fac = aug[r][col]
                aug[r] = [a - fac * b for a, b in zip(aug[r], aug[col])]
    return [row[n:] for row in aug]


def test_tables_are_the_exact_rationals():
    M, MI, Ct = orc.tables()
    assert (Ct == C).all()
    Me = exact_mode_matrix()
    MIe = exact_inverse(Me)
    for m in range(15):
        for p in range(15):
            assert M[m, p] == float(Me[m][p]), (m, p)
            assert MI[m, p] == float(MIe[m][p]), (m, p)
    # identities over Q
    for i in range(15):
        for j in range(15):
            assert sum(Me[i][k] * MIe[k][j] for k in range(15)) == int(i == j)
    for m in range(15):
        assert sum(W[i] * Me[m][i] for i in range(15)) == int(m == 0)
    for p in range(15):
        assert MIe[p][0] == W[p]
    # a few entries the reference writes as literals (src/AmrSim.cpp:1058-1072)
    assert MI[0, 10] == -2.0 / 9.0 and MI[7, 5] == 1.0 / 8.0 and MI[1, 10] == 1.0 / 18.0
    # the rounded tables do not invert each other exactly (the golden noise)
    assert np.abs(M @ MI - np.eye(15)).max() > 0


def pulse_sim():
    dims, rho_c, u_c = workloads.reference_pulse_c_order()
    s = orc.Sim(*dims, 0.5)
    assert [(tuple(lo), tuple(hi)) for lo, hi in s.boxes] == \
        [((0, 0, 24), (9, 9, 49)), ((0, 0, 0), (9, 9, 23))]
    s.init(rho_c, u_c)
    return s


def approx_mismatches(rho, u, fx, t):
    exp_rho = np.array([float(v) for v in fx["rho"][str(t)]])[:, None, None]
    exp_u = np.array([[float(v) for v in fx["u"][str(t)][n]] for n in range(3)])[:, :, None, None]
    bad_rho = ~(np.abs(rho - exp_rho) <= EPS * np.abs(exp_rho))
    bad_u = ~(np.abs(u - exp_u) <= EPS * np.abs(exp_u))
    return int(bad_rho.sum()), int(bad_u.sum())


def test_pulse_goldens():
    """tests/catch2RegressionTests.cpp:6-93 with the oracle."""
    fx = json.loads(GOLDEN.read_text())
    s = pulse_sim()
    done = 0
    for t in (0, 100, 200):
        s.iterate(t - done)
        done = t
        rho, u = s.hydro()
        assert approx_mismatches(rho, u, fx, t) == (0, 0), t
        assert s.time() == float(t) and s.step() == t
    s.close()


def numpy_restatement_step(f, omega):
    """The same model with another summation order (matrix products), used
    only to show that the goldens discriminate operation order."""
    M, MI, _ = orc.tables()
    nz, ny, nx = f.shape[1:]
    g = f.reshape(15, -1)
    mode = M @ g
    rho = mode[0]
    v = mode[1:4] / rho
    usq = (v * v).sum(0)
    S = np.empty((3, 3, g.shape[1]))
    idx = {(0, 0): 4, (0, 1): 5, (0, 2): 6, (1, 1): 7, (1, 2): 8, (2, 2): 9}
    for (a, b), m in idx.items():
        S[a, b] = S[b, a] = mode[m]
    tr = S[0, 0] + S[1, 1] + S[2, 2]
    for a in range(3):
        S[a, a] -= tr / 3
    tr = tr - omega * (tr - rho * usq)
    for (a, b), m in idx.items():
        d = usq / 15 if a == b else 0.0
        mode[m] = S[a, b] - omega * (S[a, b] - rho * (v[a] * v[b] - d)) + (tr / 3 if a == b else 0.0)
    mode[10:] = 0
    post = (MI @ mode).reshape(15, nz, ny, nx)
    return orc.stream_periodic(post)


def test_goldens_discriminate_operation_order():
    """A correct model evaluated with a different summation order (BLAS
    matrix products) drifts from the golden noise-level velocities, so
    passing the goldens is evidence of the reference's exact order."""
    fx = json.loads(GOLDEN.read_text())
    s = pulse_sim()
    f = s.get_f()
    s.close()
    omega = 1.0 / (0.5 + 0.5)
    for _ in range(100):
        f = numpy_restatement_step(f, omega)
    M, _, _ = orc.tables()
    g = f.reshape(15, -1)
    rho = (M[0] @ g).reshape(f.shape[1:])
    u = np.stack([(M[1 + a] @ g).reshape(f.shape[1:]) / rho for a in range(3)])
    bad_rho, bad_u = approx_mismatches(rho, u, fx, 100)
    assert bad_rho == 0  # the physics is right...
    assert bad_u > 0     # ...but the round-off pattern is not the reference's


def test_rest_state_is_a_fixed_point():
    M, MI, _ = orc.tables()
    for rho in (1.0, 0.37, 2.5):
        f = np.array([float(w) for w in W]) * rho
        g = orc.collide_cell(f, 1.0 / 1.3, 1.0 / 0.9)
        assert np.abs(g - f).max() <= 4 * np.finfo(float).eps * rho


def exact_collide(f, omega_s, omega_b):
    """Mode-space relaxation in exact rational arithmetic, as matrix algebra:
    modes = M f; the trace of the second-moment block relaxes to rho u^2 at
    omega_b, its traceless part to rho (u u - u^2 I/15) at omega_s; ghosts
    are zeroed; f' = M^-1 modes."""
    Me = exact_mode_matrix()
    MIe = exact_inverse(Me)
    fr = [Fraction(x) for x in f]
    ws, wb = Fraction(omega_s), Fraction(omega_b)
    mode = [sum(Me[m][p] * fr[p] for p in range(15)) for m in range(15)]
    rho = mode[0]
    v = [mode[1 + a] / rho for a in range(3)]
    usq = sum(x * x for x in v)
    S = [[None] * 3 for _ in range(3)]
    idx = {(0, 0): 4, (0, 1): 5, (0, 2): 6, (1, 1): 7, (1, 2): 8, (2, 2): 9}
    for (a, b), m in idx.items():
        S[a][b] = S[b][a] = mode[m]
    tr = S[0][0] + S[1][1] + S[2][2]
    tr_new = tr - wb * (tr - rho * usq)
    new = list(mode)
    for (a, b), m in idx.items():
        delta = 1 if a == b else 0
        traceless = S[a][b] - delta * tr / 3
        target = rho * (v[a] * v[b] - usq * delta * Fraction(1, 15))
        new[m] = traceless - ws * (traceless - target) + delta * tr_new / 3
    for m in range(10, 15):
        new[m] = Fraction(0)
    return [float(sum(MIe[p][m] * new[m] for m in range(15))) for p in range(15)]


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_collide_matches_exact_rational_relaxation(seed):
    rng = np.random.default_rng(seed)
    w = np.array([float(x) for x in W])
    f = w * (1.0 + 0.05 * rng.standard_normal(15))
    omega_s, omega_b = 1.0 / (rng.uniform(0.05, 1.5) + 0.5), 1.0 / (rng.uniform(0.05, 1.5) + 0.5)
    got = orc.collide_cell(f, omega_s, omega_b)
    exp = np.array(exact_collide(f, omega_s, omega_b))
    assert np.abs(got - exp).max() <= 1e-14 * np.abs(exp).max()


def test_collision_conserves_mass_and_momentum():
    rng = np.random.default_rng(7)
    w = np.array([float(x) for x in W])
    for _ in range(20):
        f = w * (1.0 + 0.1 * rng.standard_normal(15))
        g = orc.collide_cell(f, 1.0 / 0.9, 1.0 / 1.7)
        assert abs(g.sum() - f.sum()) <= 1e-15 * 8
        assert np.abs(C.T @ g - C.T @ f).max() <= 1e-15 * 8


def test_stream_is_numpy_roll():
    rng = np.random.default_rng(3)
    f = rng.random((15, 5, 6, 7))
    out = orc.stream_periodic(f)
    for q in range(15):
        exp = np.roll(f[q], shift=(C[q, 2], C[q, 1], C[q, 0]), axis=(0, 1, 2))
        assert (out[q] == exp).all()


def test_equilibrium_moments():
    rng = np.random.default_rng(5)
    for _ in range(10):
        rho = rng.uniform(0.5, 2.0)
        u = rng.uniform(-0.1, 0.1, 3)
        f = orc.equilibrium_cell(rho, u)
        assert abs(f.sum() - rho) <= 1e-15 * 8
        assert np.abs(C.T @ f - rho * u).max() <= 1e-15 * 8
        r2, u2 = orc.hydro_cell(f)
        assert abs(r2 - rho) <= 1e-14 and np.abs(u2 - u).max() <= 1e-14


def test_element_wise_init_layout():
    """tests/catch2InitTests.cpp:50-109: values 0,1,2,... in C order land at
    (i,j,k) with i slowest, and velocity component fastest."""
    NX, NY, NZ = 16, 9, 8
    s = orc.Sim(NX, NY, NZ, 0.01)
    s.init(np.arange(NX * NY * NZ, dtype=float), np.zeros(3 * NX * NY * NZ))
    rho, _ = s.hydro()
    s.close()
    for i, j, k in [(0, 0, 1), (1, 0, 0), (3, 4, 5), (15, 8, 7)]:
        assert rho[k, j, i] == pytest.approx(i * NY * NZ + j * NZ + k, rel=EPS, abs=0)
    s = orc.Sim(NX, NY, NZ, 0.01)
    u_c = np.arange(3 * NX * NY * NZ, dtype=float) * 1e-6
    s.init(np.ones(NX * NY * NZ), u_c)
    _, u = s.hydro()
    s.close()
    for i, j, k, n in [(0, 0, 0, 2), (1, 2, 3, 1), (15, 8, 7, 0)]:
        assert u[n, k, j, i] == pytest.approx(((i * NY + j) * NZ + k) * 3 * 1e-6 + n * 1e-6, rel=1e-9)


@pytest.mark.parametrize("tau", [0.3, 0.8])
def test_shear_wave_decays_at_nu_tau_over_3(tau):
    """u_x(y) = U sin(2 pi y / N) decays as exp(-nu k^2 t) with nu = tau/3
    (reference tau; omega = 1/(tau+1/2)).  The rate is taken from the ratio
    of two amplitudes, after the start-up transient of an equilibrium
    initialisation; the O(k^2) lattice correction at N = 64 is ~3e-4."""
    N, U, half = 64, 1e-6, 300
    dims = (2, N, 2)
    rho = np.ones((dims[2], dims[1], dims[0]))
    u = np.zeros((3,) + rho.shape)
    y = np.arange(N)
    u[0] = U * np.sin(2 * np.pi * y / N)[None, :, None]
    s = orc.Sim(*dims, tau)
    s.init(np.ascontiguousarray(rho.transpose(2, 1, 0)).reshape(-1),
           np.ascontiguousarray(u.transpose(3, 2, 1, 0)).reshape(-1))
    amps = []
    for _ in range(2):
        s.iterate(half)
        _, uu = s.hydro()
        amps.append(2.0 / N * (uu[0, 0, :, 0] * np.sin(2 * np.pi * y / N)).sum())
    s.close()
    k = 2 * np.pi / N
    nu_measured = -np.log(amps[1] / amps[0]) / (k * k * half)
    assert nu_measured == pytest.approx(tau / 3, rel=1e-3)


def test_decomposition_invariance():
    dims = (12, 10, 14)
    rho, u = workloads.random_state(dims, seed=1, rho_sigma=1e-2, u_amp=0.03)
    rho_c = np.ascontiguousarray(rho.transpose(2, 1, 0)).reshape(-1)
    u_c = np.ascontiguousarray(u.transpose(3, 2, 1, 0)).reshape(-1)
    out = []
    for mgs in (64, 5, 3):
        s = orc.Sim(*dims, 0.7, max_grid_size=mgs)
        s.init(rho_c, u_c)
        s.iterate(6)
        out.append(s.get_f())
        s.close()
    assert (out[0].view(np.uint64) == out[1].view(np.uint64)).all()
    assert (out[0].view(np.uint64) == out[2].view(np.uint64)).all()


def test_max_size_chop_matches_amrex_cuts():
    assert workloads.chop((10, 10, 50), 32) == [((0, 0, 24), (9, 9, 49)), ((0, 0, 0), (9, 9, 23))]
    assert len(workloads.chop((256, 256, 256), 64)) == 64
    assert len(workloads.chop((512, 512, 512), 64)) == 512
    # 100 = 4 * 25 with chunk 32 = 4 * 8: 4 blocks of 6-7 (x4) -> 28, 24, 24, 24 from the top
    cuts = sorted(b[0][0] for b in workloads.chop((100, 1, 1), 32))
    assert cuts == [0, 24, 48, 72]
    assert orc.max_size_chop((0, 0, 0), (99, 0, 0), 32) == [
        ([72, 0, 0], [99, 0, 0]), ([48, 0, 0], [71, 0, 0]), ([24, 0, 0], [47, 0, 0]), ([0, 0, 0], [23, 0, 0])]
