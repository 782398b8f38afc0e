"""Pins of the oracle's two-level pieces (oracle/lbx_oracle.c restating
src/AmrSim.cpp:373-388, 471-485, 592-617) against what AMReX's documented
routines and the reference's own loops fix, each checked so that a plausible
misreading fails it.

No reference test pins the two-level path and AMReX is not in the image, so
these checks reduce each piece to a case whose answer is fixed without the
oracle's code:

* InitPostCollision of a fine level (FillPatchTwoLevels with PCInterp, then
  comp 0 zeroed on the outermost ghost layer, :471-485): every cell of the
  fine `now` and the coarse `now` carries a distinct code, so each grown fine
  cell of the result names exactly where it came from.  FillPatchTwoLevels
  takes fine valid data wherever some fine box holds the periodic image and
  the piecewise-constant parent (floor(pos / 2), periodic) elsewhere; ghosts
  of `now` hold NaN and must never be read.
* ZeroInvalidComponents (:604-617): brute-force enumeration of every
  (ghost cell, population) of a box grown by 2, against the reference's test
  vbox.contains(pos - 2 c_m), on a 2^3 box and uneven ones.
* sum_fine_to_coarse (:592-602): amrex_avgdown's average over each 2x2x2
  block of the fine FAB grown by 2 (children summed with i fastest, then
  times 1/8) ADDed by ParallelCopy into every coarse valid cell holding the
  periodic image of the coarsened box grown by 1.  A coarse cell with one
  contribution is checked bit for bit; where contributions from several
  blocks meet, AMReX's addition order is internal to ParallelCopy, so those
  cells are checked to 1e-14 only (the one part that stays unpinned).
* A whole cycle on a domain refined everywhere, one box per level: the fine
  ghosts are filled once, before the collision, and never refreshed (Q7), so
  only a fine cell whose two-substep domain of dependence stays inside the
  box (two or more layers in) must equal a periodic single-level run of the
  fine domain taking two steps, bit for bit; and every coarse cell away from
  the box faces (whose pulls would read the pre-collision ghost copies) must
  equal the single-level coarse step (unmasked) or +0 (masked) plus the 1/8
  sum of its 8 fine children after the step, bit for bit.

Every check runs at refinement ratio 2 (the reference's, src/AmrSim.cpp:766)
and at ratios 3 and 4 (the generalisation: parents floor(pos / r), r^3
blocks times 1/r^3 over the box grown by 2 / r coarse cells, r fine
substeps).  test_mutants_fail_a_pin builds the oracle with each planted
mistake (ORC_MUTANT, lbx_oracle.c) and shows that some check above fails on
it.
"""
import itertools
import subprocess

import numpy as np
import pytest

from oracle import oracle as orc

C = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
              [1, 1, 1], [1, 1, -1], [1, -1, 1], [1, -1, -1],
              [-1, 1, 1], [-1, 1, -1], [-1, -1, 1], [-1, -1, -1]])
G = orc.HALO


def c_order(rho_x, u_x):
    return (np.ascontiguousarray(rho_x.transpose(2, 1, 0)).reshape(-1),
            np.ascontiguousarray(u_x.transpose(3, 2, 1, 0)).reshape(-1))


def lists(boxes):
    return [(list(a), list(b)) for a, b in boxes]


def grown_positions(lo, hi):
    """Absolute (x, y, z) of every cell of a box grown by G, as arrays shaped
    like one component of its FAB [gz][gy][gx]."""
    zs = np.arange(lo[2] - G, hi[2] + G + 1)
    ys = np.arange(lo[1] - G, hi[1] + G + 1)
    xs = np.arange(lo[0] - G, hi[0] + G + 1)
    z, y, x = np.meshgrid(zs, ys, xs, indexing="ij")
    return x, y, z


def inside(x, y, z, lo, hi):
    return ((x >= lo[0]) & (x <= hi[0]) & (y >= lo[1]) & (y <= hi[1]) & (z >= lo[2]) & (z <= hi[2]))


# --- geometry of the piece checks -------------------------------------------
CDIMS = (8, 6, 6)  # coarse domain, fine domain r * CDIMS
RATIOS = (2, 3, 4)


def fine_boxes(r):
    """Fine boxes (fine index space, aligned to r) for ratio r: two boxes
    side by side in y spanning all y and z (ghosts from each other and from
    periodic images), an isolated box (ghosts from the coarse level) and a
    one-coarse-cell box."""
    def box(lo, hi):  # coarse cells lo..hi refined by r
        return (tuple(r * a for a in lo), tuple(r * b + r - 1 for b in hi))
    return [box((0, 0, 0), (2, 2, 5)), box((0, 3, 0), (2, 5, 5)), box((5, 1, 2), (5, 2, 3)),
            box((5, 4, 0), (5, 4, 0))]


FBOXES = fine_boxes(2)


def piece_sim(library, r=2):
    cboxes = orc.max_size_chop((0, 0, 0), tuple(n - 1 for n in CDIMS), 4)
    s = orc.Sim(*CDIMS, 0.8, boxes=cboxes, library=library)
    rho = np.ones(CDIMS[::-1])
    u = np.zeros((3,) + CDIMS[::-1])
    s.init(*c_order(rho, u))
    s.set_ratio(r)
    s.add_fine(lists(fine_boxes(r)), [])
    return s


def code_coarse(q, x, y, z):
    return -(1.0 + q * 10000.0 + (z * CDIMS[1] + y) * CDIMS[0] + x)


def code_fine(q, x, y, z, r=2):
    fd = [r * n for n in CDIMS]
    return 1.0 + q * 100000.0 + (z * fd[1] + y) * fd[0] + x


def set_coded(s, level, code):
    """Valid cells of every box of `level` carry code(q, x, y, z); ghosts NaN."""
    for b in range(s.nbox(level)):
        lo, hi = s.box(level, b)
        x, y, z = grown_positions(lo, hi)
        v = inside(x, y, z, lo, hi)
        data = np.full((15,) + x.shape, np.nan)
        for q in range(15):
            data[q][v] = code(q, x[v], y[v], z[v])
        s.set_grown(level, b, data, which=0)


def check_fill(library, r=2):
    """InitPostCollision(1) on coded fields; returns a list of failures."""
    fails = []
    s = piece_sim(library, r)
    fboxes = fine_boxes(r)
    set_coded(s, 0, code_coarse)
    set_coded(s, 1, lambda q, x, y, z: code_fine(q, x, y, z, r))
    s.stage(1, orc.STAGE_INIT_POST_COLLISION)
    fd = np.array([r * n for n in CDIMS])
    for b, (lo, hi) in enumerate(fboxes):
        got = s.get_grown(1, b, which=1)
        x, y, z = grown_positions(lo, hi)
        wx, wy, wz = x % fd[0], y % fd[1], z % fd[2]
        fine_image = np.zeros(x.shape, dtype=bool)
        for blo, bhi in fboxes:
            fine_image |= inside(wx, wy, wz, blo, bhi)
        # piecewise-constant parent: floor division, then the coarse periodic image
        px, py, pz = np.floor_divide(x, r) % CDIMS[0], np.floor_divide(y, r) % CDIMS[1], np.floor_divide(z, r) % CDIMS[2]
        outer = (x == lo[0] - 2) | (x == hi[0] + 2) | (y == lo[1] - 2) | (y == hi[1] + 2) | \
                (z == lo[2] - 2) | (z == hi[2] + 2)
        for q in range(15):
            exp = np.where(fine_image, code_fine(q, wx, wy, wz, r), code_coarse(q, px, py, pz))
            if q == 0:
                exp = np.where(outer, 0.0, exp)
            if np.isnan(got[q]).any():
                fails.append(f"fill: box {b} q {q} read a ghost of now (NaN)")
            elif not np.array_equal(got[q], exp):
                fails.append(f"fill: box {b} q {q}: {int((got[q] != exp).sum())} cells from the wrong source")
    s.close()
    return fails


def check_zero_invalid(library, r=2):
    """ZeroInvalidComponents(1) on all-ones FABs against the enumeration of
    vbox.contains(pos - 2 c_m) over every ghost cell and population."""
    fails = []
    s = piece_sim(library, r)
    fboxes = fine_boxes(r)
    for b, (lo, hi) in enumerate(fboxes):
        lo, hi = np.array(lo), np.array(hi)
        shape = s.get_grown(1, b).shape
        s.set_grown(1, b, np.ones(shape), which=1)
    s.stage(1, orc.STAGE_ZERO_INVALID)
    for b, (lo, hi) in enumerate(fboxes):
        got = s.get_grown(1, b, which=1)
        lo, hi = np.array(lo), np.array(hi)
        for (k, j, i) in itertools.product(range(got.shape[1]), range(got.shape[2]), range(got.shape[3])):
            pos = lo - G + np.array([i, j, k])
            valid = bool(np.all(pos >= lo) and np.all(pos <= hi))
            for m in range(15):
                src = pos - 2 * C[m]
                keep = valid or bool(np.all(src >= lo) and np.all(src <= hi))
                if got[m, k, j, i] != (1.0 if keep else 0.0):
                    fails.append(f"zero-invalid: box {b} pos {tuple(pos)} q {m} kept={got[m, k, j, i]}")
                    if len(fails) > 20:
                        s.close()
                        return fails
    s.close()
    return fails


def check_sum(library, r=2):
    """SumFromFine(0): random fine FABs (valid + ghosts) ADDed as 1/r^3 block
    sums into random coarse `next` values, over each box coarsened and grown
    by 2 // r coarse cells."""
    fails = []
    rng = np.random.default_rng(3)
    s = piece_sim(library, r)
    fboxes = fine_boxes(r)
    ng = 2 // r
    fine = []
    for b in range(len(fboxes)):
        d = rng.standard_normal(s.get_grown(1, b).shape)
        s.set_grown(1, b, d, which=0)
        fine.append(d)
    cnb = s.nbox(0)
    before = []
    for b in range(cnb):
        d = rng.standard_normal(s.get_grown(0, b).shape)
        s.set_grown(0, b, d, which=1)
        before.append(d)
    s.stage(0, orc.STAGE_SUM_FROM_FINE)
    after = [s.get_grown(0, b, which=1) for b in range(cnb)]
    # contributions per coarse cell (domain coordinates), in box order
    contrib = {}
    for b, (lo, hi) in enumerate(fboxes):
        lo, hi = np.array(lo), np.array(hi)
        clo, chi = np.floor_divide(lo, r) - ng, np.floor_divide(hi, r) + ng
        for cz in range(clo[2], chi[2] + 1):
            for cy in range(clo[1], chi[1] + 1):
                for cx in range(clo[0], chi[0] + 1):
                    vals = []
                    for dk in range(r):  # amrex_avgdown: k outer, i inner
                        for dj in range(r):
                            for di in range(r):
                                p = np.array([r * cx + di, r * cy + dj, r * cz + dk]) - lo + G
                                vals.append(fine[b][:, p[2], p[1], p[0]])
                    acc = np.zeros(15)
                    for v in vals:
                        acc = acc + v
                    key = (cx % CDIMS[0], cy % CDIMS[1], cz % CDIMS[2])
                    contrib.setdefault(key, []).append(acc * (1.0 / r ** 3))
    for b in range(cnb):
        lo, hi = s.box(0, b)
        x, y, z = grown_positions(lo, hi)
        v = inside(x, y, z, lo, hi)
        if not np.array_equal(after[b][:, ~v], before[b][:, ~v]):
            fails.append(f"sum: coarse box {b} ghosts changed")
        for (zz, yy, xx) in zip(*np.nonzero(v)):
            key = (int(x[zz, yy, xx]), int(y[zz, yy, xx]), int(z[zz, yy, xx]))
            got = after[b][:, zz, yy, xx]
            base = before[b][:, zz, yy, xx]
            parts = contrib.get(key, [])
            if len(parts) <= 1:
                exp = base + parts[0] if parts else base
                if not np.array_equal(got, exp):
                    fails.append(f"sum: cell {key} ({len(parts)} contribution) differs")
            else:
                exp = base + sum(parts)
                if not np.allclose(got, exp, rtol=1e-14, atol=1e-14):
                    fails.append(f"sum: cell {key} ({len(parts)} contributions) differs beyond 1e-14")
    s.close()
    return fails


def check_whole_cycle(library, masked, r=2):
    """One Rohde cycle with level 1 covering the whole domain (8^3 coarse in
    one box, (8r)^3 fine in one box), random state, unmasked or masked
    coarse."""
    fails = []
    dims = (8, 8, 8)
    nf = 8 * r
    fdims = (nf, nf, nf)
    rng = np.random.default_rng(11)
    rho = 1.0 + 1e-2 * rng.standard_normal(dims[::-1])
    u = 0.02 * (rng.random((3,) + dims[::-1]) - 0.5)
    tau = 0.8
    tau_f = r * (tau - 0.5) + 0.5
    one = [((0, 0, 0), (7, 7, 7))]
    fbox = [((0, 0, 0), (nf - 1, nf - 1, nf - 1))]
    mask = list(fbox) if masked else []
    s = orc.Sim(*dims, tau, boxes=lists(one), library=library)
    s.init(*c_order(rho, u))
    s.set_ratio(r)
    s.add_fine(lists(fbox), lists(mask))
    f_fine0 = s.get_f(1)
    s.iterate(1)
    cg, fg = s.get_f(0), s.get_f(1)
    if s.tau(1) != tau_f or s.step(1) != r:
        fails.append(f"cycle: fine tau {s.tau(1)} / steps {s.step(1)} are not ComputeDt's / r")
    s.close()
    # the fine level alone: a periodic single level on the fine domain, r
    # steps; the ghosts, filled once and never refreshed, reach r layers in
    ref_f = orc.Sim(*fdims, tau_f, boxes=lists(fbox), library=library)
    ref_f.init(*c_order(np.ones(fdims[::-1]), np.zeros((3,) + fdims[::-1])))
    ref_f.set_f(f_fine0)
    ref_f.iterate(r)
    deep = (slice(None), slice(r, nf - r), slice(r, nf - r), slice(r, nf - r))
    if not np.array_equal(ref_f.get_f(0)[deep], fg[deep]):
        fails.append(f"cycle: fine cells {r} layers inside the box differ from a single-level run")
    ref_f.close()
    # the coarse level alone, one single-level step (same arithmetic per cell)
    ref_c = orc.Sim(*dims, tau, boxes=lists(one), library=library)
    ref_c.init(*c_order(rho, u))
    ref_c.iterate(1)
    a = ref_c.get_f(0)
    ref_c.close()
    inner = (slice(None), slice(1, 7), slice(1, 7), slice(1, 7))
    acc = np.zeros((15, 8, 8, 8))
    for dk in range(r):
        for dj in range(r):
            for di in range(r):
                acc = acc + fg[:, dk::r, dj::r, di::r]
    avg = acc * (1.0 / r ** 3)
    exp = (0.0 + avg) if masked else (a + avg)
    if not np.array_equal(cg[inner], exp[inner]):
        fails.append(f"cycle: interior coarse cells differ from the closed form ({'masked' if masked else 'unmasked'})")
    return fails


def all_checks(library, ratios=(2, 3)):
    fails = []
    for r in ratios:
        fails += (check_fill(library, r) + check_zero_invalid(library, r) + check_sum(library, r) +
                  check_whole_cycle(library, False, r) + check_whole_cycle(library, True, r))
    return fails


@pytest.mark.parametrize("r", RATIOS)
def test_fill_two_levels_pinned(r):
    assert check_fill(orc.lib(), r) == []


@pytest.mark.parametrize("r", RATIOS)
def test_zero_invalid_components_pinned(r):
    assert check_zero_invalid(orc.lib(), r) == []


@pytest.mark.parametrize("r", RATIOS)
def test_sum_fine_to_coarse_pinned(r):
    assert check_sum(orc.lib(), r) == []


@pytest.mark.parametrize("r", RATIOS)
@pytest.mark.parametrize("masked", [False, True])
def test_whole_cycle_on_fully_refined_domain(masked, r):
    assert check_whole_cycle(orc.lib(), masked, r) == []


MUTANTS = {
    1: "sum_fine_to_coarse without the 1/8",
    2: "sum_fine_to_coarse over valid children only",
    3: "sum_fine_to_coarse overwrites instead of ADDing",
    4: "children summed in the wrong order",
    5: "fine ghosts from the coarse level although a fine image exists",
    6: "parent index by truncation instead of floor",
    7: "comp 0 zeroed on both ghost layers",
    8: "every component zeroed on the outer layer",
    9: "ZeroInvalidComponents keeps pos + 2c",
    10: "ZeroInvalidComponents keeps pos - c",
    11: "average-down over the valid boxes only (nGrow 0) at every ratio",
    12: "average-down volume fraction 1/8 at every ratio",
    13: "the finest level takes 2 substeps at every ratio",
}


@pytest.fixture(scope="module")
def mutant_dir(tmp_path_factory):
    return tmp_path_factory.mktemp("oracle_mutants")


@pytest.mark.parametrize("mutant", sorted(MUTANTS))
def test_mutants_fail_a_pin(mutant, mutant_dir):
    """Each planted misreading of the two-level pieces fails at least one of
    the checks above (so the checks pin what they claim to pin)."""
    src = orc._HERE / "lbx_oracle.c"
    out = mutant_dir / f"liblbx_oracle_m{mutant}.so"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-std=c11", f"-DORC_MUTANT={mutant}",
                    "-shared", "-o", str(out), str(src)], check=True)
    fails = all_checks(orc.load(out))
    assert fails, f"mutant {mutant} ({MUTANTS[mutant]}) passes every pin"
