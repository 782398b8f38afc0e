"""Refinement tags from the state on the GPU (lbx_tag_cells) equal the rule
restated with numpy on the density and velocity lbx_hydro returns, bit for
bit, in both storage precisions; clustered boxes drive a level remake."""
import numpy as np
import pytest

from lambrex_amd import lbx, workloads

pytestmark = pytest.mark.gpu


def criterion_numpy(rho, u, criterion):
    if criterion == lbx.LBX_TAG_DENSITY_GRADIENT:
        g = [(np.roll(rho, -1, axis=a) - np.roll(rho, 1, axis=a)) * 0.5 for a in (2, 1, 0)]  # x, y, z
        return g[0] * g[0] + g[1] * g[1] + g[2] * g[2]
    return u[0] * u[0] + u[1] * u[1] + u[2] * u[2]


@pytest.mark.parametrize("precision", [lbx.LBX_FP64, lbx.LBX_FP32_SHIFTED])
@pytest.mark.parametrize("criterion", [lbx.LBX_TAG_DENSITY_GRADIENT, lbx.LBX_TAG_SPEED])
def test_tags_match_numpy(ctx, precision, criterion):
    dims = (24, 16, 12)
    rho, u = workloads.random_state(dims, seed=71, rho_sigma=1e-2, u_amp=0.02)
    sim = lbx.Sim(ctx, dims, precision=precision)
    try:
        sim.define_level(0, workloads.chop(dims, 8))
        sim.set_tau(0, 0.7)
        sim.init_from_fields(rho, u)
        sim.iterate(3)
        r, v = sim.get_hydro(0)
        s = criterion_numpy(r, v, criterion)
        threshold = float(np.sqrt(np.median(s)))  # tags about half the cells
        tags = sim.tag_cells(criterion, threshold)
    finally:
        sim.close()
    want = (s > threshold * threshold).astype(np.uint8)
    assert 0 < want.sum() < want.size
    assert (tags == want).all()


def test_regrid_flow_from_tags(ctx):
    """Tags -> clustered fine boxes -> a new level, then a remake from new
    tags; the fine state stays finite and the level steps on."""
    dims = (32, 16, 16)
    rho, u = workloads.plane_pulse(dims, 0.01, axis=0, plane=10)
    sim = lbx.Sim(ctx, dims)
    try:
        sim.define_level(0, workloads.chop(dims, 16))
        sim.set_tau(0, 0.8)
        sim.init_from_fields(rho, u)
        boxes = lbx.cluster_tags(dims, sim.tag_cells(lbx.LBX_TAG_DENSITY_GRADIENT, 1e-3))
        # tags at x = 9 and 11, grown by 1, in tiles of 4: coarse x in [8, 16)
        assert boxes and all(lo[0] >= 16 and hi[0] <= 31 for lo, hi in boxes)
        sim.define_level(1, boxes)
        sim.set_tau(1, 1.1)
        sim.fill_from_coarse(1)
        sim.set_fine_mask(boxes)
        sim.iterate(20)  # the pulse spreads: more cells are tagged
        boxes2 = lbx.cluster_tags(dims, sim.tag_cells(lbx.LBX_TAG_DENSITY_GRADIENT, 1e-4))
        assert boxes2 != boxes
        sim.remake_level(1, boxes2)
        sim.set_fine_mask(boxes2)
        sim.iterate(5)
        f1, _ = sim.get_f_bbox(1)
        assert np.isfinite(f1).all()
    finally:
        sim.close()
