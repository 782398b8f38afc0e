"""Clustering of refinement tags into fine boxes (lbx_cluster_tags, host
only): the boxes cover exactly the tiles that hold a tag grown by the
buffer, do not overlap, are aligned to the refinement ratio and respect
max_grid_size.  No reference test pins this (the reference only tags a
static box and leaves clustering to AMReX); the checks restate the rule
with numpy."""
import itertools

import numpy as np
import pytest

from lambrex_amd import lbx


def tiles_expected(tags, blocking, buffer):
    """Coarse-cell mask of the tiles holding a tag grown by `buffer`
    (periodic), the rule restated with numpy."""
    grown = np.zeros_like(tags, dtype=bool)
    for dz, dy, dx in itertools.product(range(-buffer, buffer + 1), repeat=3):
        grown |= np.roll(tags.astype(bool), (dz, dy, dx), axis=(0, 1, 2))
    nz, ny, nx = tags.shape
    mask = np.zeros_like(grown)
    for k in range(0, nz, blocking):
        for j in range(0, ny, blocking):
            for i in range(0, nx, blocking):
                if grown[k:k + blocking, j:j + blocking, i:i + blocking].any():
                    mask[k:k + blocking, j:j + blocking, i:i + blocking] = True
    return mask


def coverage(boxes, shape):
    """Coarse-cell cover count of fine boxes."""
    nz, ny, nx = shape
    cov = np.zeros(shape, dtype=np.int32)
    for lo, hi in boxes:
        for d in range(3):
            assert lo[d] % 2 == 0 and hi[d] % 2 == 1, (lo, hi)
        cov[lo[2] // 2:hi[2] // 2 + 1, lo[1] // 2:hi[1] // 2 + 1, lo[0] // 2:hi[0] // 2 + 1] += 1
    return cov


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("dims,blocking,buffer,mgs", [((24, 16, 12), 4, 1, 32), ((30, 18, 10), 4, 2, 16),
                                                       ((16, 16, 16), 8, 0, 64)])
def test_boxes_cover_exactly_the_tagged_tiles(seed, dims, blocking, buffer, mgs):
    rng = np.random.default_rng(seed)
    nx, ny, nz = dims
    tags = (rng.random((nz, ny, nx)) < 0.01).astype(np.uint8)
    boxes = lbx.cluster_tags(dims, tags, blocking=blocking, buffer=buffer, max_grid_size=mgs)
    cov = coverage(boxes, (nz, ny, nx))
    assert cov.max() <= 1  # no overlap
    assert ((cov == 1) == tiles_expected(tags, blocking, buffer)).all()
    for lo, hi in boxes:
        for d in range(3):
            assert 0 <= lo[d] <= hi[d] < 2 * dims[d]
            assert hi[d] - lo[d] + 1 <= mgs


def test_no_tags_no_boxes_and_all_tags_the_whole_domain():
    dims = (16, 12, 8)
    assert lbx.cluster_tags(dims, np.zeros((8, 12, 16), np.uint8)) == []
    boxes = lbx.cluster_tags(dims, np.ones((8, 12, 16), np.uint8), max_grid_size=64)
    assert boxes == [((0, 0, 0), (31, 23, 15))]


def test_a_tag_across_the_periodic_face_grows_onto_the_other_side():
    dims = (16, 8, 8)
    tags = np.zeros((8, 8, 16), np.uint8)
    tags[0, 0, 0] = 1
    boxes = lbx.cluster_tags(dims, tags, blocking=4, buffer=1)
    cov = coverage(boxes, (8, 8, 16))
    assert cov[0, 0, 0] and cov[7, 7, 15] and cov[0, 0, 15] and not cov[3, 3, 8]


def test_bad_arguments():
    lib = lbx.load()
    import ctypes
    d = np.array([8, 8, 8], np.int32)
    t = np.zeros(512, np.uint8)
    n = ctypes.c_int32()
    dp = d.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    tp = t.ctypes.data_as(ctypes.c_void_p)
    assert lib.lbx_cluster_tags(dp, tp, 0, 1, 32, 0, None, ctypes.byref(n)) == lbx.LBX_E_ARG   # blocking
    assert lib.lbx_cluster_tags(dp, tp, 4, 1, 31, 0, None, ctypes.byref(n)) == lbx.LBX_E_ARG   # odd max size
    t[0] = 1
    assert lib.lbx_cluster_tags(dp, tp, 4, 1, 32, 0, None, ctypes.byref(n)) == lbx.LBX_OK
    assert n.value >= 1
    out = np.zeros(6, np.int32)
    op = out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    if n.value > 1:
        assert lib.lbx_cluster_tags(dp, tp, 4, 1, 32, 1, op, ctypes.byref(n)) == lbx.LBX_E_ARG


@pytest.mark.parametrize("seed", range(4))
def test_allowed_mask_keeps_tiles_with_a_disallowed_cell_unrefined(seed):
    """lbx_cluster_tags_allowed: the refined tiles are exactly those holding a
    grown tag and no disallowed cell."""
    rng = np.random.default_rng(100 + seed)
    dims = (24, 20, 12)
    nx, ny, nz = dims
    tags = (rng.random((nz, ny, nx)) < 0.02).astype(np.uint8)
    allowed = np.ones((nz, ny, nx), np.uint8)
    allowed[:, :, :5] = 0                                   # a slab
    allowed[rng.random((nz, ny, nx)) < 0.003] = 0           # scattered cells
    blocking, buffer = 4, 1
    boxes = lbx.cluster_tags(dims, tags, blocking=blocking, buffer=buffer, max_grid_size=32, allowed=allowed)
    want = tiles_expected(tags, blocking, buffer)
    for k in range(0, nz, blocking):
        for j in range(0, ny, blocking):
            for i in range(0, nx, blocking):
                if not allowed[k:k + blocking, j:j + blocking, i:i + blocking].all():
                    want[k:k + blocking, j:j + blocking, i:i + blocking] = False
    cov = coverage(boxes, (nz, ny, nx))
    assert cov.max() <= 1
    assert ((cov == 1) == want).all()
    assert not cov[:, :, :5].any()
    # with everything allowed it is lbx_cluster_tags
    assert lbx.cluster_tags(dims, tags, allowed=np.ones_like(tags)) == lbx.cluster_tags(dims, tags)
