"""Synthetic initial states for the benchmark and test configurations.

Every random draw is a pure function of (seed, global cell index, stream):
SplitMix64 of a counter, so any box decomposition, any process and any run
sees the same inputs.  Normals use Box-Muller.  Fields are produced x
fastest: rho[nz][ny][nx], u[3][nz][ny][nx].
"""
import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_NSTREAM = 5  # normal pair (2) + three velocity components


def splitmix64(counter, seed):
    """SplitMix64 output for an array of uint64 counters under `seed`."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) * _GOLDEN + (counter + np.uint64(1)) * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform53(counter, seed):
    """Uniform doubles in [0,1) with 53 random bits."""
    return (splitmix64(counter, seed) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def c_order_index(dims, lo=(0, 0, 0), shape=None):
    """Reference C-order cell index i*ny*nz + j*nz + k (include/AmrSim.h:79-83)
    of the x-fastest block starting at lo with shape (nz', ny', nx')."""
    nx, ny, nz = dims
    if shape is None:
        shape = (nz, ny, nx)
    k = np.arange(lo[2], lo[2] + shape[0], dtype=np.uint64)[:, None, None]
    j = np.arange(lo[1], lo[1] + shape[1], dtype=np.uint64)[None, :, None]
    i = np.arange(lo[0], lo[0] + shape[2], dtype=np.uint64)[None, None, :]
    return i * np.uint64(ny * nz) + j * np.uint64(nz) + k


def random_state(dims, seed, rho_sigma, u_amp, lo=(0, 0, 0), shape=None):
    """rho = 1 + rho_sigma * N(0,1), u_a ~ U(-u_amp, u_amp) per component,
    for the x-fastest block at lo (whole domain by default)."""
    c = c_order_index(dims, lo, shape) * np.uint64(_NSTREAM)
    u1 = uniform53(c, seed)
    u2 = uniform53(c + np.uint64(1), seed)
    normal = np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)
    rho = 1.0 + rho_sigma * normal
    u = np.stack([u_amp * (2.0 * uniform53(c + np.uint64(2 + a), seed) - 1.0) for a in range(3)])
    return rho, u


def plane_pulse(dims, amplitude, axis=0, plane=None, lo=(0, 0, 0), shape=None):
    """rho = 1 everywhere except 1 + amplitude on one plane normal to `axis`
    (default: the middle plane), u = 0."""
    nx, ny, nz = dims
    if shape is None:
        shape = (nz, ny, nx)
    if plane is None:
        plane = dims[axis] // 2
    xx = np.arange(lo[0], lo[0] + shape[2])[None, None, :]
    yy = np.arange(lo[1], lo[1] + shape[1])[None, :, None]
    zz = np.arange(lo[2], lo[2] + shape[0])[:, None, None]
    coord = (xx, yy, zz)[axis]
    rho = np.broadcast_to(np.where(coord == plane, 1.0 + amplitude, 1.0), shape).copy()
    return rho, np.zeros((3,) + tuple(shape))


def reference_pulse_c_order():
    """The reference's pulse test input (tests/catch2RegressionTests.cpp:8-34):
    10x10x50, rho = 1 plus 0.01 on plane k = NZ/2 - 1, normalised by the mean
    of one z column, u = 0; returned in C order as the reference builds it."""
    NX, NY, NZ = 10, 10, 50
    NUMEL = NX * NY * NZ
    rho = [1.0] * NUMEL
    k = NZ // 2
    for i in range(NX):
        for j in range(NY):
            rho[i * NY * NZ + j * NZ + k - 1] += 0.01
    z_mean = 0.0
    for k in range(NZ):
        z_mean += rho[(NUMEL + NY * NZ) // 2 + k]
    z_mean /= NZ
    rho = np.array([r / z_mean for r in rho])
    return (NX, NY, NZ), rho, np.zeros(3 * NUMEL)


def chop(dims, max_grid_size):
    """Level-0 boxes as AMReX's BoxList::maxSize cuts the domain."""
    boxes = [([0, 0, 0], [dims[0] - 1, dims[1] - 1, dims[2] - 1])]
    for d in range(3):
        out = []
        for blo, bhi in boxes:
            blo, bhi = list(blo), list(bhi)
            length = bhi[d] - blo[d] + 1
            if length > max_grid_size:
                ratio, bs, nlen = 1, max_grid_size, length
                while bs % 2 == 0 and nlen % 2 == 0:
                    ratio, bs, nlen = ratio * 2, bs // 2, nlen // 2
                numblk = nlen // bs + (1 if nlen % bs else 0)
                sz, extra = nlen // numblk, nlen % numblk
                for kk in range(numblk - 1):
                    ksize = (sz + 1 if kk < extra else sz) * ratio
                    pos = bhi[d] - ksize + 1
                    top = list(blo)
                    top[d] = pos
                    out.append((top, list(bhi)))
                    bhi[d] = pos - 1
            out.append((blo, bhi))
        boxes = out
    return [(tuple(lo), tuple(hi)) for lo, hi in boxes]
