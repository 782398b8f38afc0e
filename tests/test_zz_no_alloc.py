"""lbx_conservation_noalloc, through the C ABI: the sums equal
lbx_conservation's bit for bit, on two calls in a row, the second over a
smaller input (a 64^3 level array, then a 24x20x16 level array and its
patched level 1).  The output buffer is device memory the test allocates
through the HIP runtime liblbx.so itself uses (torch bundles a HIP runtime
of its own, which cannot share the process's device); the entry point
allocates nothing."""
import ctypes

import numpy as np
import pytest

from lambrex_amd import lbx, workloads

pytestmark = pytest.mark.gpu


def same_bits(a, b):
    return bool((np.asarray(a, dtype=np.float64).view(np.uint64) ==
                 np.asarray(b, dtype=np.float64).view(np.uint64)).all())


def hip_runtime():
    """The HIP runtime liblbx.so runs on: the one libamdhip64 mapped into
    this process (opened again by its path, which returns the same copy)."""
    lbx.load()
    with open("/proc/self/maps") as maps:
        paths = {line.split()[-1] for line in maps if "libamdhip64" in line}
    assert len(paths) == 1, paths
    return ctypes.CDLL(paths.pop())


class DeviceBuffer:
    """n doubles of device memory from the HIP runtime liblbx.so is linked to."""

    def __init__(self, n):
        self.hip = hip_runtime()
        self.n = n
        self.ptr = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(8 * n)) == 0

    def read(self):
        out = np.full(self.n, np.nan)
        assert self.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, ctypes.c_size_t(8 * self.n),
                                  2) == 0  # hipMemcpyDeviceToHost
        return out

    def poison(self):
        nan = np.full(self.n, np.nan)
        assert self.hip.hipMemcpy(self.ptr, nan.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(8 * self.n),
                                  1) == 0  # hipMemcpyHostToDevice

    def free(self):
        self.hip.hipFree(self.ptr)


def noalloc_sums(ctx, sim, level, buf):
    buf.poison()
    sim.conservation_noalloc(level, buf.ptr.value)
    ctx.sync()
    return buf.read()


def test_noalloc_conservation_matches_twice(ctx):
    d_out = DeviceBuffer(4)

    dims = (64, 64, 64)
    rho, u = workloads.random_state(dims, seed=5, rho_sigma=1e-3, u_amp=0.01)
    sim = lbx.Sim(ctx, dims)
    try:
        sim.define_level(0, workloads.chop(dims, 32))
        sim.set_tau(0, 0.7)
        sim.init_from_fields(rho, u)
        sim.iterate(3)  # level 0 holds the post-collision state: the sums pull
        a = noalloc_sums(ctx, sim, 0, d_out)
        assert same_bits(a, sim.conservation(0))
    finally:
        sim.close()

    dims = (24, 20, 16)
    rho, u = workloads.random_state(dims, seed=6, rho_sigma=1e-3, u_amp=0.01)
    sim = lbx.Sim(ctx, dims)
    try:
        sim.define_level(0, workloads.chop(dims, 8))
        sim.set_tau(0, 0.8)
        sim.init_from_fields(rho, u)
        sim.define_level(1, [((8, 4, 4), (23, 19, 15)), ((24, 4, 4), (31, 19, 15))])
        sim.set_tau(1, 2 * (0.8 - 0.5) + 0.5)
        sim.fill_from_coarse(1)
        sim.set_fine_mask([((8, 4, 4), (31, 19, 15))])
        sim.iterate(2)
        b0 = noalloc_sums(ctx, sim, 0, d_out)
        b1 = noalloc_sums(ctx, sim, 1, d_out)
        assert same_bits(b0, sim.conservation(0))
        assert same_bits(b1, sim.conservation(1))
        assert not same_bits(a, b0)
    finally:
        sim.close()
        d_out.free()


def test_noalloc_rejects_bad_arguments(ctx):
    dims = (16, 16, 16)
    sim = lbx.Sim(ctx, dims)
    try:
        sim.define_level(0, workloads.chop(dims, 16))
        with pytest.raises(lbx.LbxError):
            sim.conservation_noalloc(0, 0)  # null output
        with pytest.raises(lbx.LbxError):
            sim.conservation_noalloc(1, 1 << 20)  # undefined level
    finally:
        sim.close()
