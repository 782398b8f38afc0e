"""ctypes wrapper over liblbx_oracle.so, the CPU restatement of the reference.

Test infrastructure only: imported by tests/, by the smoke entry point and by
bench.py's CPU-baseline leg, never by the product package.
"""
import ctypes
import os
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def load(path):
    """A ctypes handle on an oracle library at `path` with every entry point
    typed (the default library, or a variant built by a test)."""
    L = ctypes.CDLL(str(path))
    _bind(L)
    return L


def lib():
    global _LIB
    if _LIB is None:
        path = _HERE / "liblbx_oracle.so"
        if not path.exists():
            raise RuntimeError(f"{path} is missing; run `make -C oracle` (or __graft_entry__.build())")
        _LIB = load(path)
    return _LIB


def _bind(L):
    L.orc_tables.argtypes = [_dp, _dp, _ip]
    L.orc_collide_cell.argtypes = [_dp, ctypes.c_double, ctypes.c_double]
    L.orc_equilibrium_cell.argtypes = [ctypes.c_double, _dp, _dp]
    L.orc_hydro_cell.argtypes = [_dp, ctypes.POINTER(ctypes.c_double), _dp]
    L.orc_stream_periodic.argtypes = [_dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.orc_sim_create.restype = ctypes.c_void_p
    L.orc_sim_create.argtypes = [ctypes.c_int] * 4 + [_ip, ctypes.c_double, ctypes.c_double]
    L.orc_sim_destroy.argtypes = [ctypes.c_void_p]
    L.orc_sim_init.argtypes = [ctypes.c_void_p, _dp, _dp]
    L.orc_sim_add_fine.argtypes = [ctypes.c_void_p, ctypes.c_int, _ip, ctypes.c_int, _ip]
    L.orc_sim_set_ratio.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.orc_sim_iterate.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.orc_sim_bbox.argtypes = [ctypes.c_void_p, ctypes.c_int, _ip]
    L.orc_sim_get_f.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
    L.orc_sim_set_f.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
    L.orc_sim_hydro.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp]
    L.orc_sim_time.restype = ctypes.c_double
    L.orc_sim_time.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.orc_sim_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.orc_sim_tau.restype = ctypes.c_double
    L.orc_sim_tau.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    L.orc_sim_nbox.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.orc_sim_box.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _ip]
    L.orc_sim_get_grown.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]
    L.orc_sim_set_grown.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]
    L.orc_stage.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]


STAGE_INIT_POST_COLLISION = 0
STAGE_ZERO_INVALID = 1
STAGE_SUM_FROM_FINE = 2
HALO = 2


def tables():
    M = np.zeros((15, 15))
    MI = np.zeros((15, 15))
    C = np.zeros((15, 3), dtype=np.int32)
    lib().orc_tables(M, MI, C)
    return M, MI, C


def collide_cell(f, omega_s, omega_b):
    g = np.ascontiguousarray(f, dtype=np.float64).copy()
    lib().orc_collide_cell(g, float(omega_s), float(omega_b))
    return g


def equilibrium_cell(rho, u):
    f = np.zeros(15)
    lib().orc_equilibrium_cell(float(rho), np.ascontiguousarray(u, dtype=np.float64), f)
    return f


def hydro_cell(f):
    rho = ctypes.c_double()
    u = np.zeros(3)
    lib().orc_hydro_cell(np.ascontiguousarray(f, dtype=np.float64), ctypes.byref(rho), u)
    return rho.value, u


def stream_periodic(f):
    """Pull-stream a [15][nz][ny][nx] periodic lattice."""
    f = np.ascontiguousarray(f, dtype=np.float64)
    out = np.empty_like(f)
    lib().orc_stream_periodic(f, out, f.shape[3], f.shape[2], f.shape[1])
    return out


def max_size_chop(lo, hi, max_grid_size):
    """Chop a box into pieces no longer than max_grid_size per direction the
    way AMReX's BoxList::maxSize does for MakeBaseGrids: per direction, reduce
    by common powers of two, split into near-equal blocks, cut from the top."""
    boxes = [(list(lo), list(hi))]
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
                for k in range(numblk - 1):
                    ksize = (sz + 1 if k < extra else sz) * ratio
                    pos = bhi[d] - ksize + 1
                    top_lo = list(blo)
                    top_lo[d] = pos
                    out.append((top_lo, list(bhi)))
                    bhi[d] = pos - 1
            out.append((blo, bhi))
        boxes = out
    return boxes


def boxes_to_lohi(boxes):
    return np.array([list(lo) + list(hi) for lo, hi in boxes], dtype=np.int32).reshape(-1)


class Sim:
    """One oracle simulation (level 0 periodic; optional fine level)."""

    def __init__(self, nx, ny, nz, tau_s, tau_b=None, boxes=None, max_grid_size=32, library=None):
        if tau_b is None:
            tau_b = tau_s
        if boxes is None:
            boxes = max_size_chop((0, 0, 0), (nx - 1, ny - 1, nz - 1), max_grid_size)
        self.dims = (nx, ny, nz)
        self.boxes = boxes
        self._L = library if library is not None else lib()
        lohi = boxes_to_lohi(boxes)
        self._h = self._L.orc_sim_create(nx, ny, nz, len(boxes), lohi, float(tau_s), float(tau_b))
        if not self._h:
            raise ValueError("oracle rejected the box layout")

    def close(self):
        if self._h:
            self._L.orc_sim_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def init(self, rho_c, u_c):
        """rho_c: nx*ny*nz values in C order (i slowest); u_c: 3 per cell."""
        self._L.orc_sim_init(self._h, np.ascontiguousarray(rho_c, dtype=np.float64).reshape(-1),
                             np.ascontiguousarray(u_c, dtype=np.float64).reshape(-1))

    def add_fine(self, fine_boxes, mask_boxes):
        rc = self._L.orc_sim_add_fine(self._h, len(fine_boxes), boxes_to_lohi(fine_boxes),
                                      len(mask_boxes), boxes_to_lohi(mask_boxes))
        if rc != 0:
            raise ValueError(f"orc_sim_add_fine failed ({rc})")

    def set_ratio(self, ratio):
        """Refinement ratio (2, 3 or 4) of the levels added after this call."""
        if self._L.orc_sim_set_ratio(self._h, int(ratio)) != 0:
            raise ValueError(f"unsupported refinement ratio {ratio}")

    def iterate(self, n):
        self._L.orc_sim_iterate(self._h, int(n))

    def bbox(self, level=0):
        bb = np.zeros(6, dtype=np.int32)
        self._L.orc_sim_bbox(self._h, level, bb)
        return bb

    def _shape(self, level):
        bb = self.bbox(level)
        return (int(bb[5] - bb[2] + 1), int(bb[4] - bb[1] + 1), int(bb[3] - bb[0] + 1))

    def get_f(self, level=0):
        nz, ny, nx = self._shape(level)
        f = np.zeros((15, nz, ny, nx))
        self._L.orc_sim_get_f(self._h, level, f)
        return f

    def set_f(self, f, level=0):
        self._L.orc_sim_set_f(self._h, level, np.ascontiguousarray(f, dtype=np.float64))

    def hydro(self, level=0):
        nz, ny, nx = self._shape(level)
        rho = np.zeros((nz, ny, nx))
        u = np.zeros((3, nz, ny, nx))
        self._L.orc_sim_hydro(self._h, level, rho, u)
        return rho, u

    def time(self, level=0):
        return self._L.orc_sim_time(self._h, level)

    def step(self, level=0):
        return self._L.orc_sim_step(self._h, level)

    def tau(self, level=0, which=0):
        return self._L.orc_sim_tau(self._h, level, which)

    # ---- test hooks: whole FABs (valid + 2 ghost layers) and single stages
    def nbox(self, level=0):
        return self._L.orc_sim_nbox(self._h, level)

    def box(self, level, b):
        lohi = np.zeros(6, dtype=np.int32)
        self._L.orc_sim_box(self._h, level, b, lohi)
        return lohi[:3].copy(), lohi[3:].copy()

    def _grown_shape(self, level, b):
        lo, hi = self.box(level, b)
        n = hi - lo + 1 + 2 * HALO
        return (15, int(n[2]), int(n[1]), int(n[0]))

    def get_grown(self, level, b, which=0):
        """Box b's FAB grown by 2 ghost layers, [15][gz][gy][gx]; which 0 =
        the `now` state, 1 = `next`."""
        out = np.zeros(self._grown_shape(level, b))
        self._L.orc_sim_get_grown(self._h, level, which, b, out)
        return out

    def set_grown(self, level, b, data, which=0):
        data = np.ascontiguousarray(data, dtype=np.float64).reshape(self._grown_shape(level, b))
        self._L.orc_sim_set_grown(self._h, level, which, b, data)

    def stage(self, level, stage):
        if self._L.orc_stage(self._h, level, stage) != 0:
            raise ValueError(f"stage {stage} on level {level} rejected")
