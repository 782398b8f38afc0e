"""ctypes binding of the C ABI in include/lbx.h (liblbx.so, built in-tree).

This is the Python face of the same drop-in boundary the C++ AmrSim facade
uses.  There is no fallback: if liblbx.so is missing or no GPU is visible,
every entry point raises LbxError.
"""
import ctypes
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "liblbx.so"

LBX_OK = 0
LBX_E_ARG = -1
LBX_E_STATE = -2
LBX_E_HIP = -3
LBX_E_NOMEM = -4
LBX_E_UNSUPPORTED = -5
LBX_FP64 = 0
LBX_FP32_SHIFTED = 1
LBX_MAX_LEVELS = 6
LBX_SERVE_FIFO = 0
LBX_SERVE_SHORTEST_FIRST = 1

# Every extern "C" entry point of include/lbx.h with its ctypes signature.
_vp = ctypes.c_void_p
_i32p = ctypes.POINTER(ctypes.c_int32)
_dp = ctypes.POINTER(ctypes.c_double)
SIGNATURES = {
    "lbx_ctx_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp)]),
    "lbx_ctx_destroy": (ctypes.c_int, [_vp]),
    "lbx_last_error": (ctypes.c_char_p, [_vp]),
    "lbx_sync": (ctypes.c_int, [_vp]),
    "lbx_mem_info": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]),
    "lbx_sim_create": (ctypes.c_int, [_vp, _i32p, _i32p, ctypes.c_int, ctypes.c_double, ctypes.POINTER(_vp)]),
    "lbx_sim_destroy": (ctypes.c_int, [_vp]),
    "lbx_level_define": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _i32p]),
    "lbx_level_remake": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _i32p]),
    "lbx_level_clear": (ctypes.c_int, [_vp, ctypes.c_int]),
    "lbx_hydro_level": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp]),
    "lbx_tag_cells": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_void_p]),
    "lbx_cluster_tags": (ctypes.c_int, [_i32p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                        ctypes.c_int32, _i32p, _i32p]),
    "lbx_cluster_tags_allowed": (ctypes.c_int, [_i32p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                                ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _i32p, _i32p]),
    "lbx_set_ref_ratio": (ctypes.c_int, [_vp, ctypes.c_int]),
    "lbx_set_relax": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_double]),
    "lbx_set_fine_mask": (ctypes.c_int, [_vp, ctypes.c_int, _i32p]),
    "lbx_set_level_mask": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _i32p]),
    "lbx_fill_from_coarse": (ctypes.c_int, [_vp, ctypes.c_int]),
    "lbx_init_equilibrium": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _dp, _dp]),
    "lbx_upload_f": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _dp]),
    "lbx_download_f": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _dp]),
    "lbx_iterate": (ctypes.c_int, [_vp, ctypes.c_int]),
    "lbx_hydro": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _dp, _dp]),
    "lbx_conservation": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    "lbx_conservation_noalloc": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_void_p]),
    "lbx_profile_enable": (ctypes.c_int, [_vp, ctypes.c_int]),
    "lbx_profile_read": (ctypes.c_int, [_vp, _dp, ctypes.POINTER(ctypes.c_int64)]),
    "lbx_bench_copy": (ctypes.c_int, [_vp, ctypes.c_size_t, ctypes.c_int, _dp]),
    "lbx_set_autotune": (ctypes.c_int, [_vp, ctypes.c_int]),
    "lbx_tuned_shape": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _i32p, _i32p, _i32p,
                                       ctypes.POINTER(ctypes.c_float)]),
}

# Serving structs (include/lbx.h).
class Request(ctypes.Structure):
    _fields_ = [("dims", ctypes.c_int32 * 3), ("steps", ctypes.c_int32),
                ("tau_s", ctypes.c_double), ("tau_b", ctypes.c_double),
                ("init_kind", ctypes.c_int32), ("axis", ctypes.c_int32), ("plane", ctypes.c_int32),
                ("amplitude", ctypes.c_double),
                ("rho", _dp), ("u", _dp), ("out_rho", _dp), ("out_u", _dp), ("out_f", _dp),
                ("n_boxes", ctypes.c_int32), ("box_lohi", _i32p),
                ("n_fine_boxes", ctypes.c_int32), ("fine_lohi", _i32p), ("ref_ratio", ctypes.c_int32)]


class Result(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint64), ("status", ctypes.c_int32), ("steps", ctypes.c_int32),
                ("cells", ctypes.c_int64), ("mass", ctypes.c_double), ("momentum", ctypes.c_double * 3),
                ("t_submit", ctypes.c_double), ("t_admit", ctypes.c_double), ("t_done", ctypes.c_double)]


LBX_TAG_DENSITY_GRADIENT = 0
LBX_TAG_SPEED = 1
LBX_INIT_PLANE_PULSE = 0
LBX_INIT_FIELDS = 1

SIGNATURES.update({
    "lbx_serve_create": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_size_t, ctypes.c_int,
                                        ctypes.POINTER(_vp)]),
    "lbx_serve_destroy": (ctypes.c_int, [_vp]),
    "lbx_serve_set_policy": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double]),
    "lbx_serve_last_error": (ctypes.c_char_p, [_vp]),
    "lbx_serve_submit": (ctypes.c_int, [_vp, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_uint64)]),
    "lbx_serve_progress": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                          ctypes.POINTER(ctypes.c_int)]),
    "lbx_serve_drain": (ctypes.c_int, [_vp]),
    "lbx_serve_poll": (ctypes.c_int, [_vp, ctypes.POINTER(Result), ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
    "lbx_serve_stats": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64),
                                       ctypes.POINTER(ctypes.c_size_t)]),
})

_LIB = None


class LbxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"lbx error {code}: {msg}")
        self.code = code


def load():
    """Load liblbx.so and bind every entry point (no GPU call is made)."""
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise LbxError(LBX_E_STATE, f"{LIB_PATH} is missing: build it first "
                                        "(python -c 'import __graft_entry__ as g; g.build()')")
        lib = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = lib
    return _LIB


def _ptr(a, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def _lohi(boxes):
    arr = np.array([list(lo) + list(hi) for lo, hi in boxes], dtype=np.int32).reshape(-1)
    return arr


class Context:
    """A device context: one HIP stream and scratch on one GPU."""

    def __init__(self, device=0):
        self._lib = load()
        h = ctypes.c_void_p()
        rc = self._lib.lbx_ctx_create(int(device), ctypes.byref(h))
        if rc != LBX_OK:
            raise LbxError(rc, self._lib.lbx_last_error(None).decode())
        self.h = h

    def check(self, rc):
        if rc != LBX_OK:
            raise LbxError(rc, self._lib.lbx_last_error(self.h).decode())

    def sync(self):
        self.check(self._lib.lbx_sync(self.h))

    def mem_info(self):
        f, t = ctypes.c_size_t(), ctypes.c_size_t()
        self.check(self._lib.lbx_mem_info(self.h, ctypes.byref(f), ctypes.byref(t)))
        return f.value, t.value

    def profile_enable(self, on=True):
        self.check(self._lib.lbx_profile_enable(self.h, 1 if on else 0))

    def profile_read(self):
        ms, n = ctypes.c_double(), ctypes.c_int64()
        self.check(self._lib.lbx_profile_read(self.h, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def bench_copy(self, nbytes, reps=10):
        """Measured read+write copy bandwidth in GB/s."""
        g = ctypes.c_double()
        self.check(self._lib.lbx_bench_copy(self.h, int(nbytes), int(reps), ctypes.byref(g)))
        return g.value

    def set_autotune(self, on=True):
        self.check(self._lib.lbx_set_autotune(self.h, 1 if on else 0))

    def tuned_shape(self, nx, ny, precision=LBX_FP64):
        """((bx, by, workgroups_per_cu_cap), ms per proxy launch) that launches over
        rows of nx x ny use: the autotuner's choice, or the default shape (ms 0)
        if that geometry was never tuned."""
        bx, by, occ, ms = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32(), ctypes.c_float()
        self.check(self._lib.lbx_tuned_shape(self.h, int(nx), int(ny), int(precision), ctypes.byref(bx),
                                             ctypes.byref(by), ctypes.byref(occ), ctypes.byref(ms)))
        return (bx.value, by.value, occ.value), ms.value

    def close(self):
        if self.h:
            self._lib.lbx_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def cluster_tags(dims, tags, blocking=4, buffer=1, max_grid_size=32, allowed=None):
    """Boxes ((lo), (hi)) of the next finer level covering the tagged cells of
    a tag array tags[nz][ny][nx] over [0, dims) (lbx_cluster_tags; with an
    `allowed` mask of the same shape, lbx_cluster_tags_allowed: tiles holding
    a disallowed cell are not refined).  Host only, no GPU."""
    lib = load()
    d = np.array([int(x) for x in dims], dtype=np.int32)
    t = np.ascontiguousarray(tags, dtype=np.uint8).reshape(-1)
    if t.size != int(np.prod(d)):
        raise LbxError(LBX_E_ARG, "tag array does not match dims")
    a = None
    if allowed is not None:
        a = np.ascontiguousarray(allowed, dtype=np.uint8).reshape(-1)
        if a.size != t.size:
            raise LbxError(LBX_E_ARG, "allowed mask does not match dims")

    def call(max_boxes, out, n):
        if a is None:
            return lib.lbx_cluster_tags(_ptr(d, ctypes.c_int32), t.ctypes.data_as(ctypes.c_void_p), int(blocking),
                                        int(buffer), int(max_grid_size), max_boxes, out, ctypes.byref(n))
        return lib.lbx_cluster_tags_allowed(_ptr(d, ctypes.c_int32), t.ctypes.data_as(ctypes.c_void_p),
                                            a.ctypes.data_as(ctypes.c_void_p), int(blocking), int(buffer),
                                            int(max_grid_size), max_boxes, out, ctypes.byref(n))

    n = ctypes.c_int32()
    rc = call(0, None, n)
    if rc != LBX_OK:
        raise LbxError(rc, "lbx_cluster_tags failed")
    out = np.zeros(6 * max(n.value, 1), dtype=np.int32)
    rc = call(n.value, _ptr(out, ctypes.c_int32), n)
    if rc != LBX_OK:
        raise LbxError(rc, "lbx_cluster_tags failed")
    return [(tuple(int(v) for v in out[6 * b:6 * b + 3]), tuple(int(v) for v in out[6 * b + 3:6 * b + 6]))
            for b in range(n.value)]


class Sim:
    """A periodic D3Q15 simulation held on the GPU."""

    def __init__(self, ctx, dims, precision=LBX_FP64, rho0=1.0, periodic=(1, 1, 1)):
        self.ctx = ctx
        self._lib = ctx._lib
        self.dims = tuple(int(d) for d in dims)
        self.precision = precision
        d = np.array(self.dims, dtype=np.int32)
        p = np.array(periodic, dtype=np.int32)
        h = ctypes.c_void_p()
        ctx.check(self._lib.lbx_sim_create(ctx.h, _ptr(d, ctypes.c_int32), _ptr(p, ctypes.c_int32),
                                           int(precision), float(rho0), ctypes.byref(h)))
        self.h = h
        self.boxes = {}

    def close(self):
        if self.h:
            self._lib.lbx_sim_destroy(self.h)
            self.h = None

    def define_level(self, level, boxes):
        """Boxes of `level`; finer levels are released (lbx_level_define)."""
        lohi = _lohi(boxes)
        self.ctx.check(self._lib.lbx_level_define(self.h, level, len(boxes), _ptr(lohi, ctypes.c_int32)))
        for l in [l for l in self.boxes if l > level]:
            del self.boxes[l]
        self.boxes[level] = [(tuple(lo), tuple(hi)) for lo, hi in boxes]

    def remake_level(self, level, boxes):
        """New boxes for the finest level, keeping its data (old values where
        an old box covered the cell, else the parent's)."""
        lohi = _lohi(boxes)
        self.ctx.check(self._lib.lbx_level_remake(self.h, level, len(boxes), _ptr(lohi, ctypes.c_int32)))
        self.boxes[level] = [(tuple(lo), tuple(hi)) for lo, hi in boxes]

    def tag_cells(self, criterion, threshold, level=0):
        """Refinement tags of level 0, [nz][ny][nx] uint8 (lbx_tag_cells)."""
        nx, ny, nz = self.dims
        tags = np.zeros((nz, ny, nx), dtype=np.uint8)
        self.ctx.check(self._lib.lbx_tag_cells(self.h, level, int(criterion), float(threshold),
                                               tags.ctypes.data_as(ctypes.c_void_p)))
        return tags

    def set_ref_ratio(self, ratio):
        """Refinement ratio (2, 3 or 4) between consecutive levels, the same on
        every level; before level 1 is defined."""
        self.ctx.check(self._lib.lbx_set_ref_ratio(self.h, int(ratio)))

    def set_relax(self, level, omega_s, omega_b):
        self.ctx.check(self._lib.lbx_set_relax(self.h, level, float(omega_s), float(omega_b)))

    def set_fine_mask(self, boxes):
        """Coarse cells under coarsen(box, r) of these (fine index space)
        boxes are masked in the two-level coarse collide."""
        lohi = _lohi(boxes) if boxes else np.zeros(6, dtype=np.int32)
        self.ctx.check(self._lib.lbx_set_fine_mask(self.h, len(boxes), _ptr(lohi, ctypes.c_int32)))

    def set_level_mask(self, level, boxes):
        """Cells of `level` under coarsen(box, r) of these boxes (index space
        of level + 1) are masked in that level's collide."""
        lohi = _lohi(boxes) if boxes else np.zeros(6, dtype=np.int32)
        self.ctx.check(self._lib.lbx_set_level_mask(self.h, level, len(boxes), _ptr(lohi, ctypes.c_int32)))

    def clear_level(self, level):
        """Release `level` and every finer level."""
        self.ctx.check(self._lib.lbx_level_clear(self.h, level))
        for l in list(self.boxes):
            if l >= level:
                del self.boxes[l]

    def fill_from_coarse(self, level=1):
        self.ctx.check(self._lib.lbx_fill_from_coarse(self.h, level))

    def set_tau(self, level, tau_s, tau_b=None):
        """omega = 1/(tau + 1/2) in double, as src/AmrSim.cpp:125-126."""
        tau_b = tau_s if tau_b is None else tau_b
        self.set_relax(level, 1.0 / (tau_s + 0.5), 1.0 / (tau_b + 0.5))

    @staticmethod
    def _shape(box):
        lo, hi = box
        return (hi[2] - lo[2] + 1, hi[1] - lo[1] + 1, hi[0] - lo[0] + 1)

    def init_equilibrium(self, level, box, rho, u):
        shape = self._shape(self.boxes[level][box])
        rho = np.ascontiguousarray(rho, dtype=np.float64).reshape(shape)
        u = np.ascontiguousarray(u, dtype=np.float64).reshape((3,) + shape)
        self.ctx.check(self._lib.lbx_init_equilibrium(self.h, level, box, _ptr(rho, ctypes.c_double),
                                                      _ptr(u, ctypes.c_double)))

    def upload_f(self, level, box, f):
        shape = (15,) + self._shape(self.boxes[level][box])
        f = np.ascontiguousarray(f, dtype=np.float64).reshape(shape)
        self.ctx.check(self._lib.lbx_upload_f(self.h, level, box, _ptr(f, ctypes.c_double)))

    def download_f(self, level, box):
        f = np.empty((15,) + self._shape(self.boxes[level][box]))
        self.ctx.check(self._lib.lbx_download_f(self.h, level, box, _ptr(f, ctypes.c_double)))
        return f

    def iterate(self, n):
        self.ctx.check(self._lib.lbx_iterate(self.h, int(n)))

    def hydro(self, level, box):
        shape = self._shape(self.boxes[level][box])
        rho = np.empty(shape)
        u = np.empty((3,) + shape)
        self.ctx.check(self._lib.lbx_hydro(self.h, level, box, _ptr(rho, ctypes.c_double),
                                           _ptr(u, ctypes.c_double)))
        return rho, u

    def hydro_level(self, level=0):
        """Per-box (rho, u) of every box of a level from one bulk call."""
        bl = self.boxes[level]
        sizes = [int(np.prod(self._shape(b))) for b in bl]
        rho = np.empty(sum(sizes))
        u = np.empty(3 * sum(sizes))
        self.ctx.check(self._lib.lbx_hydro_level(self.h, level, _ptr(rho, ctypes.c_double),
                                                 _ptr(u, ctypes.c_double)))
        out, off = [], 0
        for b, n in zip(bl, sizes):
            shape = self._shape(b)
            out.append((rho[off:off + n].reshape(shape), u[3 * off:3 * off + 3 * n].reshape((3,) + shape)))
            off += n
        return out

    def conservation(self, level=0):
        out = np.zeros(4)
        self.ctx.check(self._lib.lbx_conservation(self.h, level, _ptr(out, ctypes.c_double)))
        return out

    def conservation_noalloc(self, level, d_out):
        """lbx_conservation_noalloc: the 4 sums written in stream order to
        d_out, a device address of at least 4 doubles; no allocation, no wait."""
        self.ctx.check(self._lib.lbx_conservation_noalloc(self.h, level, ctypes.c_void_p(d_out)))

    # --- whole-level helpers (level 0 tiles the domain)
    def level_array(self, level, per_box, ncomp):
        """Assemble per-box arrays ([ncomp][nz][ny][nx] or [nz][ny][nx]) into
        one domain array."""
        nx, ny, nz = self.dims
        out = np.empty(((ncomp,) if ncomp else ()) + (nz, ny, nx))
        for (lo, hi), a in zip(self.boxes[level], per_box):
            out[..., lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] = a
        return out

    def get_f(self, level=0):
        return self.level_array(level, [self.download_f(level, b) for b in range(len(self.boxes[level]))], 15)

    def get_f_bbox(self, level):
        """Populations over the bounding box of the level's boxes (cells
        outside every box are zero): [15][nz][ny][nx] and the box lo."""
        bl = self.boxes[level]
        lo = [min(b[0][d] for b in bl) for d in range(3)]
        hi = [max(b[1][d] for b in bl) for d in range(3)]
        out = np.zeros((15, hi[2] - lo[2] + 1, hi[1] - lo[1] + 1, hi[0] - lo[0] + 1))
        for b, (blo, bhi) in enumerate(bl):
            out[:, blo[2] - lo[2]:bhi[2] - lo[2] + 1, blo[1] - lo[1]:bhi[1] - lo[1] + 1,
                blo[0] - lo[0]:bhi[0] - lo[0] + 1] = self.download_f(level, b)
        return out, lo

    def get_hydro_bbox(self, level):
        """rho [nz][ny][nx] and u [3][nz][ny][nx] over the bounding box of the
        level's boxes (cells outside every box are zero), and the box lo."""
        bl = self.boxes[level]
        lo = [min(b[0][d] for b in bl) for d in range(3)]
        hi = [max(b[1][d] for b in bl) for d in range(3)]
        shape = (hi[2] - lo[2] + 1, hi[1] - lo[1] + 1, hi[0] - lo[0] + 1)
        rho, u = np.zeros(shape), np.zeros((3,) + shape)
        for b, (blo, bhi) in enumerate(bl):
            sl = (slice(blo[2] - lo[2], bhi[2] - lo[2] + 1), slice(blo[1] - lo[1], bhi[1] - lo[1] + 1),
                  slice(blo[0] - lo[0], bhi[0] - lo[0] + 1))
            rho[sl], u[(slice(None),) + sl] = self.hydro(level, b)
        return rho, u, lo

    def set_f(self, f, level=0):
        for b, (lo, hi) in enumerate(self.boxes[level]):
            self.upload_f(level, b, f[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1])

    def get_hydro(self, level=0):
        parts = [self.hydro(level, b) for b in range(len(self.boxes[level]))]
        return (self.level_array(level, [p[0] for p in parts], 0),
                self.level_array(level, [p[1] for p in parts], 3))

    def init_from_fields(self, rho, u, level=0):
        """rho [nz][ny][nx], u [3][nz][ny][nx] over the domain."""
        for b, (lo, hi) in enumerate(self.boxes[level]):
            sl = (slice(lo[2], hi[2] + 1), slice(lo[1], hi[1] + 1), slice(lo[0], hi[0] + 1))
            self.init_equilibrium(level, b, rho[sl], u[(slice(None),) + sl])


class Server:
    """Batched serving of independent periodic runs (lbx_serve_*)."""

    def __init__(self, ctx, precision=LBX_FP32_SHIFTED, rho0=1.0, arena_bytes=0, max_active=0,
                 policy=None, max_wait_s=0.0):
        self.ctx = ctx
        self._lib = ctx._lib
        h = ctypes.c_void_p()
        rc = self._lib.lbx_serve_create(ctx.h, int(precision), float(rho0), int(arena_bytes),
                                        int(max_active), ctypes.byref(h))
        if rc != LBX_OK:
            raise LbxError(rc, "lbx_serve_create failed")
        self.h = h
        self._keep = {}  # request id -> buffers that must outlive the run
        if policy is not None:
            self.set_policy(policy, max_wait_s)

    def set_policy(self, policy, max_wait_s=0.0):
        """LBX_SERVE_FIFO or LBX_SERVE_SHORTEST_FIRST admission; with the
        latter, requests queued longer than max_wait_s seconds go first."""
        self._check(self._lib.lbx_serve_set_policy(self.h, int(policy), float(max_wait_s)))

    def _check(self, rc):
        if rc != LBX_OK:
            raise LbxError(rc, self._lib.lbx_serve_last_error(self.h).decode())

    def submit(self, dims, steps, tau_s, tau_b=None, amplitude=0.0, axis=0, plane=None,
               rho=None, u=None, want_rho=False, want_f=False, boxes=None, fine_boxes=None, ref_ratio=0):
        """Queue one run; returns (id, outputs) where outputs holds the numpy
        arrays the server fills at completion.  fine_boxes (fine index space,
        ref_ratio * dims; ratio 2 unless ref_ratio is 3 or 4) makes it a
        two-level run over level-0 `boxes` (default: one box); outputs are
        then level 0's."""
        nx, ny, nz = (int(d) for d in dims)
        r = Request()
        r.dims[:] = (nx, ny, nz)
        r.steps = int(steps)
        r.tau_s = float(tau_s)
        r.tau_b = float(tau_s if tau_b is None else tau_b)
        keep = {}
        if rho is not None:
            keep["rho_in"] = np.ascontiguousarray(rho, dtype=np.float64).reshape(nz, ny, nx)
            keep["u_in"] = np.ascontiguousarray(u, dtype=np.float64).reshape(3, nz, ny, nx)
            r.init_kind = LBX_INIT_FIELDS
            r.rho = _ptr(keep["rho_in"], ctypes.c_double)
            r.u = _ptr(keep["u_in"], ctypes.c_double)
        else:
            r.init_kind = LBX_INIT_PLANE_PULSE
            r.axis = int(axis)
            r.plane = int(dims[axis] // 2 if plane is None else plane)
            r.amplitude = float(amplitude)
        outs = {}
        if want_rho:
            outs["rho"] = np.empty((nz, ny, nx))
            outs["u"] = np.empty((3, nz, ny, nx))
            r.out_rho = _ptr(outs["rho"], ctypes.c_double)
            r.out_u = _ptr(outs["u"], ctypes.c_double)
        if want_f:
            outs["f"] = np.empty((15, nz, ny, nx))
            r.out_f = _ptr(outs["f"], ctypes.c_double)
        if boxes:
            keep["boxes"] = _lohi(boxes)
            r.n_boxes = len(boxes)
            r.box_lohi = _ptr(keep["boxes"], ctypes.c_int32)
        if fine_boxes:
            keep["fine"] = _lohi(fine_boxes)
            r.n_fine_boxes = len(fine_boxes)
            r.fine_lohi = _ptr(keep["fine"], ctypes.c_int32)
        r.ref_ratio = int(ref_ratio)
        rid = ctypes.c_uint64()
        self._check(self._lib.lbx_serve_submit(self.h, ctypes.byref(r), ctypes.byref(rid)))
        keep.update(outs)
        self._keep[rid.value] = keep
        return rid.value, outs

    def progress(self, max_steps=8):
        a, q = ctypes.c_int(), ctypes.c_int()
        self._check(self._lib.lbx_serve_progress(self.h, int(max_steps), ctypes.byref(a), ctypes.byref(q)))
        return a.value, q.value

    def drain(self):
        self._check(self._lib.lbx_serve_drain(self.h))

    def poll(self, max_results=1024):
        buf = (Result * max_results)()
        n = ctypes.c_int()
        self._check(self._lib.lbx_serve_poll(self.h, buf, max_results, ctypes.byref(n)))
        out = []
        for i in range(n.value):
            r = buf[i]
            out.append({"id": r.id, "status": r.status, "steps": r.steps, "cells": r.cells, "mass": r.mass,
                        "momentum": tuple(r.momentum), "t_submit": r.t_submit, "t_admit": r.t_admit,
                        "t_done": r.t_done})
            self._keep.pop(r.id, None)
        return out

    def stats(self):
        lu, st, fr = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_size_t()
        self._check(self._lib.lbx_serve_stats(self.h, ctypes.byref(lu), ctypes.byref(st), ctypes.byref(fr)))
        return {"lattice_updates": lu.value, "steps_launched": st.value, "arena_largest_free": fr.value}

    def close(self):
        if self.h:
            self._lib.lbx_serve_destroy(self.h)
            self.h = None


def c_order_to_xfast(rho_c, u_c, dims):
    """Reference C-order inputs (i slowest, velocity component fastest,
    include/AmrSim.h:79-83) to x-fastest rho[nz][ny][nx], u[3][nz][ny][nx]."""
    nx, ny, nz = dims
    rho = np.asarray(rho_c, dtype=np.float64).reshape(nx, ny, nz).transpose(2, 1, 0)
    u = np.asarray(u_c, dtype=np.float64).reshape(nx, ny, nz, 3).transpose(3, 2, 1, 0)
    return np.ascontiguousarray(rho), np.ascontiguousarray(u)
