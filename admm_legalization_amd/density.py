"""Python mirror of skyplace::DensityGradient and skyplace::PoissonSolver
(skyplace/src/density_gradient/DensityGradient.h:23-152, PoissonSolver.h:9-68)
over the C ABI in include/skyline_density.h.

Device arrays are torch CUDA float32 tensors; calls are ordered after the work
already queued on torch's current stream, and that stream's later work is
ordered after them.  Bins are y-major (bin = x + y * numBinX); 2-D maps are returned as
tensors of shape (numBinY, numBinX).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from .lg import lib as _lglib

PREC_BF16, PREC_SPLIT_BF16, PREC_F32 = 0, 1, 2
# Poisson GEMM tile shapes (poisson_set_tile 1..NUM_TILES)
TILE_NAMES = ["tuned", "64x64", "64x128", "128x64", "128x128", "256x128"]
NUM_TILES = len(TILE_NAMES) - 1
PRECISIONS = {"bf16": PREC_BF16, "split_bf16": PREC_SPLIT_BF16, "f32": PREC_F32}

_f32p = C.POINTER(C.c_float)
_i32p = C.POINTER(C.c_int32)


class DensDesign(C.Structure):
    _fields_ = [("die_lx", C.c_float), ("die_ly", C.c_float), ("die_ux", C.c_float), ("die_uy", C.c_float),
                ("target_density", C.c_float), ("num_movable", C.c_int32), ("cell_dx", _f32p),
                ("cell_dy", _f32p), ("is_macro", _i32p), ("is_filler", _i32p), ("num_fixed", C.c_int32),
                ("fixed_lx", _f32p), ("fixed_ly", _f32p), ("fixed_ux", _f32p), ("fixed_uy", _f32p),
                ("num_bin_x", C.c_int32), ("num_bin_y", C.c_int32)]


class DensGeom(C.Structure):
    _fields_ = [("num_bin_x", C.c_int32), ("num_bin_y", C.c_int32), ("die_lx", C.c_float),
                ("die_ly", C.c_float), ("die_ux", C.c_float), ("die_uy", C.c_float), ("bin_w", C.c_float),
                ("bin_h", C.c_float), ("target_density", C.c_float), ("bin_lx", _f32p), ("bin_ux", _f32p),
                ("bin_ly", _f32p), ("bin_uy", _f32p), ("fixed_area", _f32p), ("scaled_bin_area", _f32p),
                ("bin_lambda", _f32p), ("num_movable", C.c_int32), ("cell_dw", _f32p), ("cell_dh", _f32p),
                ("cell_scale", _f32p), ("is_filler", _i32p), ("is_macro", _i32p),
                ("sum_movable_area", C.c_float), ("precision", C.c_int32)]


EXPORTED = ["dens_last_error", "dens_prepare", "dens_create", "dens_compute_grad", "dens_get_overflow",
            "dens_get_overflow_device",
            "dens_get_bin_density", "dens_get_potential", "dens_get_field_x", "dens_get_field_y",
            "dens_get_preconditioner", "dens_get_movable_area", "dens_get_filler_area",
            "dens_set_precision", "dens_destroy", "poisson_create", "poisson_solve", "poisson_solve_noalloc",
            "poisson_set_tile",
            "poisson_get_tile", "poisson_destroy"]

_bound = None
_hip = None


class DensityError(RuntimeError):
    pass


def lib():
    global _bound
    L = _lglib()
    if _bound is None:
        L.dens_last_error.restype = C.c_char_p
        L.dens_prepare.argtypes = [C.POINTER(DensDesign), _i32p] + [_f32p] * 10
        L.dens_create.argtypes = [C.POINTER(DensGeom), C.POINTER(C.c_void_p)]
        L.dens_compute_grad.argtypes = [C.c_void_p] + [C.c_void_p] * 4 + [_f32p, C.c_void_p]
        L.dens_get_overflow.argtypes = [C.c_void_p, _f32p]
        for name in ("dens_get_overflow_device", "dens_get_bin_density", "dens_get_potential", "dens_get_field_x", "dens_get_field_y",
                     "dens_get_preconditioner", "dens_get_movable_area", "dens_get_filler_area"):
            getattr(L, name).argtypes = [C.c_void_p]
            getattr(L, name).restype = C.c_void_p
        L.dens_set_precision.argtypes = [C.c_void_p, C.c_int32]
        L.dens_destroy.argtypes = [C.c_void_p]
        L.poisson_create.argtypes = [C.c_int32, C.c_int32, C.POINTER(C.c_void_p)]
        L.poisson_solve.argtypes = [C.c_void_p] + [C.c_void_p] * 4 + [C.c_int32, C.c_void_p]
        L.poisson_solve_noalloc.argtypes = [C.c_void_p] + [C.c_void_p] * 4 + [C.c_int32, C.c_void_p]
        L.poisson_set_tile.argtypes = [C.c_void_p, C.c_int32]
        L.poisson_get_tile.argtypes = [C.c_void_p, C.c_int32, _i32p, _f32p]
        L.poisson_destroy.argtypes = [C.c_void_p]
        _bound = L
    return _bound


def _check(st):
    if st != 0:
        raise DensityError(f"status {st}: {lib().dens_last_error().decode()}")


def _hiplib():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
    return _hip


_side = None


class _OnStream:
    """Stream for library work, ordered with torch's current stream.  torch's
    default stream has handle 0, which the library (built with the legacy
    null-stream semantics) would not order against torch's own work, so work
    for the default stream goes to a side stream that waits on it and that it
    waits on in turn (event-based, no host synchronisation)."""

    def __enter__(self):
        global _side
        import torch
        self.cur = torch.cuda.current_stream()
        if self.cur.cuda_stream != 0:
            self.side = None
            return C.c_void_p(self.cur.cuda_stream)
        if _side is None:
            _side = torch.cuda.Stream()
        self.side = _side
        self.side.wait_stream(self.cur)
        return C.c_void_p(self.side.cuda_stream)

    def __exit__(self, *exc):
        if self.side is not None:
            self.cur.wait_stream(self.side)
        return False


def _ptr(t, n=None):
    import torch
    if not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != torch.float32 or not t.is_contiguous():
        raise DensityError("expected a contiguous float32 CUDA tensor")
    if n is not None and t.numel() != n:
        raise DensityError(f"expected {n} elements, got {t.numel()}")
    return C.c_void_p(t.data_ptr())


def _device_copy(src_ptr: int, shape):
    """Copy of a library-owned device array into a new torch tensor."""
    import torch
    out = torch.empty(shape, dtype=torch.float32, device="cuda")
    with _OnStream() as s:
        st = _hiplib().hipMemcpyAsync(C.c_void_p(out.data_ptr()), C.c_void_p(src_ptr), out.numel() * 4, 3, s)
    if st != 0:
        raise DensityError(f"hipMemcpyAsync failed ({st})")
    return out


def _f(a):
    a = np.ascontiguousarray(a, np.float32)
    return a, a.ctypes.data_as(_f32p)


def _i(a):
    a = np.ascontiguousarray(a, np.int32)
    return a, a.ctypes.data_as(_i32p)


@dataclass
class Geometry:
    """Bin and cell geometry (host arrays) of one placement."""
    num_bin_x: int
    num_bin_y: int
    die: tuple
    bin_w: float
    bin_h: float
    target_density: float
    bin_lx: np.ndarray
    bin_ux: np.ndarray
    bin_ly: np.ndarray
    bin_uy: np.ndarray
    fixed_area: np.ndarray
    scaled_bin_area: np.ndarray
    cell_dw: np.ndarray
    cell_dh: np.ndarray
    cell_scale: np.ndarray
    is_filler: np.ndarray
    is_macro: np.ndarray
    sum_movable_area: float
    bin_lambda: np.ndarray | None = None


def prepare(die, cell_dx, cell_dy, *, is_macro=None, is_filler=None, fixed_rects=None,
            target_density: float = 1.0, num_bins=None) -> Geometry:
    """SkyPlaceDB::createBins + updateDensitySize + updateFixedOverlapArea
    (SkyPlaceDB.cpp:446-549, 821-852, 909-938) through dens_prepare.
    fixed_rects: (F, 4) array of lx, ly, ux, uy; num_bins: (nbx, nby) override."""
    n = len(cell_dx)
    keep = []
    d = DensDesign()
    d.die_lx, d.die_ly, d.die_ux, d.die_uy = (float(v) for v in die)
    d.target_density = float(target_density)
    d.num_movable = n
    for name, arr in (("cell_dx", cell_dx), ("cell_dy", cell_dy)):
        a, p = _f(arr)
        keep.append(a)
        setattr(d, name, p)
    im = np.zeros(n, np.int32) if is_macro is None else np.asarray(is_macro, np.int32)
    fl = np.zeros(n, np.int32) if is_filler is None else np.asarray(is_filler, np.int32)
    for name, arr in (("is_macro", im), ("is_filler", fl)):
        a, p = _i(arr)
        keep.append(a)
        setattr(d, name, p)
    fr = np.zeros((0, 4), np.float32) if fixed_rects is None else np.asarray(fixed_rects, np.float32).reshape(-1, 4)
    d.num_fixed = len(fr)
    for k, name in enumerate(("fixed_lx", "fixed_ly", "fixed_ux", "fixed_uy")):
        a, p = _f(fr[:, k] if len(fr) else np.zeros(1))
        keep.append(a)
        setattr(d, name, p)
    d.num_bin_x, d.num_bin_y = (int(num_bins[0]), int(num_bins[1])) if num_bins else (0, 0)
    shape = np.zeros(2, np.int32)
    nullf = C.cast(None, _f32p)
    _check(lib().dens_prepare(C.byref(d), shape.ctypes.data_as(_i32p), *([nullf] * 10)))
    nbx, nby = int(shape[0]), int(shape[1])
    B = nbx * nby
    blx, bux = np.zeros(nbx, np.float32), np.zeros(nbx, np.float32)
    bly, buy = np.zeros(nby, np.float32), np.zeros(nby, np.float32)
    fixed, scaled = np.zeros(B, np.float32), np.zeros(B, np.float32)
    dw, dh, sc = np.zeros(n, np.float32), np.zeros(n, np.float32), np.zeros(n, np.float32)
    scal = np.zeros(3, np.float32)
    _check(lib().dens_prepare(C.byref(d), shape.ctypes.data_as(_i32p),
                              *[a.ctypes.data_as(_f32p) for a in (blx, bux, bly, buy, fixed, scaled, dw, dh, sc,
                                                                   scal)]))
    return Geometry(nbx, nby, tuple(float(v) for v in die), float(scal[0]), float(scal[1]),
                    float(target_density), blx, bux, bly, buy, fixed, scaled, dw, dh, sc, fl, im, float(scal[2]))


class PoissonSolver:
    """skyplace::PoissonSolver (PoissonSolver.h:9-68)."""

    def __init__(self, numBinX: int, numBinY: int):
        self.numBinX, self.numBinY = int(numBinX), int(numBinY)
        h = C.c_void_p()
        _check(lib().poisson_create(self.numBinX, self.numBinY, C.byref(h)))
        self.h = h

    def solvePoisson(self, binDensity, potential, electroForceX, electroForceY, precision=PREC_SPLIT_BF16):
        B = self.numBinX * self.numBinY
        args = (_ptr(binDensity, B), _ptr(potential, B), _ptr(electroForceX, B), _ptr(electroForceY, B))
        with _OnStream() as s:
            _check(lib().poisson_solve(self.h, *args, int(precision), s))

    def setTile(self, tile: int):
        """GEMM tile shape: 0 = tuned per precision, 1 = 64x64, 2 = 64x128, 3 = 128x64, 4 = 128x128,
        5 = 256x128."""
        _check(lib().poisson_set_tile(self.h, int(tile)))

    def getTile(self, precision=PREC_SPLIT_BF16):
        """(shape in use, tuner's whole-solve ms per shape) for `precision`."""
        t = C.c_int32()
        ms = (C.c_float * NUM_TILES)()
        _check(lib().poisson_get_tile(self.h, int(precision), C.byref(t), ms))
        return t.value, list(ms)

    def close(self):
        if getattr(self, "h", None):
            lib().poisson_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DensityGradient:
    """skyplace::DensityGradient (DensityGradient.h:23-152) on a Geometry."""

    def __init__(self, geom: Geometry, precision: int = PREC_SPLIT_BF16):
        self.geom = geom
        g = DensGeom()
        keep = []

        def fp(a):
            a, p = _f(a)
            keep.append(a)
            return p

        def ip(a):
            a, p = _i(a)
            keep.append(a)
            return p

        g.num_bin_x, g.num_bin_y = geom.num_bin_x, geom.num_bin_y
        g.die_lx, g.die_ly, g.die_ux, g.die_uy = geom.die
        g.bin_w, g.bin_h, g.target_density = geom.bin_w, geom.bin_h, geom.target_density
        g.bin_lx, g.bin_ux, g.bin_ly, g.bin_uy = (fp(a) for a in (geom.bin_lx, geom.bin_ux, geom.bin_ly,
                                                                  geom.bin_uy))
        g.fixed_area, g.scaled_bin_area = fp(geom.fixed_area), fp(geom.scaled_bin_area)
        g.bin_lambda = fp(geom.bin_lambda) if geom.bin_lambda is not None else C.cast(None, _f32p)
        g.num_movable = len(geom.cell_dw)
        g.cell_dw, g.cell_dh, g.cell_scale = fp(geom.cell_dw), fp(geom.cell_dh), fp(geom.cell_scale)
        g.is_filler, g.is_macro = ip(geom.is_filler), ip(geom.is_macro)
        g.sum_movable_area = geom.sum_movable_area
        g.precision = int(precision)
        h = C.c_void_p()
        _check(lib().dens_create(C.byref(g), C.byref(h)))
        self.h = h
        self.numMovable = g.num_movable

    def numBinX(self):
        return self.geom.num_bin_x

    def numBinY(self):
        return self.geom.num_bin_y

    def setPrecision(self, precision: int):
        _check(lib().dens_set_precision(self.h, int(precision)))

    def computeGrad(self, densityGradX, densityGradY, cellCx, cellCy):
        n = self.numMovable
        args = (_ptr(cellCx, n), _ptr(cellCy, n), _ptr(densityGradX, n), _ptr(densityGradY, n))
        with _OnStream() as s:
            _check(lib().dens_compute_grad(self.h, *args, None, s))

    def overflow(self) -> float:
        v = C.c_float()
        _check(lib().dens_get_overflow(self.h, C.byref(v)))
        return float(v.value)

    def _map(self, fn):
        return _device_copy(getattr(lib(), fn)(self.h), (self.geom.num_bin_y, self.geom.num_bin_x))

    def getDevicePotential(self):
        return self._map("dens_get_potential")

    def getDeviceBinDensity(self):
        return self._map("dens_get_bin_density")

    def getDeviceFieldX(self):
        return self._map("dens_get_field_x")

    def getDeviceFieldY(self):
        return self._map("dens_get_field_y")

    def getDeviceMovableArea(self):
        return self._map("dens_get_movable_area")

    def getDeviceFillerArea(self):
        return self._map("dens_get_filler_area")

    def getDevicePreconditioner(self):
        return _device_copy(lib().dens_get_preconditioner(self.h), (self.numMovable,))

    def close(self):
        if getattr(self, "h", None):
            lib().dens_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
