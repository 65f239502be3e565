"""Python mirror of skyplace::WireLengthGradient
(skyplace/src/wirelength_gradient/WireLengthGradient.h:13-139) and of the
preconditioned total gradient of TargetFunction::updatePointAndGetGrad
(problem_instance/TargetFunction.cu:88-141) over include/skyline_wl.h.

Device arrays are torch CUDA float32 tensors; calls are ordered with torch's
current stream like the density calls.  The netlist is host numpy arrays in
the reference's layout: pins numbered net by net, net_start[n] the first pin
of net n, pin2cell the movable cell of a pin or -1.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .density import _OnStream, _ptr
from .lg import lib as _lglib

_f32p = C.POINTER(C.c_float)
_i32p = C.POINTER(C.c_int32)

EXPORTED = ["wl_last_error", "wl_create", "wl_update_pin_coordinates", "wl_compute_grad",
            "wl_compute_grad_at", "wl_compute_hpwl", "wl_compute_hpwl_async", "wl_get_pins", "wl_destroy", "tf_precondition"]


class WlNetlist(C.Structure):
    _fields_ = [("num_cells", C.c_int32), ("num_nets", C.c_int32), ("num_pins", C.c_int32),
                ("net_start", _i32p), ("pin2cell", _i32p), ("pin_offset_x", _f32p), ("pin_offset_y", _f32p),
                ("pin_x", _f32p), ("pin_y", _f32p), ("net_weight", _f32p)]


class WirelengthError(RuntimeError):
    pass


_bound = None


def lib():
    global _bound
    L = _lglib()
    if _bound is None:
        L.wl_last_error.restype = C.c_char_p
        L.wl_create.argtypes = [C.POINTER(WlNetlist), C.POINTER(C.c_void_p)]
        L.wl_update_pin_coordinates.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.wl_compute_grad.argtypes = [C.c_void_p, C.c_float, C.c_void_p, C.c_void_p, C.c_void_p]
        L.wl_compute_grad_at.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_float, C.c_void_p, C.c_void_p,
                                         C.c_void_p]
        L.wl_compute_hpwl.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.c_void_p]
        L.wl_compute_hpwl_async.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.wl_get_pins.argtypes = [C.c_void_p, _f32p, _f32p]
        L.wl_destroy.argtypes = [C.c_void_p]
        L.tf_precondition.argtypes = [C.c_int32, C.c_float, C.c_float] + [C.c_void_p] * 8 + [C.c_void_p]
        _bound = L
    return _bound


def _check(st):
    if st != 0:
        raise WirelengthError(f"status {st}: {lib().wl_last_error().decode()}")


class WireLengthGradient:
    """skyplace::WireLengthGradient on the GPU."""

    def __init__(self, num_cells, net_start, pin2cell, pin_offset_x, pin_offset_y, pin_x, pin_y,
                 net_weight=None):
        keep = []

        def arr(a, dt, t):
            a = np.ascontiguousarray(a, dt)
            keep.append(a)
            return a.ctypes.data_as(C.POINTER(t))

        n = WlNetlist()
        n.num_cells = int(num_cells)
        n.num_nets = int(len(net_start) - 1)
        n.num_pins = int(len(pin2cell))
        n.net_start = arr(net_start, np.int32, C.c_int32)
        n.pin2cell = arr(pin2cell, np.int32, C.c_int32)
        n.pin_offset_x = arr(pin_offset_x, np.float32, C.c_float)
        n.pin_offset_y = arr(pin_offset_y, np.float32, C.c_float)
        n.pin_x = arr(pin_x, np.float32, C.c_float)
        n.pin_y = arr(pin_y, np.float32, C.c_float)
        n.net_weight = arr(net_weight, np.float32, C.c_float) if net_weight is not None else None
        h = C.c_void_p()
        _check(lib().wl_create(C.byref(n), C.byref(h)))
        self.h = h
        self.num_cells, self.num_nets, self.num_pins = n.num_cells, n.num_nets, n.num_pins
        self.gamma_inv = 0.0

    def setGammaInv(self, g: float):
        self.gamma_inv = float(g)

    def updatePinCoordinates(self, cell_cx, cell_cy):
        with _OnStream() as s:
            _check(lib().wl_update_pin_coordinates(self.h, _ptr(cell_cx, self.num_cells),
                                                   _ptr(cell_cy, self.num_cells), s))

    def computeGrad(self, grad_x, grad_y):
        with _OnStream() as s:
            _check(lib().wl_compute_grad(self.h, self.gamma_inv, _ptr(grad_x, self.num_cells),
                                         _ptr(grad_y, self.num_cells), s))

    def computeGradAt(self, cell_cx, cell_cy, grad_x, grad_y):
        """updatePinCoordinates + computeGrad in one pass."""
        with _OnStream() as s:
            _check(lib().wl_compute_grad_at(self.h, _ptr(cell_cx, self.num_cells), _ptr(cell_cy, self.num_cells),
                                            self.gamma_inv, _ptr(grad_x, self.num_cells),
                                            _ptr(grad_y, self.num_cells), s))

    def computeHPWL(self) -> float:
        v = C.c_double()
        with _OnStream() as s:
            _check(lib().wl_compute_hpwl(self.h, C.byref(v), s))
        return v.value

    def pins(self):
        x = np.zeros(max(self.num_pins, 1), np.float32)
        y = np.zeros(max(self.num_pins, 1), np.float32)
        _check(lib().wl_get_pins(self.h, x.ctypes.data_as(_f32p), y.ctypes.data_as(_f32p)))
        return x[:self.num_pins], y[:self.num_pins]

    def close(self):
        if self.h:
            lib().wl_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def precondition(lam, min_precond, num_pins, dens_precond, wl_gx, wl_gy, dens_gx, dens_gy, gx, gy):
    """jacobianPrecondition (ProblemInstanceKernel.hpp:74-100) on torch tensors."""
    n = int(gx.numel())
    with _OnStream() as s:
        _check(lib().tf_precondition(n, float(lam), float(min_precond), _ptr(num_pins, n), _ptr(dens_precond, n),
                                     _ptr(wl_gx, n), _ptr(wl_gy, n), _ptr(dens_gx, n), _ptr(dens_gy, n),
                                     _ptr(gx, n), _ptr(gy, n), s))
