"""Python mirror of skyplace::SkyPlace (skyplace/include/skyplace/SkyPlace.h)
over include/skyline_gp.h: the global placer that turns a design into the
legalizer's input (random initial placement, fillers, Nesterov or Adam over
the WA wirelength + electrostatic density objective).

`SkyPlace(design)` keeps the reference's setters (setTargetOverflow,
setInitLambda, setMaxPhiCoef, setRefHpwl, setInitGammaInv, setTargetDensity,
setSolver, setInitialPlaceMethod, setInitialPlaceOnly) and `run()` returns
the placed design; `GlobalPlacer` exposes the handle steps (initial place,
solver init, iterations, state) for step-by-step checks.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .design import Design, MarshalledDesign
from .lg import lib as _lglib

_f32p = C.POINTER(C.c_float)
_i32p = C.POINTER(C.c_int32)

SOLVER_NESTEROV, SOLVER_ADAM = 0, 1

EXPORTED = ["gp_last_error", "gp_default_params", "gp_create", "gp_initial_place", "gp_init_solver", "gp_iterate",
            "gp_solve", "gp_get_positions", "gp_set_positions", "gp_get_state", "gp_export", "gp_get_stats",
            "gp_destroy", "gp_place_design"]


class GpParams(C.Structure):
    _fields_ = [("max_opt_iter", C.c_int32), ("max_backtrack_iter", C.c_int32), ("init_lambda", C.c_float),
                ("init_gamma_inv_coef", C.c_float), ("target_overflow", C.c_float), ("min_phi_coef", C.c_float),
                ("max_phi_coef", C.c_float), ("min_precond", C.c_float), ("init_opt_coef", C.c_float),
                ("reference_hpwl", C.c_float), ("reference_hpwl_ratio", C.c_float),
                ("min_step_length", C.c_float), ("adam_alpha", C.c_float),
                ("adam_beta1", C.c_float), ("adam_beta2", C.c_float), ("adam_epsilon", C.c_float),
                ("target_density", C.c_float), ("scaling_factor", C.c_float), ("solver", C.c_int32), ("initial_place_only", C.c_int32),
                ("num_bin_x", C.c_int32), ("num_bin_y", C.c_int32), ("precision", C.c_int32),
                ("seed", C.c_uint64), ("log_freq", C.c_int32), ("verbose", C.c_int32)]


class GpTrace(C.Structure):
    _fields_ = [("cap", C.c_int32), ("hpwl", _f32p), ("overflow", _f32p), ("lambda_", _f32p), ("step", _f32p),
                ("backtracks", _i32p)]


class GpStats(C.Structure):
    _fields_ = [("num_movable", C.c_int32), ("num_fillers", C.c_int32), ("num_bin_x", C.c_int32),
                ("num_bin_y", C.c_int32), ("filler_width", C.c_float), ("filler_height", C.c_float),
                ("iterations", C.c_int32), ("converged", C.c_int32), ("evaluations", C.c_int32),
                ("hpwl", C.c_float), ("overflow", C.c_float), ("lambda_", C.c_float), ("gamma_inv", C.c_float),
                ("reference_hpwl", C.c_float), ("hpwl_initial_db", C.c_double), ("hpwl_db", C.c_double), ("t_init_s", C.c_double),
                ("t_solve_s", C.c_double)]


class PlacerError(RuntimeError):
    pass


_bound = None


def lib():
    global _bound
    L = _lglib()
    if _bound is None:
        L.gp_last_error.restype = C.c_char_p
        L.gp_default_params.argtypes = [C.POINTER(GpParams)]
        L.gp_default_params.restype = None
        L.gp_create.argtypes = [C.c_void_p, C.POINTER(GpParams), C.POINTER(C.c_void_p)]
        for name in ("gp_initial_place", "gp_init_solver", "gp_destroy"):
            getattr(L, name).argtypes = [C.c_void_p]
        L.gp_iterate.argtypes = [C.c_void_p, C.c_int32, _i32p]
        L.gp_solve.argtypes = [C.c_void_p, C.POINTER(GpTrace)]
        L.gp_get_positions.argtypes = [C.c_void_p, _f32p, _f32p]
        L.gp_set_positions.argtypes = [C.c_void_p, _f32p, _f32p]
        L.gp_get_state.argtypes = [C.c_void_p, _f32p, _f32p, _f32p, _f32p]
        L.gp_export.argtypes = [C.c_void_p, _i32p, _i32p]
        L.gp_get_stats.argtypes = [C.c_void_p, C.POINTER(GpStats)]
        L.gp_place_design.argtypes = [C.c_void_p, C.POINTER(GpParams), _i32p, _i32p, C.POINTER(GpStats),
                                      C.POINTER(GpTrace)]
        _bound = L
    return _bound


def _check(st):
    if st != 0:
        raise PlacerError(f"status {st}: {lib().gp_last_error().decode()}")


def default_params() -> GpParams:
    p = GpParams()
    lib().gp_default_params(C.byref(p))
    return p


def _fp(a):
    return a.ctypes.data_as(_f32p)


def _ip(a):
    return a.ctypes.data_as(_i32p)


class _TraceBuffers:
    def __init__(self, cap: int):
        self.cap = max(int(cap), 0)
        m = max(self.cap, 1)
        self.hpwl, self.overflow, self.lam, self.step = (np.zeros(m, np.float32) for _ in range(4))
        self.backtracks = np.zeros(m, np.int32)
        t = GpTrace()
        t.cap = self.cap
        t.hpwl, t.overflow, t.lambda_, t.step = _fp(self.hpwl), _fp(self.overflow), _fp(self.lam), _fp(self.step)
        t.backtracks = _ip(self.backtracks)
        self.struct = t

    def as_dict(self, iterations: int) -> dict:
        n = min(int(iterations), self.cap)
        return dict(hpwl=self.hpwl[:n].copy(), overflow=self.overflow[:n].copy(), lam=self.lam[:n].copy(),
                    step=self.step[:n].copy(), backtracks=self.backtracks[:n].copy())


def _stats_dict(s: GpStats) -> dict:
    out = {k: getattr(s, k) for k, _ in GpStats._fields_}
    out["lambda"] = out.pop("lambda_")
    return out


class GlobalPlacer:
    """The handle sequence of include/skyline_gp.h on one design."""

    def __init__(self, design: Design, params: GpParams | None = None):
        self.design = design
        self.params = params or default_params()
        self._md = MarshalledDesign(design)
        h = C.c_void_p()
        _check(lib().gp_create(C.byref(self._md.struct), C.byref(self.params), C.byref(h)))
        self.h = h
        st = self.stats()
        self.num_movable, self.num_fillers = st["num_movable"], st["num_fillers"]
        self.n = self.num_movable + self.num_fillers

    def initial_place(self):
        _check(lib().gp_initial_place(self.h))

    def init_solver(self):
        _check(lib().gp_init_solver(self.h))

    def iterate(self, n: int = 1) -> bool:
        done = np.zeros(1, np.int32)
        _check(lib().gp_iterate(self.h, int(n), _ip(done)))
        return bool(done[0])

    def solve(self, trace_cap: int = 0) -> dict:
        tb = _TraceBuffers(trace_cap)
        _check(lib().gp_solve(self.h, C.byref(tb.struct)))
        st = self.stats()
        st["trace"] = tb.as_dict(st["iterations"])
        return st

    def positions(self):
        cx, cy = np.zeros(self.n, np.float32), np.zeros(self.n, np.float32)
        _check(lib().gp_get_positions(self.h, _fp(cx), _fp(cy)))
        return cx, cy

    def set_positions(self, cx, cy):
        cx = np.ascontiguousarray(cx, np.float32)
        cy = np.ascontiguousarray(cy, np.float32)
        _check(lib().gp_set_positions(self.h, _fp(cx), _fp(cy)))

    def state(self) -> dict:
        sc = np.zeros(8, np.float32)
        v, p, g = (np.zeros(2 * self.n, np.float32) for _ in range(3))
        _check(lib().gp_get_state(self.h, _fp(sc), _fp(v), _fp(p), _fp(g)))
        keys = ("lambda", "gamma_inv", "step", "cur_a", "hpwl", "prev_hpwl", "overflow", "iter")
        out = {k: float(x) for k, x in zip(keys, sc)}
        out["iter"] = int(sc[7])
        out.update(cur_var=v, cur_pre=p, cur_grad=g)
        return out

    def export(self):
        lx, ly = np.zeros(self.design.num_insts, np.int32), np.zeros(self.design.num_insts, np.int32)
        _check(lib().gp_export(self.h, _ip(lx), _ip(ly)))
        return lx, ly

    def stats(self) -> dict:
        s = GpStats()
        _check(lib().gp_get_stats(self.h, C.byref(s)))
        return _stats_dict(s)

    def close(self):
        if getattr(self, "h", None):
            lib().gp_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def place_design(design: Design, params: GpParams | None = None, trace_cap: int = 0):
    """gp_place_design: (lx, ly, stats) with stats["trace"] of the loop."""
    params = params or default_params()
    md = MarshalledDesign(design)
    lx, ly = np.zeros(design.num_insts, np.int32), np.zeros(design.num_insts, np.int32)
    s = GpStats()
    tb = _TraceBuffers(trace_cap)
    _check(lib().gp_place_design(C.byref(md.struct), C.byref(params), _ip(lx), _ip(ly), C.byref(s),
                                 C.byref(tb.struct)))
    st = _stats_dict(s)
    st["trace"] = tb.as_dict(st["iterations"])
    return lx, ly, st


class SkyPlace:
    """skyplace::SkyPlace (SkyPlace.h:19-85): setters, then run() places the
    design and returns a copy at the global-placement positions."""

    def __init__(self, design: Design):
        self.design = design
        self.params = default_params()
        self.stats: dict | None = None

    def setTargetOverflow(self, v: float):
        self.params.target_overflow = float(v)

    def setInitLambda(self, v: float):
        self.params.init_lambda = float(v)

    def setMaxPhiCoef(self, v: float):
        self.params.max_phi_coef = float(v)

    def setRefHpwl(self, v: float):
        self.params.reference_hpwl = float(v)

    def setInitGammaInv(self, v: float):
        self.params.init_gamma_inv_coef = float(v)

    def setAdamAlpha(self, v: float):
        self.params.adam_alpha = float(v)

    def setAdamBeta1(self, v: float):
        self.params.adam_beta1 = float(v)

    def setAdamBeta2(self, v: float):
        self.params.adam_beta2 = float(v)

    def setTargetDensity(self, density: float):
        # SkyPlace::setTargetDensity (SkyPlace.cpp:99-114): out of (0, 100 %] -> 1.0
        pct = float(density) * 100.0
        self.params.target_density = 1.0 if (pct > 100.0 or pct <= 0.0) else float(density)

    def setSolver(self, name: str):
        if name in ("Adam", "adam"):
            self.params.solver = SOLVER_ADAM
        elif name in ("Nesterov", "nesterov"):
            self.params.solver = SOLVER_NESTEROV
        else:
            raise ValueError(f"undefined solver {name}")

    def setInitialPlaceMethod(self, method: str):
        if method != "random":
            raise ValueError(f"initial placement method {method!r} is not supported (random only)")

    def setInitialPlaceOnly(self):
        self.params.initial_place_only = 1

    def setNumBins(self, nbx: int, nby: int):
        self.params.num_bin_x, self.params.num_bin_y = int(nbx), int(nby)

    def run(self, trace_cap: int = 0) -> Design:
        lx, ly, self.stats = place_design(self.design, self.params, trace_cap)
        return self.design.copy_positions(lx, ly)
