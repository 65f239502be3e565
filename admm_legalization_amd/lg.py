"""Python front of the legalizer C ABI (include/skyline_lg.h).

`Legalizer` mirrors the reference's public class legalizer::Legalizer
(src/skyline/legalizer/include/legalizer/Legalizer.h:66-97): the same setters
with the same meanings, then `run()`.  `DeviceLegalizer` drives the device
layer (lg_create / lg_solve / lg_postprocess / lg_check / lg_copy_out) on
explicit grid and cell arrays.

All compute runs in libskyline_amd.so on the GPU; if the library is missing
or no device is present the calls raise -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .design import (SORT_ALGO, SORT_POLICY, Design, LgCellsSoA, LgGridDesc, LgParams, LgStats,
                     LgVioStat, MarshalledDesign, ResultBuffers, default_params, marshal_device_inputs)

_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "libskyline_amd.so")
_lib = None

_i32p = C.POINTER(C.c_int32)


EXPORTED = ["lg_default_params", "lg_last_error", "lg_run_design", "lg_import_and_preplace",
            "lg_import_and_preplace_device",
            "lg_create", "lg_solve", "lg_postprocess", "lg_check", "lg_copy_out", "lg_get_stats",
            "lg_get_pixel_state", "lg_get_ovf_trace", "lg_destroy", "lg_profile_partition",
            "lg_profile_read", "lg_use_graphs", "lg_debug_phase_ticks", "lg_srv_default_config", "lg_srv_create",
            "lg_srv_submit", "lg_srv_poll", "lg_srv_wait", "lg_srv_counters", "lg_srv_profile",
            "lg_srv_profile_read", "lg_srv_destroy"]


class LegalizerError(RuntimeError):
    pass


def lib():
    """Loads libskyline_amd.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise LegalizerError(f"{_LIB_PATH} not built; run admm_legalization_amd.build.build()")
        L = C.CDLL(_LIB_PATH)
        L.lg_last_error.restype = C.c_char_p
        L.lg_default_params.argtypes = [C.POINTER(LgParams)]
        L.lg_run_design.argtypes = [C.c_void_p, C.POINTER(LgParams), C.c_void_p]
        L.lg_import_and_preplace.argtypes = [C.c_void_p, C.POINTER(LgParams)] + [_i32p] * 8
        L.lg_import_and_preplace_device.argtypes = [C.c_void_p, C.POINTER(LgParams)] + [_i32p] * 5
        L.lg_create.argtypes = [C.POINTER(LgGridDesc), C.POINTER(LgCellsSoA), C.POINTER(LgParams),
                                C.POINTER(C.c_void_p)]
        L.lg_solve.argtypes = [C.c_void_p, C.c_void_p]
        L.lg_postprocess.argtypes = [C.c_void_p, C.c_void_p]
        L.lg_check.argtypes = [C.c_void_p, C.POINTER(LgVioStat)]
        L.lg_copy_out.argtypes = [C.c_void_p, _i32p, _i32p, _i32p]
        L.lg_get_stats.argtypes = [C.c_void_p, C.POINTER(LgStats)]
        L.lg_get_pixel_state.argtypes = [C.c_void_p, _i32p, C.POINTER(C.c_float)]
        L.lg_get_ovf_trace.argtypes = [C.c_void_p, _i32p, C.c_int32, _i32p]
        L.lg_destroy.argtypes = [C.c_void_p]
        L.lg_profile_partition.argtypes = [C.c_void_p, C.c_int]
        L.lg_use_graphs.argtypes = [C.c_void_p, C.c_int32]
        L.lg_profile_read.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int64)]
        L.lg_debug_phase_ticks.argtypes = [C.c_void_p, C.POINTER(C.c_uint64), C.c_int]
        _lib = L
    return _lib


def _check(st: int):
    if st != 0:
        raise LegalizerError(f"status {st}: {lib().lg_last_error().decode()}")


def _p(a, t=C.c_int32):
    return a.ctypes.data_as(C.POINTER(t))


class Legalizer:
    """legalizer::Legalizer (Legalizer.h:66-97) over lg_run_design."""

    def __init__(self, design: Design):
        self.design = design
        self.constraint = None
        self.size_file = None
        self.params = LgParams()
        lib().lg_default_params(C.byref(self.params))
        self.result = None

    # -- setters of the reference (Legalizer.cpp:95-195) ---------------------
    def setICCAD17Constraint(self, path: str): self.constraint = path or None  # Legalizer.h:74
    def setSizeByFile(self, path: str): self.size_file = path or None  # Legalizer.h:75
    def setTechPenalty(self, v: float): self.params.tech_penalty = float(v)
    def setStandardAdmm(self, v: bool): self.params.standard_admm = int(bool(v))
    def setQpRefine(self, v: bool): self.params.qp_admm_on = int(bool(v))
    def setCoeffToInitRho(self, v: float): self.params.coeff_to_init_rho = float(v)
    def setCoeffToInitDual(self, v: float): self.params.coeff_to_init_dual = float(v)
    def setMaxAdmmIter(self, v: int): self.params.max_admm_iter = int(v)
    def setFlagICCAD2017(self, v: bool): self.params.flag_iccad17 = int(bool(v))
    def setAdaptiveSearch(self, v: bool): self.params.use_adaptive_search = int(bool(v))
    def setMinIterToCallAdaptiveSearch(self, v: int): self.params.min_iter_to_call_adaptive_search = int(v)
    def setOvfRatioToCallAdaptiveSearch(self, v: float): self.params.ovf_ratio_to_call_adaptive_search = float(v)
    def setCoeffToNeglectDisplace(self, v: float): self.params.coeff_to_neglect_displace = float(v)
    def setRandomSortingSeed(self, v: int): self.params.random_sorting_seed = int(v)

    def setRandomSorting(self, v: bool):
        self.params.sorting_policy = SORT_POLICY["RANDOM" if v else "OVERFLOW"]

    def setSortingPolicy(self, name: str):  # unknown names -> OVERFLOW (Legalizer.cpp:161-172)
        self.params.sorting_policy = SORT_POLICY.get(name, SORT_POLICY["OVERFLOW"])

    def setSortingAlgorithm(self, name: str):  # unknown names -> M_BITONIC (Legalizer.cpp:174-189)
        self.params.sorting_algorithm = SORT_ALGO.get(name, SORT_ALGO["M_BITONIC"])

    def run(self, trace_cap: int = 0):
        """Legalizes; returns (lx, ly, flipped, stats) per instance (dbu)."""
        md = MarshalledDesign(self.design, self.size_file, self.constraint)
        rb = ResultBuffers(self.design.num_insts, trace_cap)
        _check(lib().lg_run_design(C.byref(md.struct), C.byref(self.params), C.byref(rb.struct)))
        self.result = (rb.lx, rb.ly, rb.flipped, rb.stats_dict())
        return self.result


def run_design(design: Design, params: LgParams | None = None, trace_cap: int = 0, size_file: str | None = None,
               constraint_file: str | None = None):
    lg = Legalizer(design)
    if params is not None:
        lg.params = params
    lg.setSizeByFile(size_file)
    lg.setICCAD17Constraint(constraint_file)
    return lg.run(trace_cap)


def import_and_preplace(design: Design, params: LgParams | None = None, size_file: str | None = None) -> dict:
    """Host import + preplace: pixel maps (x-major), region bounds and the
    preplaced movable-cell positions (dbu).  Pure host code (no device)."""
    params = params or default_params()
    md = MarshalledDesign(design, size_file)
    shape = np.zeros(4, np.int32)
    null = C.cast(None, _i32p)
    _check(lib().lg_import_and_preplace(C.byref(md.struct), C.byref(params), _p(shape), null, null,
                                        null, null, null, null, null))
    xs, ys, n, nr = (int(v) for v in shape)
    P = xs * ys
    valid, group, region = (np.zeros(P, np.int32) for _ in range(3))
    vdd = np.zeros(ys, np.int32)
    rb = np.zeros(max(4 * nr, 1), np.int32)
    lx, ly = np.zeros(max(n, 1), np.int32), np.zeros(max(n, 1), np.int32)
    _check(lib().lg_import_and_preplace(C.byref(md.struct), C.byref(params), _p(shape), _p(valid),
                                        _p(group), _p(region), _p(vdd), _p(rb), _p(lx), _p(ly)))
    rb = rb[:4 * nr].reshape(nr, 4)
    return dict(x_size=xs, y_size=ys, valid=valid, group=group, region=region, vdd_up=vdd,
                region_bounds=rb, cell_lx=lx[:n], cell_ly=ly[:n])


def import_and_preplace_device(design: Design, params: LgParams | None = None) -> dict:
    """lg_import_and_preplace_device: the pixel grid built and the cells
    preplaced on the GPU (valid / group x-major, group -1 on invalid pixels)."""
    params = params or default_params()
    md = MarshalledDesign(design)
    shape = np.zeros(4, np.int32)
    null = C.cast(None, _i32p)
    _check(lib().lg_import_and_preplace_device(C.byref(md.struct), C.byref(params), _p(shape), null, null, null,
                                               null))
    xs, ys, n, _ = (int(v) for v in shape)
    valid, group = np.zeros(xs * ys, np.int32), np.zeros(xs * ys, np.int32)
    lx, ly = np.zeros(max(n, 1), np.int32), np.zeros(max(n, 1), np.int32)
    _check(lib().lg_import_and_preplace_device(C.byref(md.struct), C.byref(params), _p(shape), _p(valid),
                                               _p(group), _p(lx), _p(ly)))
    return dict(x_size=xs, y_size=ys, valid=valid, group=group, cell_lx=lx[:n], cell_ly=ly[:n])


def device_inputs(design: Design, params: LgParams | None = None):
    """Grid / cell dicts for DeviceLegalizer from the host import + preplace,
    and the parameters with y_hint adjusted as Legalizer::adjustYHint does
    (Legalizer.cpp:230-247)."""
    params = params or default_params()
    pre = import_and_preplace(design, params)
    mv = design.movable_mask()
    sw, rh = design.site_width, design.row_height
    h_grid = (design.inst_dy[mv] // rh).astype(np.int32)
    p2 = LgParams.from_buffer_copy(params)
    if h_grid.size:
        p2.y_hint = max(int(p2.y_hint), 3 * int(design.inst_dy[mv].max() // rh))
    rb = pre["region_bounds"]
    grid = dict(x_size=pre["x_size"], y_size=pre["y_size"], site_width=sw, row_height=rh,
                core_lx=design.core[0], core_ly=design.core[1], valid=pre["valid"],
                group=pre["group"], region=pre["region"], vdd_up=pre["vdd_up"],
                region_x_min=rb[:, 0].copy(), region_y_min=rb[:, 1].copy(),
                region_x_max=rb[:, 2].copy(), region_y_max=rb[:, 3].copy())
    cells = dict(orig_lx_dbu=design.inst_lx[mv], orig_ly_dbu=design.inst_ly[mv],
                 lx_dbu=pre["cell_lx"], ly_dbu=pre["cell_ly"],
                 w_grid=(design.inst_dx[mv] // sw).astype(np.int32), h_grid=h_grid,
                 group=design.inst_group[mv].astype(np.int32),
                 ground_at_bottom=design.inst_ground_at_bottom[mv].astype(np.int32))
    return grid, cells, p2


class DeviceLegalizer:
    """The device layer on explicit arrays (CudaDatabase + CudaAdmmLegalizer +
    CudaPostProcess of the reference).  `grid` and `cells` are dicts of numpy
    arrays with the fields of LgGridDesc / LgCellsSoA."""

    def __init__(self, grid: dict, cells: dict, params: LgParams):
        g, c, self._keep = marshal_device_inputs(grid, cells)
        self.num_cells = c.num_cells
        self.P = int(grid["x_size"]) * int(grid["y_size"])
        self.params = params
        h = C.c_void_p()
        _check(lib().lg_create(C.byref(g), C.byref(c), C.byref(params), C.byref(h)))
        self.h = h

    def solve(self, stream: int | None = None):
        """Runs the ADMM loop from the positions given at construction
        (repeatable).  `stream` is a raw hipStream_t handle or None for the
        handle's own stream."""
        _check(lib().lg_solve(self.h, stream))

    def postprocess(self, stream: int | None = None):
        _check(lib().lg_postprocess(self.h, stream))

    def use_graphs(self, enable: bool = True):
        """HIP graphs for the ADMM iterations (off by default; same results)."""
        _check(lib().lg_use_graphs(self.h, int(bool(enable))))

    def profile_partition(self, enable: bool = True):
        _check(lib().lg_profile_partition(self.h, int(bool(enable))))

    def profile_read(self):
        """(ms[2], launches[2]) of the partition kernel per LDS class."""
        ms = (C.c_double * 2)()
        n = (C.c_int64 * 2)()
        _check(lib().lg_profile_read(self.h, ms, n))
        return list(ms), list(n)

    def check(self) -> dict:
        v = LgVioStat()
        _check(lib().lg_check(self.h, C.byref(v)))
        return {k: getattr(v, k) for k, _ in LgVioStat._fields_}

    def positions(self):
        lx = np.zeros(self.num_cells, np.int32)
        ly = np.zeros(self.num_cells, np.int32)
        rot = np.zeros(self.num_cells, np.int32)
        _check(lib().lg_copy_out(self.h, _p(lx), _p(ly), _p(rot)))
        return lx, ly, rot

    def pixel_state(self):
        u = np.zeros(self.P, np.int32)
        d = np.zeros(self.P, np.float32)
        _check(lib().lg_get_pixel_state(self.h, _p(u), _p(d, C.c_float)))
        return u, d

    def stats(self) -> dict:
        s = LgStats()
        _check(lib().lg_get_stats(self.h, C.byref(s)))
        out = {k: getattr(s, k) for k, _ in LgStats._fields_ if k != "vio"}
        out["vio"] = {k: getattr(s.vio, k) for k, _ in LgVioStat._fields_}
        return out

    def trace(self) -> np.ndarray:
        n = C.c_int32()
        _check(lib().lg_get_ovf_trace(self.h, None, 0, C.byref(n)))
        t = np.zeros(max(n.value, 1), np.int32)
        _check(lib().lg_get_ovf_trace(self.h, _p(t), n.value, C.byref(n)))
        return t[:n.value]

    def close(self):
        if self.h:
            lib().lg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
