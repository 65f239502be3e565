"""In-memory design (the legalizer's view of a placed netlist) and its C ABI.

`Design` holds the subset of the reference's ``db::dbDatabase`` that
``legalizer::Legalizer`` reads (reference: src/skyline/legalizer/src/
DatabaseIO.cpp:31-379, Grid.cpp:24-67): the core box, the placement rows, the
instances with their global-placement positions and the fence groups.  All
coordinates are integer database units (dbu).

The ctypes mirrors of ``include/skyline_lg.h`` live here as well so that every
caller of the C ABI marshals structures the same way.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

ORIENT_N, ORIENT_S, ORIENT_W, ORIENT_E, ORIENT_FN, ORIENT_FS, ORIENT_FW, ORIENT_FE = range(8)

SORT_POLICY = {"RANDOM": 0, "SIZE": 1, "OVERFLOW": 2}
SORT_ALGO = {"S_MERGE": 0, "M_MERGE": 1, "S_BITONIC": 2, "M_BITONIC": 3, "CUB": 4}


@dataclass
class Design:
    """Rows, instances and fences of one design, all int32 dbu arrays."""

    core: tuple  # (lx, ly, ux, uy)
    row_lx: np.ndarray
    row_ly: np.ndarray
    row_dx: np.ndarray
    row_dy: np.ndarray
    row_site_width: np.ndarray
    row_orient: np.ndarray
    inst_lx: np.ndarray
    inst_ly: np.ndarray
    inst_dx: np.ndarray
    inst_dy: np.ndarray
    inst_fixed: np.ndarray
    inst_macro: np.ndarray
    inst_group: np.ndarray
    inst_ground_at_bottom: np.ndarray
    num_groups: int = 0
    box_group: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    box_lx: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    box_ly: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    box_ux: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    box_uy: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    # optional netlist (HPWL): pins of net k are net_start[k] .. net_start[k+1]-1,
    # at (pin_dx, pin_dy) from instance pin_inst's lower-left (orientation N),
    # or absolute for pin_inst = -1 (IO pins)
    net_start: np.ndarray = field(default_factory=lambda: np.zeros(1, np.int32))
    pin_inst: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    pin_dx: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    pin_dy: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    inst_name: list = None  # optional instance names (size files)
    dbu_per_micron: int = 1000
    meta: dict = field(default_factory=dict)
    # optional ICCAD 2017 technology: the LgDesign technology arrays by name
    # (inst_master, master_ledge, ..., es_spacing_um; include/skyline_lg.h)
    tech: dict = None

    @property
    def num_insts(self) -> int:
        return int(self.inst_lx.shape[0])

    @property
    def site_width(self) -> int:
        return int(self.row_site_width[-1])

    @property
    def row_height(self) -> int:
        return int(self.row_dy.min())

    def movable_mask(self) -> np.ndarray:
        """Instances the legalizer moves (DatabaseIO.cpp:100-149): non-macro,
        non-fixed, nonzero width and at least one row tall."""
        ok = (self.inst_dx != 0) & (self.inst_dy >= int(self.row_dy[-1]))
        return ok & (self.inst_macro == 0) & (self.inst_fixed == 0)

    def copy_positions(self, lx, ly) -> "Design":
        d = Design(**{k: getattr(self, k) for k in self.__dataclass_fields__})
        d.inst_lx = np.ascontiguousarray(lx, dtype=np.int32)
        d.inst_ly = np.ascontiguousarray(ly, dtype=np.int32)
        return d

    # -- persistence (golden fixtures) ------------------------------------
    def to_npz(self, path: str) -> None:
        arrays = {k: np.asarray(getattr(self, k)) for k in self.__dataclass_fields__
                  if isinstance(getattr(self, k), np.ndarray)}
        np.savez_compressed(path, core=np.asarray(self.core, np.int32),
                            num_groups=np.int32(self.num_groups),
                            dbu_per_micron=np.int32(self.dbu_per_micron), **arrays)

    def save_lgd(self, path: str) -> None:
        """The design in the LGD1 binary format of the command driver
        (admm_legalization_amd/csrc/lgd_file.hpp)."""
        i32 = lambda a: np.ascontiguousarray(a, np.int32).tobytes()
        u8 = lambda a: np.ascontiguousarray(a, np.uint8).tobytes()
        n, nr = self.num_insts, int(self.row_lx.shape[0])
        nb = int(np.asarray(self.box_group).shape[0])
        ns = np.asarray(self.net_start, np.int32)
        with open(path, "wb") as f:
            f.write(b"LGD1")
            f.write(i32([self.dbu_per_micron, *self.core, nr]))
            for a in (self.row_lx, self.row_ly, self.row_dx, self.row_dy, self.row_site_width, self.row_orient):
                f.write(i32(a))
            f.write(i32([n]))
            for a in (self.inst_lx, self.inst_ly, self.inst_dx, self.inst_dy):
                f.write(i32(a))
            f.write(u8(self.inst_fixed))
            f.write(u8(self.inst_macro))
            f.write(i32(self.inst_group))
            f.write(u8(self.inst_ground_at_bottom))
            f.write(i32([self.num_groups, nb]))
            for a in (self.box_group, self.box_lx, self.box_ly, self.box_ux, self.box_uy):
                f.write(i32(a))
            f.write(i32([max(len(ns) - 1, 0)]))
            f.write(i32(ns if len(ns) else [0]))
            f.write(i32([len(self.pin_inst)]))
            for a in (self.pin_inst, self.pin_dx, self.pin_dy):
                f.write(i32(a))
            f.write(i32([1 if self.inst_name is not None else 0]))
            if self.inst_name is not None:
                for name in self.inst_name:
                    b = name.encode()
                    f.write(i32([len(b)]))
                    f.write(b)

    @staticmethod
    def from_npz(path: str) -> "Design":
        z = np.load(path, allow_pickle=False)
        kw = {k: z[k] for k in z.files if k not in ("core", "num_groups", "dbu_per_micron")}
        return Design(core=tuple(int(v) for v in z["core"]), num_groups=int(z["num_groups"]),
                      dbu_per_micron=int(z["dbu_per_micron"]), **kw)


# ---------------------------------------------------------------------------
# ctypes mirrors of include/skyline_lg.h
# ---------------------------------------------------------------------------
_i32p = C.POINTER(C.c_int32)
_u8p = C.POINTER(C.c_uint8)


class LgParams(C.Structure):
    _fields_ = [
        ("flag_iccad17", C.c_int32), ("standard_admm", C.c_int32), ("log_freq", C.c_int32),
        ("num_cpu_threads", C.c_int32), ("max_disp_in_row_height", C.c_int32),
        ("max_admm_iter", C.c_int32), ("admm_terminate_thr", C.c_int32),
        ("x_hint", C.c_int32), ("y_hint", C.c_int32), ("iter_update_partition", C.c_int32),
        ("max_disp_coeff", C.c_float), ("tech_penalty", C.c_float),
        ("edge_spacing_penalty", C.c_float),
        ("iter_threshold_dual_update", C.c_int32), ("iter_update_dual_step", C.c_int32),
        ("iter_perturbation_on", C.c_int32), ("coeff_to_init_rho", C.c_float),
        ("coeff_to_init_dual", C.c_float), ("coeff_to_update_dual_step", C.c_float),
        ("qp_admm_on", C.c_int32), ("qp_admm_dp_site_align", C.c_int32),
        ("qp_admm_max_iter", C.c_int32), ("qp_admm_min_iter", C.c_int32),
        ("qp_admm_rho", C.c_float),
        ("use_adaptive_search", C.c_int32), ("stagnation_window", C.c_int32),
        ("min_iter_to_call_adaptive_search", C.c_int32),
        ("cooltime_to_call_adaptive_search", C.c_int32), ("stagnation_threshold", C.c_float),
        ("ovf_ratio_to_call_adaptive_search", C.c_float),
        ("coeff_to_neglect_displace", C.c_float),
        ("random_sorting_seed", C.c_int32), ("sorting_policy", C.c_int32),
        ("sorting_algorithm", C.c_int32), ("verbose", C.c_int32),
    ]


def default_params() -> LgParams:
    """Defaults of Legalizer::initParameters (Legalizer.cpp:47-93)."""
    p = LgParams()
    p.flag_iccad17 = 0
    p.standard_admm = 0
    p.log_freq = 100
    p.num_cpu_threads = 16
    p.max_disp_in_row_height = 3
    p.max_admm_iter = 25000
    p.admm_terminate_thr = 0
    p.x_hint = 250
    p.y_hint = 12
    p.iter_update_partition = 1
    p.max_disp_coeff = 40.0
    p.tech_penalty = 1.0
    p.edge_spacing_penalty = 10000.0
    p.iter_threshold_dual_update = 300
    p.iter_update_dual_step = 100
    p.iter_perturbation_on = 2
    p.coeff_to_init_rho = 3.0
    p.coeff_to_init_dual = 0.5
    p.coeff_to_update_dual_step = 0.2
    p.qp_admm_on = 0
    p.qp_admm_dp_site_align = 1
    p.qp_admm_max_iter = 10000
    p.qp_admm_min_iter = 5
    p.qp_admm_rho = 2.0
    p.use_adaptive_search = 0
    p.stagnation_window = 10
    p.stagnation_threshold = 0.01
    p.min_iter_to_call_adaptive_search = 300
    p.cooltime_to_call_adaptive_search = 100
    p.ovf_ratio_to_call_adaptive_search = 1.05
    p.coeff_to_neglect_displace = 20.0
    p.random_sorting_seed = 0
    p.sorting_policy = SORT_POLICY["OVERFLOW"]
    p.sorting_algorithm = SORT_ALGO["M_BITONIC"]
    p.verbose = 0
    return p


class LgDesign(C.Structure):
    _fields_ = [
        ("dbu_per_micron", C.c_int32),
        ("core_lx", C.c_int32), ("core_ly", C.c_int32), ("core_ux", C.c_int32), ("core_uy", C.c_int32),
        ("num_rows", C.c_int32),
        ("row_lx", _i32p), ("row_ly", _i32p), ("row_dx", _i32p), ("row_dy", _i32p),
        ("row_site_width", _i32p), ("row_orient", _i32p),
        ("num_insts", C.c_int32),
        ("inst_lx", _i32p), ("inst_ly", _i32p), ("inst_dx", _i32p), ("inst_dy", _i32p),
        ("inst_fixed", _u8p), ("inst_macro", _u8p), ("inst_group", _i32p),
        ("inst_ground_at_bottom", _u8p),
        ("num_groups", C.c_int32), ("num_region_boxes", C.c_int32),
        ("box_group", _i32p), ("box_lx", _i32p), ("box_ly", _i32p), ("box_ux", _i32p),
        ("box_uy", _i32p),
        ("num_nets", C.c_int32), ("net_start", _i32p), ("pin_inst", _i32p), ("pin_dx", _i32p),
        ("pin_dy", _i32p), ("inst_name", C.POINTER(C.c_char_p)), ("size_file", C.c_char_p),
        ("constraint_file", C.c_char_p),
        ("num_masters", C.c_int32), ("inst_master", _i32p), ("master_ledge", _i32p), ("master_redge", _i32p),
        ("master_is_block", _u8p), ("master_pin_start", _i32p), ("mpin_layer", _i32p), ("mpin_lx", _i32p),
        ("mpin_ly", _i32p), ("mpin_ux", _i32p), ("mpin_uy", _i32p),
        ("num_pwires", C.c_int32), ("pwire_layer", _i32p), ("pwire_x1", _i32p), ("pwire_y1", _i32p),
        ("pwire_x2", _i32p), ("pwire_y2", _i32p), ("pwire_width", _i32p),
        ("num_edge_spacings", C.c_int32), ("es_type1", _i32p), ("es_type2", _i32p),
        ("es_spacing_um", C.POINTER(C.c_double)),
    ]

# design-level technology arrays (Design.tech keys) and their C types
TECH_FIELDS = {
    "inst_master": C.c_int32, "master_ledge": C.c_int32, "master_redge": C.c_int32,
    "master_is_block": C.c_uint8, "master_pin_start": C.c_int32, "mpin_layer": C.c_int32,
    "mpin_lx": C.c_int32, "mpin_ly": C.c_int32, "mpin_ux": C.c_int32, "mpin_uy": C.c_int32,
    "pwire_layer": C.c_int32, "pwire_x1": C.c_int32, "pwire_y1": C.c_int32, "pwire_x2": C.c_int32,
    "pwire_y2": C.c_int32, "pwire_width": C.c_int32, "es_type1": C.c_int32, "es_type2": C.c_int32,
    "es_spacing_um": C.c_double,
}


class LgVioStat(C.Structure):
    _fields_ = [("total_ovf", C.c_int32), ("num_row_orient_vio", C.c_int32),
                ("num_pin_short_vio", C.c_int32), ("num_pin_access_vio", C.c_int32),
                ("num_edge_spacing_vio", C.c_int32), ("num_site_vio", C.c_int32),
                ("num_fence_vio", C.c_int32), ("num_core_vio", C.c_int32), ("num_rail_vio", C.c_int32)]


STRICT_VIO = ("total_ovf", "num_row_orient_vio", "num_site_vio", "num_fence_vio", "num_core_vio",
              "num_rail_vio")


def strictly_legal(vio: dict) -> bool:
    """No overlap, no orientation or rail error, every footprint on valid
    sites of the cell's own fence group and inside the core."""
    return all(int(vio[k]) == 0 for k in STRICT_VIO)


class LgStats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int32), ("converged", C.c_int32),
        ("final_overflow", C.c_int64), ("total_disp_dbu", C.c_int64), ("max_disp_dbu", C.c_int64),
        ("total_disp_um", C.c_double), ("avg_disp_um", C.c_double), ("max_disp_um", C.c_double),
        ("rho", C.c_float), ("initial_dual", C.c_float), ("num_cells", C.c_int32),
        ("num_adaptive_calls", C.c_int32), ("vio", LgVioStat),
        ("t_import_s", C.c_double), ("t_preplace_s", C.c_double), ("t_admm_s", C.c_double),
        ("t_post_s", C.c_double), ("t_total_s", C.c_double),
        ("hpwl_before_um", C.c_double), ("hpwl_after_um", C.c_double), ("delta_hpwl_pct", C.c_double),
        ("max_util_constraint", C.c_double), ("max_disp_constraint", C.c_double),
        ("iccad17_s_am", C.c_double), ("iccad17_m_max", C.c_double), ("iccad17_score", C.c_double),
    ]


class LgResult(C.Structure):
    _fields_ = [("out_lx", _i32p), ("out_ly", _i32p), ("out_flipped", _u8p),
                ("ovf_trace", _i32p), ("ovf_trace_cap", C.c_int32), ("stats", LgStats)]


class LgGridDesc(C.Structure):
    _fields_ = [("x_size", C.c_int32), ("y_size", C.c_int32), ("site_width", C.c_int32),
                ("row_height", C.c_int32), ("core_lx", C.c_int32), ("core_ly", C.c_int32),
                ("pixel_valid", _i32p), ("pixel_group", _i32p), ("pixel_region", _i32p),
                ("row_vdd_up", _i32p), ("num_regions", C.c_int32), ("region_x_min", _i32p),
                ("region_y_min", _i32p), ("region_x_max", _i32p), ("region_y_max", _i32p),
                ("pixel_power_layer", _i32p), ("num_edge_rules", C.c_int32), ("edge_spacing_rules", _i32p)]


class LgCellsSoA(C.Structure):
    _fields_ = [("num_cells", C.c_int32), ("orig_lx_dbu", _i32p), ("orig_ly_dbu", _i32p),
                ("lx_dbu", _i32p), ("ly_dbu", _i32p), ("w_grid", _i32p), ("h_grid", _i32p),
                ("group", _i32p), ("ground_at_bottom", _i32p),
                ("ledge_type", _i32p), ("redge_type", _i32p), ("macro_id", _i32p), ("num_macros", C.c_int32),
                ("macro_pin_start", _i32p), ("pin_bgn", _i32p), ("pin_end", _i32p), ("pin_layer", _i32p)]


def marshal_device_inputs(grid: dict, cells: dict):
    """LgGridDesc / LgCellsSoA over the arrays of `grid` and `cells` (dicts of
    numpy arrays with the fields of the two structs; the technology keys
    power_layer, edge_rules, ledge_type, redge_type, macro_id,
    macro_pin_start, pin_bgn, pin_end, pin_layer are optional).  The third
    value keeps the int32 copies alive for as long as the structs are used."""
    keep = []

    def arr(a):
        a = np.ascontiguousarray(a, np.int32)
        keep.append(a)
        return a.ctypes.data_as(_i32p)

    g = LgGridDesc()
    for k in ("x_size", "y_size", "site_width", "row_height", "core_lx", "core_ly"):
        setattr(g, k, int(grid[k]))
    g.pixel_valid = arr(grid["valid"])
    g.pixel_group = arr(grid["group"])
    g.pixel_region = arr(grid.get("region", np.full(int(grid["x_size"]) * int(grid["y_size"]), -1)))
    g.row_vdd_up = arr(grid["vdd_up"])
    g.num_regions = len(grid.get("region_x_min", []))
    for k in ("region_x_min", "region_y_min", "region_x_max", "region_y_max"):
        setattr(g, k, arr(grid.get(k, np.zeros(0, np.int32))))
    if grid.get("power_layer") is not None:
        g.pixel_power_layer = arr(grid["power_layer"])
    if grid.get("edge_rules") is not None:
        g.num_edge_rules = len(grid["edge_rules"])
        g.edge_spacing_rules = arr(grid["edge_rules"])
    c = LgCellsSoA()
    c.num_cells = int(len(cells["w_grid"]))
    for k in ("orig_lx_dbu", "orig_ly_dbu", "lx_dbu", "ly_dbu", "w_grid", "h_grid", "group",
              "ground_at_bottom"):
        setattr(c, k, arr(cells[k]))
    for k in ("ledge_type", "redge_type", "macro_id", "macro_pin_start", "pin_bgn", "pin_end", "pin_layer"):
        if cells.get(k) is not None:
            setattr(c, k, arr(cells[k]))
    if cells.get("macro_pin_start") is not None:
        c.num_macros = len(cells["macro_pin_start"]) - 1
    return g, c, keep



def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


class MarshalledDesign:
    """Keeps contiguous copies alive while an LgDesign points into them."""

    def __init__(self, d: Design, size_file: str | None = None, constraint_file: str | None = None):
        conv = lambda a, t: np.ascontiguousarray(a, dtype=t)
        self.keep = {}
        s = LgDesign()
        s.dbu_per_micron = int(d.dbu_per_micron)
        s.core_lx, s.core_ly, s.core_ux, s.core_uy = (int(v) for v in d.core)
        s.num_rows = int(d.row_lx.shape[0])
        for name in ("row_lx", "row_ly", "row_dx", "row_dy", "row_site_width", "row_orient",
                     "inst_lx", "inst_ly", "inst_dx", "inst_dy", "inst_group",
                     "box_group", "box_lx", "box_ly", "box_ux", "box_uy"):
            a = conv(getattr(d, name), np.int32)
            self.keep[name] = a
            setattr(s, name, _ptr(a, C.c_int32))
        for name in ("inst_fixed", "inst_macro", "inst_ground_at_bottom"):
            a = conv(getattr(d, name), np.uint8)
            self.keep[name] = a
            setattr(s, name, _ptr(a, C.c_uint8))
        s.num_insts = d.num_insts
        s.num_groups = int(d.num_groups)
        s.num_region_boxes = int(np.asarray(d.box_group).shape[0])
        ns = np.asarray(d.net_start)
        s.num_nets = max(int(ns.shape[0]) - 1, 0)
        for name in ("net_start", "pin_inst", "pin_dx", "pin_dy"):
            a = conv(getattr(d, name), np.int32)
            self.keep[name] = a
            setattr(s, name, _ptr(a, C.c_int32))
        if d.inst_name is not None:
            names = (C.c_char_p * d.num_insts)(*[n.encode() for n in d.inst_name])
            self.keep["inst_name"] = names
            s.inst_name = C.cast(names, C.POINTER(C.c_char_p))
        self.keep["files"] = (size_file.encode() if size_file else None,
                              constraint_file.encode() if constraint_file else None)
        s.size_file, s.constraint_file = self.keep["files"]
        if d.tech:
            t = d.tech
            for name, ct in TECH_FIELDS.items():
                a = conv(t[name], np.float64 if ct is C.c_double else (np.uint8 if ct is C.c_uint8 else np.int32))
                self.keep["tech_" + name] = a
                setattr(s, name, _ptr(a, ct))
            s.num_masters = len(t["master_ledge"])
            s.num_pwires = len(t["pwire_layer"])
            s.num_edge_spacings = len(t["es_type1"])
        self.struct = s


class ResultBuffers:
    """Caller-owned output arrays of lg_run_design."""

    def __init__(self, num_insts: int, trace_cap: int = 0):
        self.lx = np.zeros(num_insts, np.int32)
        self.ly = np.zeros(num_insts, np.int32)
        self.flipped = np.zeros(num_insts, np.uint8)
        self.trace = np.zeros(max(trace_cap, 1), np.int32)
        r = LgResult()
        r.out_lx = _ptr(self.lx, C.c_int32)
        r.out_ly = _ptr(self.ly, C.c_int32)
        r.out_flipped = _ptr(self.flipped, C.c_uint8)
        r.ovf_trace = _ptr(self.trace, C.c_int32)
        r.ovf_trace_cap = int(trace_cap)
        self.struct = r

    def stats_dict(self) -> dict:
        st = self.struct.stats
        out = {k: getattr(st, k) for k, _ in LgStats._fields_ if k != "vio"}
        out["vio"] = {k: getattr(st.vio, k) for k, _ in LgVioStat._fields_}
        n = min(int(st.iterations), int(self.struct.ovf_trace_cap))
        out["ovf_trace"] = self.trace[:n].copy()
        return out
