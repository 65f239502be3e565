"""LEF / DEF reading and DEF writing through the library's C ABI
(include/skyline_lefdef.h): the design-level input of the legalizer as the
reference's db::dbDatabase::readLef / readDef / writeDef provide it
(src/skyline/database/db/dbDatabase.h:29-34).

    db = LefDef(["tech.lef", "cells.lef"], "placed.def")
    d = db.design()                      # a Design, ICCAD 2017 technology included
    lx, ly, fl, st = lg.run_design(d)
    db.write_def("legalized.def", lx, ly, fl)
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .design import Design, LgDesign
from .lg import LegalizerError, lib

_i32p = C.POINTER(C.c_int32)
_u8p = C.POINTER(C.c_uint8)

# the entry points of include/skyline_lefdef.h
EXPORTED = ["ld_last_error", "ld_create", "ld_read_lef", "ld_read_def", "ld_get_design", "ld_num_insts",
            "ld_inst_name", "ld_write_def", "ld_destroy"]


def _setup(L):
    if getattr(L, "_ld_ready", False):
        return L
    L.ld_last_error.restype = C.c_char_p
    L.ld_create.argtypes = [C.POINTER(C.c_void_p)]
    L.ld_read_lef.argtypes = [C.c_void_p, C.c_char_p]
    L.ld_read_def.argtypes = [C.c_void_p, C.c_char_p]
    L.ld_get_design.argtypes = [C.c_void_p, C.POINTER(LgDesign)]
    L.ld_num_insts.argtypes = [C.c_void_p]
    L.ld_num_insts.restype = C.c_int32
    L.ld_inst_name.argtypes = [C.c_void_p, C.c_int32]
    L.ld_inst_name.restype = C.c_char_p
    L.ld_write_def.argtypes = [C.c_void_p, C.c_char_p, _i32p, _i32p, _u8p]
    L.ld_destroy.argtypes = [C.c_void_p]
    L._ld_ready = True
    return L


def _check(st):
    if st != 0:
        raise LegalizerError(f"status {st}: {lib().ld_last_error().decode()}")


def _arr(p, n, dtype=np.int32):
    if n <= 0 or not p:
        return np.zeros(0, dtype)
    return np.ctypeslib.as_array(p, shape=(n,)).astype(dtype, copy=True)


class LefDef:
    """A LEF / DEF database: LEF files in order, then one DEF."""

    def __init__(self, lef_paths, def_path: str | None = None):
        L = _setup(lib())
        h = C.c_void_p()
        _check(L.ld_create(C.byref(h)))
        self.h = h
        for p in ([lef_paths] if isinstance(lef_paths, str) else lef_paths):
            _check(L.ld_read_lef(self.h, str(p).encode()))
        if def_path:
            self.read_def(def_path)

    def read_def(self, path: str):
        _check(lib().ld_read_def(self.h, str(path).encode()))

    def design(self) -> Design:
        """The design read, as the legalizer's Design (with its netlist,
        instance names and ICCAD 2017 technology)."""
        L = lib()
        s = LgDesign()
        _check(L.ld_get_design(self.h, C.byref(s)))
        n, nr = s.num_insts, s.num_rows
        nb, nn = s.num_region_boxes, s.num_nets
        ns = _arr(s.net_start, nn + 1)
        npins = int(ns[-1]) if len(ns) else 0
        M = s.num_masters
        mps = _arr(s.master_pin_start, M + 1)
        nmp = int(mps[-1]) if M else 0
        nw, ne = s.num_pwires, s.num_edge_spacings
        tech = None
        if M > 0:
            tech = dict(inst_master=_arr(s.inst_master, n), master_ledge=_arr(s.master_ledge, M),
                        master_redge=_arr(s.master_redge, M),
                        master_is_block=_arr(s.master_is_block, M, np.uint8), master_pin_start=mps,
                        mpin_layer=_arr(s.mpin_layer, nmp), mpin_lx=_arr(s.mpin_lx, nmp),
                        mpin_ly=_arr(s.mpin_ly, nmp), mpin_ux=_arr(s.mpin_ux, nmp), mpin_uy=_arr(s.mpin_uy, nmp),
                        pwire_layer=_arr(s.pwire_layer, nw), pwire_x1=_arr(s.pwire_x1, nw),
                        pwire_y1=_arr(s.pwire_y1, nw), pwire_x2=_arr(s.pwire_x2, nw),
                        pwire_y2=_arr(s.pwire_y2, nw), pwire_width=_arr(s.pwire_width, nw),
                        es_type1=_arr(s.es_type1, ne), es_type2=_arr(s.es_type2, ne),
                        es_spacing_um=_arr(s.es_spacing_um, ne, np.float64))
        return Design(
            core=(s.core_lx, s.core_ly, s.core_ux, s.core_uy),
            row_lx=_arr(s.row_lx, nr), row_ly=_arr(s.row_ly, nr), row_dx=_arr(s.row_dx, nr),
            row_dy=_arr(s.row_dy, nr), row_site_width=_arr(s.row_site_width, nr),
            row_orient=_arr(s.row_orient, nr),
            inst_lx=_arr(s.inst_lx, n), inst_ly=_arr(s.inst_ly, n), inst_dx=_arr(s.inst_dx, n),
            inst_dy=_arr(s.inst_dy, n), inst_fixed=_arr(s.inst_fixed, n, np.uint8),
            inst_macro=_arr(s.inst_macro, n, np.uint8), inst_group=_arr(s.inst_group, n),
            inst_ground_at_bottom=_arr(s.inst_ground_at_bottom, n, np.uint8),
            num_groups=s.num_groups, box_group=_arr(s.box_group, nb), box_lx=_arr(s.box_lx, nb),
            box_ly=_arr(s.box_ly, nb), box_ux=_arr(s.box_ux, nb), box_uy=_arr(s.box_uy, nb),
            net_start=ns if nn else np.zeros(1, np.int32), pin_inst=_arr(s.pin_inst, npins),
            pin_dx=_arr(s.pin_dx, npins), pin_dy=_arr(s.pin_dy, npins),
            inst_name=[L.ld_inst_name(self.h, i).decode() for i in range(n)],
            dbu_per_micron=s.dbu_per_micron, tech=tech, meta={"name": "def"})

    def write_def(self, path: str, lx=None, ly=None, flipped=None):
        """The DEF that was read, with the COMPONENTS section regenerated from
        the given instance positions (dbDesign::writeDef)."""
        keep = []

        def p(a, t, ct):
            if a is None:
                return None
            a = np.ascontiguousarray(a, t)
            keep.append(a)
            return a.ctypes.data_as(C.POINTER(ct))

        _check(lib().ld_write_def(self.h, str(path).encode(), p(lx, np.int32, C.c_int32),
                                  p(ly, np.int32, C.c_int32), p(flipped, np.uint8, C.c_uint8)))

    def close(self):
        if self.h:
            lib().ld_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
