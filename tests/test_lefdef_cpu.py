"""LEF / DEF reading and DEF writing (include/skyline_lefdef.h) on the CPU.

* A synthetic design written as LEF + DEF by tests/lefdef_writer.py (rows,
  fixed macros, fences by pattern and by name, nets with IO pins, other
  sections the reader skips) reads back as the Design the writer expects:
  rows, instances, sizes, fixed / macro flags, rails from the POWER / GROUND
  pins, groups, fence boxes, pin positions for HPWL.
* The design-level ICCAD 2017 technology (LEF58_EDGETYPE, signal-pin layers,
  SPECIALNETS stripes, LEF58_CELLEDGESPACINGTABLE) becomes the grid-level
  inputs of the legalizer as Legalizer::importTechInfo makes them: checked
  against an independent numpy rasterisation of the stripes and pin boxes.
* ld_write_def writes the legalized positions: read back, the instances sit
  where the legalizer put them (FS where flipped), fixed ones unchanged, and
  every section but COMPONENTS is the input's text.
"""
import os
import sys

import numpy as np
import pytest

from admm_legalization_amd import synth
from admm_legalization_amd.design import strictly_legal
from admm_legalization_amd.lefdef import LefDef
from oracle import ref

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import lefdef_writer  # noqa: E402

FIELDS = ("row_lx", "row_ly", "row_dx", "row_dy", "row_site_width", "row_orient", "inst_lx", "inst_ly",
          "inst_dx", "inst_dy", "inst_fixed", "inst_macro", "inst_group", "inst_ground_at_bottom",
          "box_group", "box_lx", "box_ly", "box_ux", "box_uy", "net_start", "pin_inst", "pin_dx", "pin_dy")


def _design(n=2500, seed=3, fences=2):
    d = synth.mixed(n, util=0.75, seed=seed, num_fences=fences, name="lefdef")
    return synth.with_netlist(d, seed=seed)


@pytest.fixture(scope="module")
def files(tmp_path_factory):
    tmp = tmp_path_factory.mktemp("lefdef")
    d = _design()
    lef, dfile = str(tmp / "cells.lef"), str(tmp / "design.def")
    exp, info = lefdef_writer.write(d, lef, dfile, tech=True, seed=4)
    return dict(d=d, lef=lef, def_=dfile, exp=exp, info=info, tmp=tmp)


def test_read_reconstructs_the_design(files):
    got = LefDef(files["lef"], files["def_"]).design()
    exp = files["exp"]
    assert got.core == exp.core and got.num_groups == exp.num_groups and got.dbu_per_micron == exp.dbu_per_micron
    for f in FIELDS:
        np.testing.assert_array_equal(np.asarray(getattr(got, f)), np.asarray(getattr(exp, f)), err_msg=f)
    assert got.inst_name == exp.inst_name
    assert got.tech is not None and len(got.tech["master_ledge"]) == len(files["info"]["masters"])


def test_tech_import_matches_independent_rasterisation(files):
    """importTechInfo's grid inputs from the LEF / DEF technology: power
    metal on routing layers 1-2 only, widened by half the wire width,
    clipped to the core, end points as Grid::getEndPointIndexExpanded; the
    pin boxes in whole sites / rows; the rule table in sites."""
    d = LefDef(files["lef"], files["def_"]).design()
    t = ref.import_tech(d)
    xs, ys = t["x_size"], t["y_size"]
    lx, ly, ux, uy = d.core
    sw, rh = d.site_width, d.row_height
    pm = np.full((xs, ys), -1, np.int64)

    def end(v, o, s, size):
        q, r = divmod(v - o, s)
        return min(max(q - 1 if r == 0 else q, 0), size - 1)

    for layer, x1, y1, x2, y2, w in files["info"]["wires"]:
        if layer not in (1, 2):
            continue
        a, b = max(lx, min(x1, x2) - w // 2), max(ly, min(y1, y2) - w // 2)
        c, e = min(ux, max(x1, x2) + w // 2), min(uy, max(y1, y2) + w // 2)
        x0, x1g = (a - lx) // sw, end(c, lx, sw, xs)
        y0, y1g = (b - ly) // rh, end(e, ly, rh, ys)
        blk = pm[x0:x1g + 1, y0:y1g + 1]
        pm[x0:x1g + 1, y0:y1g + 1] = np.where(blk == -1, layer, np.minimum(blk, layer))
    np.testing.assert_array_equal(t["power_layer"].reshape(xs, ys), pm)
    assert (pm == 1).any() and (pm == 2).any()
    # rules: spacing (um) * dbu truncated, in whole sites, at index a + b
    want = np.zeros(5, np.int64)
    for a, b, s in files["info"]["rules"]:
        want[a + b] = int(s * float(np.float32(d.dbu_per_micron))) // sw
    np.testing.assert_array_equal(t["rules"], want)
    # pins: each tech master's signal pins in grid units
    info = files["info"]
    tech_id = {}
    for m, key in enumerate(info["masters"]):
        if not key[2]:
            tech_id[m] = len(tech_id)
    mv = d.movable_mask()
    inst = np.nonzero(mv)[0]
    for k, i in enumerate(inst[:300]):
        m = int(info["master_of"][i])
        tid = tech_id[m]
        assert t["macro"][k] == tid
        assert (t["ledge"][k], t["redge"][k]) == (info["ledge"][m], info["redge"][m])
        wg, hg = int(d.inst_dx[i]) // sw, int(d.inst_dy[i]) // rh
        s0, s1 = t["macro_pin_start"][tid], t["macro_pin_start"][tid + 1]
        assert s1 - s0 == len(info["sig"][m])
        for j, (layer, bx0, by0, bx1, by1) in enumerate(info["sig"][m]):
            assert t["pin_layer"][s0 + j] == layer
            assert t["pin_bgn"][s0 + j] == (bx0 // sw) * hg + by0 // rh
            assert t["pin_end"][s0 + j] == min(wg - 1, bx1 // sw) * hg + min(hg - 1, by1 // rh)


def test_oracle_legalizes_the_def_design_with_technology(files):
    d = LefDef(files["lef"], files["def_"]).design()
    lx, ly, fl, st = ref.run_design(d)
    assert st["converged"] == 1 and strictly_legal(st["vio"]), st["vio"]
    assert st["num_cells"] == int(d.movable_mask().sum())


def test_write_def_round_trip(files):
    db = LefDef(files["lef"], files["def_"])
    d = db.design()
    lx, ly, fl, st = ref.run_design(d)
    out = str(files["tmp"] / "out.def")
    db.write_def(out, lx, ly, fl)
    back = LefDef(files["lef"], out).design()
    np.testing.assert_array_equal(back.inst_lx, lx)
    np.testing.assert_array_equal(back.inst_ly, ly)
    np.testing.assert_array_equal(back.inst_fixed, d.inst_fixed)
    text_in, text_out = open(files["def_"]).read(), open(out).read()
    # FS for the flipped movable cells, N for the others
    comps = text_out[text_out.index("COMPONENTS"):text_out.index("END COMPONENTS")].splitlines()[1:]
    orients = [line.split()[-2] for line in comps]
    mv = d.movable_mask()
    assert all(o == ("FS" if f else "N") for o, f, m in zip(orients, fl, mv) if m)
    # every other section is the input's
    cut = lambda t: t[:t.index("COMPONENTS")] + t[t.index("END COMPONENTS") + len("END COMPONENTS"):].lstrip()
    assert cut(text_in) == cut(text_out)
