"""LEF / DEF through the drop-in entry point and the command driver on the GPU.

* lg_run_design on a design read from LEF / DEF with ICCAD 2017 technology
  (edge types, pin layers, power stripes, edge spacing table) equals the
  oracle's run_design bit for bit: positions, flips, iterations, the
  checkPlace counts including pin shorts, pin access and edge spacing.
* The command driver's read_lef / read_def / legalize / write_def writes a
  DEF whose components sit where lg_run_design puts them.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from admm_legalization_amd import build, lg, synth
from admm_legalization_amd.design import strictly_legal
from admm_legalization_amd.lefdef import LefDef
from oracle import ref

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import lefdef_writer  # noqa: E402

pytestmark = pytest.mark.gpu


def _files(tmp_path, n=4000, seed=8, fences=2):
    d = synth.with_netlist(synth.mixed(n, util=0.78, seed=seed, num_fences=fences, name="lefdef"), seed=seed)
    lef, dfile = str(tmp_path / "cells.lef"), str(tmp_path / "design.def")
    lefdef_writer.write(d, lef, dfile, tech=True, seed=seed)
    return lef, dfile


def test_run_design_from_lefdef_matches_oracle(gpu, tmp_path):
    lef, dfile = _files(tmp_path)
    d = LefDef(lef, dfile).design()
    print(f"[case] lefdef run_design n={d.num_insts}", flush=True)
    lx, ly, fl, st = lg.run_design(d)
    olx, oly, ofl, ost = ref.run_design(d)
    np.testing.assert_array_equal(lx, olx)
    np.testing.assert_array_equal(ly, oly)
    np.testing.assert_array_equal(fl, ofl)
    assert st["iterations"] == ost["iterations"]
    assert st["vio"] == ost["vio"], (st["vio"], ost["vio"])
    assert st["total_disp_dbu"] == ost["total_disp_dbu"]
    assert st["converged"] == 1 and strictly_legal(st["vio"])
    assert st["vio"]["num_pin_short_vio"] + st["vio"]["num_pin_access_vio"] > 0  # the terms are live


def test_command_driver_lefdef_flow(gpu, tmp_path):
    lef, dfile = _files(tmp_path, n=3000, seed=9, fences=1)
    out = str(tmp_path / "legal.def")
    script = tmp_path / "flow.tcl"
    script.write_text(f"read_lef {lef}\nread_def {dfile}\nlegalize -tech_penalty 2.0\nwrite_def {out}\n")
    print("[case] skyline_cmd read_lef / read_def / legalize / write_def", flush=True)
    r = subprocess.run([build.CMD, str(script)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=300)
    print(r.stdout[-1200:], flush=True)
    assert r.returncode == 0 and "[summary] converged=1" in r.stdout
    d = LefDef(lef, dfile).design()
    p = lg.default_params()
    p.tech_penalty = 2.0
    lx, ly, fl, st = lg.run_design(d, p)
    back = LefDef(lef, out).design()
    np.testing.assert_array_equal(back.inst_lx, lx)
    np.testing.assert_array_equal(back.inst_ly, ly)
    assert f"pin_short_vio={st['vio']['num_pin_short_vio']} " in r.stdout
