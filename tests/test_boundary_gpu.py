"""The drop-in boundary on the GPU, against the oracle end to end:

* lg_run_design (legalizer::Legalizer::run) with a netlist, a size file and
  an ICCAD 2017 constraint file: per-instance positions and flips, fixed
  instances passed through, displacement averaged over all instances, HPWL
  before / after and the contest metric, all equal to ref.run_design;
* the command driver skyline_cmd (read_design, legalize with the Tcl
  options, write_result), which drives the C++ class legalizer::Legalizer,
  and its global_place (the C++ class skyplace::SkyPlace) followed by
  legalize, equal to the Python path through the same C ABI;
* a C++ caller of the device layer (tests/cpp/device_caller.cpp:
  lg_import_and_preplace, lg_create .. lg_destroy) against ref.solve.
"""
import os
import subprocess

import numpy as np
import pytest

from admm_legalization_amd import build, gp, lg, synth
from admm_legalization_amd.design import strictly_legal
from oracle import ref

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _design_and_files(tmp_path):
    d = synth.with_netlist(synth.mixed(4000, seed=12, util=0.75), seed=4)
    mv = np.flatnonzero(d.movable_mask())
    rng = np.random.default_rng(3)
    sizes = tmp_path / "sizes.txt"
    sizes.write_text("".join(f"o{i}={int(rng.integers(2, 10))}=1\n" for i in rng.choice(mv, 30, replace=False)))
    cons = tmp_path / "placement.constraints"
    cons.write_text("maximum_utilization=80%\nmaximum_movement=10rows\n")
    return d, str(sizes), str(cons)


def _same_stats(a, b):
    for k in ("iterations", "converged", "total_disp_dbu", "max_disp_dbu", "num_cells"):
        assert a[k] == b[k], k
    for k in ("total_disp_um", "avg_disp_um", "max_disp_um", "hpwl_before_um", "hpwl_after_um", "delta_hpwl_pct",
              "max_util_constraint", "max_disp_constraint", "iccad17_s_am", "iccad17_m_max", "iccad17_score"):
        assert a[k] == pytest.approx(b[k], rel=1e-12, abs=1e-12), k
    assert a["vio"] == b["vio"]


def test_run_design_with_nets_sizes_and_constraints(gpu, tmp_path):
    d, sizes, cons = _design_and_files(tmp_path)
    p = lg.default_params()
    p.flag_iccad17 = 1
    print(f"[case] lg_run_design n={d.num_insts} nets={len(d.net_start) - 1}", flush=True)
    lx, ly, fl, st = lg.run_design(d, p, trace_cap=30000, size_file=sizes, constraint_file=cons)
    olx, oly, ofl, ost = ref.run_design(d, p, trace_cap=30000, size_file=sizes, constraint_file=cons)
    np.testing.assert_array_equal(lx, olx)
    np.testing.assert_array_equal(ly, oly)
    np.testing.assert_array_equal(fl, ofl)
    np.testing.assert_array_equal(st["ovf_trace"], ost["ovf_trace"])
    _same_stats(st, ost)
    assert st["converged"] == 1 and strictly_legal(st["vio"])
    assert st["hpwl_before_um"] > 0 and st["iccad17_score"] > 0
    fixed = ~d.movable_mask()
    assert np.array_equal(lx[fixed], d.inst_lx[fixed]) and np.array_equal(ly[fixed], d.inst_ly[fixed])
    assert st["avg_disp_um"] == pytest.approx(st["total_disp_um"] / d.num_insts)


def test_command_driver_end_to_end(gpu, tmp_path):
    d, sizes, cons = _design_and_files(tmp_path)
    dpath, rpath = str(tmp_path / "d.lgd"), str(tmp_path / "r.lgr")
    d.save_lgd(dpath)
    script = tmp_path / "flow.tcl"
    script.write_text(f"read_design {dpath}\n"
                      f"legalize -constraint {cons} -size_file {sizes} -max_iter 20000 -sorting_algorithm M_BITONIC\n"
                      f"write_result {rpath}\n")
    print("[case] skyline_cmd", flush=True)
    r = subprocess.run([build.CMD, str(script)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=300)
    print(r.stdout[-1500:], flush=True)
    assert r.returncode == 0
    for line in ("HPWL Before", "Delta HPWL", "NumOverlap", "ICCAD 17 Contest Metric", "[summary] converged=1"):
        assert line in r.stdout
    raw = open(rpath, "rb").read()
    assert raw[:4] == b"LGR1"
    n = int(np.frombuffer(raw[4:8], np.int32)[0])
    lx = np.frombuffer(raw[8:8 + 4 * n], np.int32)
    ly = np.frombuffer(raw[8 + 4 * n:8 + 8 * n], np.int32)
    fl = np.frombuffer(raw[8 + 8 * n:8 + 9 * n], np.uint8)
    p = lg.default_params()
    p.flag_iccad17 = 1  # -constraint turns the ICCAD17 weighting on (CmdLegalize.hpp:42-49)
    p.max_admm_iter = 20000
    olx, oly, ofl, ost = ref.run_design(d, p, size_file=sizes, constraint_file=cons)
    np.testing.assert_array_equal(lx, olx)
    np.testing.assert_array_equal(ly, oly)
    np.testing.assert_array_equal(fl, ofl)
    assert f"iccad17_score={ost['iccad17_score']:.9f}" in r.stdout


def test_cpp_device_layer_caller(gpu, tmp_path):
    d = synth.mixed(6000, seed=21, util=0.8)
    dpath, opath = str(tmp_path / "d.lgd"), str(tmp_path / "o.bin")
    d.save_lgd(dpath)
    exe = os.path.join(ROOT, "tests", "cpp", "_build", "device_caller")
    assert os.path.exists(exe), "build() compiles tests/cpp/device_caller.cpp"
    print("[case] device_caller", flush=True)
    r = subprocess.run([exe, dpath, opath], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    print(r.stdout, flush=True)
    assert r.returncode == 0 and "converged=1 overlap=0" in r.stdout
    raw = np.fromfile(opath, np.int32)
    n = int(raw[0])
    grid, cells, p2 = lg.device_inputs(d)
    o = ref.solve(grid, cells, p2)
    assert n == len(cells["w_grid"])
    np.testing.assert_array_equal(raw[1:1 + n], o["lx"])
    np.testing.assert_array_equal(raw[1 + n:1 + 2 * n], o["ly"])
    np.testing.assert_array_equal(raw[1 + 2 * n:1 + 3 * n], o["rot"])
    assert f"total_disp_dbu={o['stats']['total_disp_dbu']}" in r.stdout


def test_command_driver_global_place_then_legalize(gpu, tmp_path):
    d = synth.e2e(8000, seed=3)
    dpath, rpath = str(tmp_path / "d.lgd"), str(tmp_path / "r.lgr")
    d.save_lgd(dpath)
    script = tmp_path / "flow.tcl"
    script.write_text(f"read_design {dpath}\n"
                      "global_place -target_density 0.95 -bins 64 -scaling_factor 100 -ref_hpwl_ratio 0.01 "
                      "-precision F32\n"
                      "legalize\n"
                      f"write_result {rpath}\n")
    print("[case] skyline_cmd global_place + legalize", flush=True)
    r = subprocess.run([build.CMD, str(script)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=300)
    print(r.stdout[-2000:], flush=True)
    assert r.returncode == 0
    assert "[gp_summary] converged=1" in r.stdout and "[summary] converged=1" in r.stdout
    assert "Nesterov Solve Start" in r.stdout
    raw = open(rpath, "rb").read()
    n = int(np.frombuffer(raw[4:8], np.int32)[0])
    lx = np.frombuffer(raw[8:8 + 4 * n], np.int32)
    ly = np.frombuffer(raw[8 + 4 * n:8 + 8 * n], np.int32)
    # the same flow through the Python mirrors of the C ABI
    p = gp.default_params()
    p.target_density, p.num_bin_x, p.num_bin_y = 0.95, 64, 64
    p.scaling_factor, p.reference_hpwl_ratio, p.precision = 100.0, 0.01, 2
    glx, gly, gst = gp.place_design(d, p)
    olx, oly, ofl, ost = lg.run_design(d.copy_positions(glx, gly))
    np.testing.assert_array_equal(lx, olx)
    np.testing.assert_array_equal(ly, oly)
    assert f"iterations={gst['iterations']}" in r.stdout and ost["converged"] == 1 and strictly_legal(ost["vio"])
