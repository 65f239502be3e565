"""The benchmark configurations at their full sizes on the GPU.

* plain_20k (20k single-row cells, seed 0): the device layer equals the
  oracle run live on the same inputs bit for bit (positions, rotation flags,
  pixel usage, dual prices, overflow trace), and the drop-in entry point
  lg_run_design equals the oracle's end-to-end result (digest fixture).
* mixed_1m (1M mixed-height cells, 1600 x 12000, 3 fences, 200 macros):
  lg_run_design converges, is strictly legal (no overlap, on valid sites of
  the cell's own fence, inside the core, rails right) and equals the
  oracle's end-to-end result (digest fixture written by
  tests/golden/make_config_golden.py).
* density_1m (1M cells, 1024 x 1024 bins): scatter, overflow, Poisson and
  gather against the FP64 oracle for all three Poisson precisions.
* stream (2000 requests, 500-200k cells): every request through the request
  server, each strictly legal and converged; every 50th request equal to
  the oracle's end-to-end result (digest fixture).
* e2e_2m (2M mixed-height cells, 90 % utilisation, 300 macros, 2.1M nets):
  global placement from scratch on 4096 x 4096 bins (BF16 transforms)
  converges and gives the same bits twice; the legalizer then converges
  strictly legal from its result.  The loop's parity with the oracle is
  checked step by step and end to end at 3000 cells (test_gp_gpu.py).
"""
import json
import os
import sys
import time

import numpy as np
import pytest

from admm_legalization_amd import lg, synth
from admm_legalization_amd.design import strictly_legal
from oracle import ref

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
from config_digest import array_digest, grid_to_dbu, position_digest  # noqa: E402

pytestmark = pytest.mark.gpu

GOLD_PATH = os.path.join(HERE, "golden", "config_oracle.json")
GOLD = json.load(open(GOLD_PATH)) if os.path.exists(GOLD_PATH) else {}


def _gold(name):
    if name not in GOLD:
        pytest.skip(f"no oracle fixture for {name} (tests/golden/make_config_golden.py {name})")
    return GOLD[name]


def _check_design_result(design, lx, ly, fl, st, want):
    mov = design.movable_mask()
    assert st["converged"] == 1, st
    assert strictly_legal(st["vio"]), st["vio"]
    assert int(mov.sum()) == want["num_movable"]
    assert st["iterations"] == want["iterations"], (st["iterations"], want["iterations"])
    assert array_digest(st["ovf_trace"]) == want["trace"]
    assert st["total_disp_dbu"] == want["total_disp_dbu"]
    assert st["max_disp_dbu"] == want["max_disp_dbu"]
    assert position_digest(lx[mov], ly[mov], fl[mov]) == want["positions"]
    # fixed instances pass through unchanged
    fixed = ~mov
    assert np.array_equal(lx[fixed], design.inst_lx[fixed]) and np.array_equal(ly[fixed], design.inst_ly[fixed])


def test_config1_plain_20k_bit_exact(gpu):
    d = synth.plain_20k(0)
    p = lg.default_params()
    grid, cells, p2 = lg.device_inputs(d, p)
    print(f"[case] plain_20k device layer n={len(cells['w_grid'])}", flush=True)
    dev = lg.DeviceLegalizer(grid, cells, p2)
    dev.solve()
    usage, dual = dev.pixel_state()
    dev.postprocess()
    lx, ly, rot = dev.positions()
    st = dev.stats()
    tr = dev.trace()
    dev.close()
    t = time.perf_counter()
    o = ref.solve(grid, cells, p2)
    print(f"  oracle {time.perf_counter() - t:.1f}s, {o['stats']['iterations']} iterations", flush=True)
    assert st["rho"] == o["stats"]["rho"] and st["initial_dual"] == o["stats"]["initial_dual"]
    assert st["iterations"] == o["stats"]["iterations"]
    np.testing.assert_array_equal(tr, o["trace"])
    np.testing.assert_array_equal(usage, o["usage"])
    np.testing.assert_array_equal(dual.view(np.int32), o["dual"].view(np.int32))
    np.testing.assert_array_equal(lx, o["lx"])
    np.testing.assert_array_equal(ly, o["ly"])
    np.testing.assert_array_equal(rot, o["rot"])
    assert st["vio"] == o["stats"]["vio"] and strictly_legal(st["vio"])
    assert st["total_disp_dbu"] == o["stats"]["total_disp_dbu"]


def test_config1_plain_20k_run_design(gpu):
    """The drop-in path (lg_run_design: import, grid, preplace, device
    layer, statistics) against the oracle's end-to-end result."""
    want = _gold("plain_20k")
    d = synth.plain_20k(0)
    print("[case] plain_20k lg_run_design", flush=True)
    lx, ly, fl, st = lg.run_design(d, trace_cap=200000)
    _check_design_result(d, lx, ly, fl, st, want)


def test_config3_mixed_1m_run_design(gpu):
    want = _gold("mixed_1m")
    t = time.perf_counter()
    d = synth.mixed_1m(0)
    print(f"[case] mixed_1m lg_run_design (synth {time.perf_counter() - t:.1f}s)", flush=True)
    t = time.perf_counter()
    lx, ly, fl, st = lg.run_design(d, trace_cap=200000)
    print(f"  {time.perf_counter() - t:.2f}s end to end, {st['iterations']} iterations, vio {st['vio']}",
          flush=True)
    _check_design_result(d, lx, ly, fl, st, want)


@pytest.mark.parametrize("prec_name", ["SPLIT_BF16", "F32", "BF16"])
def test_config2_density_1m(gpu, prec_name):
    """Scatter, overflow, Poisson solve and gather of 1M cells at 1024 x 1024
    bins.  The potential and field are compared twice: against the FP64
    transforms of the device's own bin density (the Poisson error alone,
    bounded by the precision's tolerance), and end to end against the FP64
    oracle of the whole chain."""
    from admm_legalization_amd import density as D
    from oracle import ref_density as R
    import test_density_gpu as TD
    prec = getattr(D, "PREC_" + prec_name)
    pl = _density_1m()
    out = TD._run_density(pl, prec=prec)
    assert (out["geom"].num_bin_x, out["geom"].num_bin_y) == (1024, 1024)
    own = R.poisson(np.asarray(out["dens"], np.float64).reshape(1024, 1024))
    e_own = [TD._rel(out[k], r) for k, r in zip(("psi", "ex", "ey"), own)]
    o = _density_1m_oracle(pl, out["geom"])
    errs = {k: TD._rel(out[k], o[k]) for k in ("mov", "psi", "ex", "ey", "gx", "gy")}
    print(f"  {prec_name}: Poisson vs FP64 on the device density {e_own}; end to end {errs}", flush=True)
    assert max(e_own) <= TD.TOL[prec], e_own
    TD._compare(out, o, field_tol=TD.TOL[prec])


_DENS = {}


def _density_1m():
    if "pl" not in _DENS:
        _DENS["pl"] = synth.density_placement(1_000_000, seed=0)
    return _DENS["pl"]


def _density_1m_oracle(pl, g):
    import test_density_gpu as TD
    if "o" not in _DENS:
        t = time.perf_counter()
        _DENS["o"] = TD._oracle(pl, g)
        print(f"  FP64 oracle {time.perf_counter() - t:.1f}s", flush=True)
    return _DENS["o"]


def test_config4_stream_2000_requests(gpu):
    """All 2000 requests of the stream, submitted as they are prepared, to
    one request server; every result strictly legal, every 50th equal to the
    oracle."""
    from admm_legalization_amd.server import LegalizationServer
    gold = _gold("stream")["requests"]
    spec = synth.stream_spec(2000)
    t0 = time.perf_counter()
    tickets, geo, cells_total = [], {}, 0
    with LegalizationServer(max_slots=64, iters_per_sync=16) as srv:
        for i, (n, u, sd) in enumerate(spec):
            g, c, p = lg.device_inputs(synth.stream_request(n, u, sd))
            tickets.append(srv.submit(g, c, p))
            cells_total += n
            if str(i) in gold:
                geo[i] = {k: g[k] for k in ("site_width", "row_height", "core_lx", "core_ly")}
            if (i + 1) % 250 == 0:
                print(f"  submitted {i + 1}/2000 at {time.perf_counter() - t0:.1f}s", flush=True)
        bad, checked = [], 0
        for i, tk in enumerate(tickets):
            r = srv.wait(tk)
            st = r["stats"]
            if st["converged"] != 1 or not strictly_legal(st["vio"]):
                bad.append((i, st["converged"], st["vio"]))
            if i in geo:
                w = gold[str(i)]
                lx, ly = grid_to_dbu(geo[i], r["lx"], r["ly"])
                ok = (position_digest(lx, ly, r["rot"]) == w["positions"] and st["iterations"] == w["iterations"]
                      and st["total_disp_dbu"] == w["total_disp_dbu"])
                if not ok:
                    bad.append((i, "differs from the oracle", st["iterations"], w["iterations"]))
                checked += 1
            if (i + 1) % 250 == 0:
                print(f"  collected {i + 1}/2000 at {time.perf_counter() - t0:.1f}s", flush=True)
        ctr = srv.counters()
    print(f"  {cells_total} cells in {time.perf_counter() - t0:.1f}s, peak slots {ctr['peak_slots']}, "
          f"{checked} requests against the oracle", flush=True)
    assert checked == len(gold)
    assert not bad, bad[:10]


def test_config5_e2e_2m(gpu):
    from admm_legalization_amd import gp
    d = synth.e2e(2_000_000, seed=0)
    p = synth.e2e_placer_params(d, 4096)
    print(f"[case] e2e_2m: {d.num_insts} instances, {len(d.net_start) - 1} nets, {int(d.net_start[-1])} pins",
          flush=True)
    t = time.perf_counter()
    lx, ly, st = gp.place_design(d, p)
    t_gp = time.perf_counter() - t
    print(f"global placement: {st['iterations']} iterations, overflow {st['overflow']:.4f}, "
          f"HPWL {st['hpwl_db']:.6g} (design {st['hpwl_initial_db']:.6g}), {t_gp:.2f}s", flush=True)
    assert st["converged"] == 1 and st["overflow"] <= p.target_overflow
    assert st["num_bin_x"] == st["num_bin_y"] == 4096
    assert st["hpwl_db"] < st["hpwl_initial_db"]  # the placer recovers the netlist's locality
    lx2, ly2, st2 = gp.place_design(d, p)
    assert np.array_equal(lx, lx2) and np.array_equal(ly, ly2) and st2["iterations"] == st["iterations"]
    placed = d.copy_positions(lx, ly)
    t = time.perf_counter()
    olx, oly, fl, ls = lg.run_design(placed)
    print(f"legalization: {ls['iterations']} iterations, {time.perf_counter() - t:.2f}s, "
          f"avg disp {ls['avg_disp_um']:.3f} um", flush=True)
    assert ls["converged"] == 1 and strictly_legal(ls["vio"]), ls["vio"]
    fixed = ~d.movable_mask()
    assert np.array_equal(olx[fixed], d.inst_lx[fixed]) and np.array_equal(oly[fixed], d.inst_ly[fixed])
