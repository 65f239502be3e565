"""GPU tests of the request server (continuous batching): every request's
result in the mix is the result it gets alone, bit for bit -- positions,
rotation flags, iteration count, displacement -- whatever it is batched with,
however many slots and whatever memory budget the server has."""
import time

import numpy as np
import pytest

from admm_legalization_amd import lg, synth
from admm_legalization_amd.design import SORT_ALGO
from admm_legalization_amd.server import LegalizationServer, LegalizerError

pytestmark = pytest.mark.gpu

# (cells, utilisation, seed, fences)
REQS = [(600, 0.70, 11, 0), (4000, 0.85, 12, 3), (1500, 0.90, 13, 0), (9000, 0.80, 14, 3),
        (800, 0.60, 15, 0), (2500, 0.88, 16, 3), (12000, 0.75, 17, 3)]


def _inputs(n, util, seed, fences, **pkw):
    d = synth.mixed(n, util=util, seed=seed, num_fences=fences, name=f"req{seed}")
    p = lg.default_params()
    for k, v in pkw.items():
        setattr(p, k, v)
    return lg.device_inputs(d, p)


def _solo(grid, cells, p):
    dev = lg.DeviceLegalizer(grid, cells, p)
    dev.solve()
    dev.postprocess()
    lx, ly, rot = dev.positions()
    st = dev.stats()
    dev.close()
    return dict(lx=lx, ly=ly, rot=rot, stats=st)


def _same(got, want, tag):
    np.testing.assert_array_equal(got["lx"], want["lx"], err_msg=tag)
    np.testing.assert_array_equal(got["ly"], want["ly"], err_msg=tag)
    np.testing.assert_array_equal(got["rot"], want["rot"], err_msg=tag)
    for k in ("iterations", "converged", "total_disp_dbu", "max_disp_dbu", "final_overflow", "rho",
              "initial_dual", "num_cells"):
        assert got["stats"][k] == want["stats"][k], (tag, k, got["stats"][k], want["stats"][k])
    assert got["stats"]["vio"] == want["stats"]["vio"], tag


@pytest.fixture(scope="module")
def cases(gpu):
    out = []
    for n, u, s, f in REQS:
        grid, cells, p = _inputs(n, u, s, f)
        out.append((f"n{n}_u{u}_s{s}", grid, cells, p, _solo(grid, cells, p)))
        print(f"[solo] n={n} util={u} seed={s}: {out[-1][4]['stats']['iterations']} iterations", flush=True)
    return out


@pytest.mark.parametrize("slots,iters,graphs", [(64, 16, True), (3, 8, True), (1, 5, True), (64, 16, False),
                                                 (3, 8, False)])
def test_batched_results_equal_solo(cases, slots, iters, graphs):
    """Iterations as graph launches (recaptured whenever the resident set
    changes) or as kernel launches: the same results."""
    with LegalizationServer(max_slots=slots, iters_per_sync=iters, mem_budget_bytes=4 << 30,
                            pool_chunk_bytes=64 << 20, use_graphs=graphs) as srv:
        tickets = [srv.submit(g, c, p) for _, g, c, p, _ in cases]
        res = [srv.wait(t) for t in tickets]
        ctr = srv.counters()
    for (tag, _, _, _, want), got in zip(cases, res):
        _same(got, want, tag)
        assert got["stats"]["vio"]["total_ovf"] == 0 and got["stats"]["converged"] == 1, tag
    assert ctr["completed"] == len(cases) and ctr["failed"] == 0
    assert (ctr["graph_captures"] > 0) == graphs
    assert 1 <= ctr["peak_slots"] <= min(slots, len(cases))
    if slots >= len(cases):
        assert ctr["peak_slots"] == len(cases)  # all resident together


def test_staggered_arrivals_equal_solo(cases):
    """Requests submitted while others are mid-solve join the running batch."""
    with LegalizationServer(max_slots=8, iters_per_sync=4, mem_budget_bytes=4 << 30) as srv:
        tickets = []
        for i, (_, g, c, p, _) in enumerate(cases):
            tickets.append(srv.submit(g, c, p))
            time.sleep(0.01 * (i % 3))
        res = {t: srv.wait(t) for t in reversed(tickets)}
    for (tag, _, _, _, want), t in zip(cases, tickets):
        _same(res[t], want, tag)
        assert res[t]["latency_s"] >= res[t]["run_s"] >= 0


def test_mixed_sorting_algorithms_in_one_batch(gpu):
    """Register-sorted (bitonic) and LDS-sorted (merge / stable) requests and
    an adaptive-search request share the batched launches."""
    specs = [dict(sorting_algorithm=SORT_ALGO["S_MERGE"]), dict(sorting_algorithm=SORT_ALGO["CUB"]),
             dict(sorting_algorithm=SORT_ALGO["M_BITONIC"]),
             dict(use_adaptive_search=1, min_iter_to_call_adaptive_search=20),
             dict(standard_admm=1, max_admm_iter=300)]
    cases = []
    for i, kw in enumerate(specs):
        g, c, p = _inputs(2500 + 500 * i, 0.85, 40 + i, 3 if i % 2 else 0, **kw)
        cases.append((str(kw), g, c, p, _solo(g, c, p)))
    with LegalizationServer(max_slots=16, iters_per_sync=8, mem_budget_bytes=2 << 30) as srv:
        tickets = [srv.submit(g, c, p) for _, g, c, p, _ in cases]
        for (tag, _, _, _, want), t in zip(cases, tickets):
            _same(srv.wait(t), want, tag)


def _min_bytes(grid, cells):
    """A lower bound of a request's device buffers: 14 B per pixel (codes,
    regions, usage, duals) and 120 B per cell."""
    return 14 * int(grid["x_size"]) * int(grid["y_size"]) + 120 * len(cells["w_grid"])


def test_memory_budget_limits_residency(cases):
    """A budget of 1.5x the largest request: all still finish, fewer are
    resident at once, results unchanged."""
    need = [_min_bytes(g, c) for _, g, c, _, _ in cases]
    budget = int(1.5 * max(need)) + (64 << 10)
    assert sum(need) > budget
    with LegalizationServer(max_slots=64, iters_per_sync=8, mem_budget_bytes=budget,
                            pool_chunk_bytes=1 << 20) as srv:
        tickets = [srv.submit(g, c, p) for _, g, c, p, _ in cases]
        res = [srv.wait(t) for t in tickets]
        ctr = srv.counters()
    for (tag, _, _, _, want), got in zip(cases, res):
        _same(got, want, tag)
    assert ctr["pool_reserved_bytes"] <= budget
    assert ctr["peak_slots"] < len(cases)


def test_request_over_budget_is_rejected(cases):
    _, g, c, p, _ = cases[-1]
    with LegalizationServer(max_slots=4, mem_budget_bytes=1 << 20) as srv:
        with pytest.raises(LegalizerError, match="budget"):
            srv.submit(g, c, p)


def test_zero_iteration_request(cases):
    _, g, c, p, _ = cases[0]
    p0 = lg.LgParams.from_buffer_copy(p)
    p0.max_admm_iter = 0
    want = _solo(g, c, p0)
    with LegalizationServer(max_slots=2) as srv:
        got = srv.wait(srv.submit(g, c, p0))
    _same(got, want, "max_admm_iter=0")


def test_destroy_while_a_thread_waits(gpu):
    """lg_srv_destroy with a thread blocked in lg_srv_wait: the waiter gets
    LG_ERR_STATE (or its result, had it finished first) and destroy returns
    only after it has left the call."""
    import threading
    from admm_legalization_amd.server import LegalizationServer, LegalizerError
    d = synth.plain_20k(0)  # thousands of iterations: still running at destroy
    g, c, p = lg.device_inputs(d)
    srv = LegalizationServer(max_slots=4, iters_per_sync=8)
    t1 = srv.submit(g, c, p)
    t2 = srv.submit(g, c, p)
    out = {}

    def waiter():
        try:
            out["result"] = srv.wait(t1)
        except LegalizerError as e:
            out["error"] = str(e)

    th = threading.Thread(target=waiter)
    th.start()
    import time
    time.sleep(0.3)
    srv.close()
    th.join(timeout=60)
    assert not th.is_alive()
    assert "error" in out and "stopped" in out["error"], out
    print(f"[case] destroy while waiting: {out}", flush=True)
