"""GPU parity of the pin-access and edge-spacing terms (ICCAD 2017
technology, AdmmKernels.hpp:193-331, 803-830) against the oracle.

The device evaluates the terms in its one-wave partition launches, with the
tile's edge maps and power metal in LDS and the pixels beside the partition
read from the maps as they stood when the step began (the oracle's rule).
Positions, rotation flags, pixel usage, dual bits, the overflow trace and
the checkPlace counts (pin shorts, pin access, edge-spacing pairs) must
equal the oracle's.
"""
import numpy as np
import pytest

from admm_legalization_amd import lg, synth
from admm_legalization_amd.design import strictly_legal
from oracle import ref

pytestmark = pytest.mark.gpu


def _case(n, seed, util=0.8, fences=1, x_hint=None, **kw):
    d = synth.mixed(n, util=util, seed=seed, num_fences=fences, name=f"tech{seed}")
    p = lg.default_params()
    if x_hint:
        p.x_hint = x_hint
    g, c, p2 = lg.device_inputs(d, p)
    synth.with_tech(g, c, seed=seed, **kw)
    return g, c, p2


def _both(g, c, p, do_post=True):
    print(f"[case] tech n={len(c['w_grid'])} grid={g['x_size']}x{g['y_size']} x_hint={p.x_hint} "
          f"qp={p.qp_admm_on} adaptive={p.use_adaptive_search} max_iter={p.max_admm_iter}", flush=True)
    dev = lg.DeviceLegalizer(g, c, p)
    dev.solve()
    usage, dual = dev.pixel_state()
    if do_post:
        dev.postprocess()
    lx, ly, rot = dev.positions()
    st = dev.stats()
    tr = dev.trace()
    dev.close()
    o = ref.solve(g, c, p, do_post=do_post)
    assert st["iterations"] == o["stats"]["iterations"]
    np.testing.assert_array_equal(tr, o["trace"])
    np.testing.assert_array_equal(usage, o["usage"])
    np.testing.assert_array_equal(dual.view(np.int32), o["dual"].view(np.int32))
    np.testing.assert_array_equal(lx, o["lx"])
    np.testing.assert_array_equal(ly, o["ly"])
    np.testing.assert_array_equal(rot, o["rot"])
    assert st["vio"] == o["stats"]["vio"], (st["vio"], o["stats"]["vio"])
    return st


@pytest.mark.parametrize("seed", [0, 1])
def test_tech_matches_oracle(gpu, seed):
    st = _both(*_case(4000, seed))
    assert st["converged"] == 1 and strictly_legal(st["vio"])


def test_tech_many_partition_columns(gpu):
    """x_hint 40: most candidates' spacing windows reach past the partition,
    where the snapshot of the step's start is read."""
    g, c, p = _case(6000, 3, x_hint=40, rules=(0, 0, 3, 4, 2))
    st = _both(g, c, p)
    assert st["converged"] == 1


def test_tech_partial_iterations(gpu):
    g, c, p = _case(3000, 4)
    p.max_admm_iter = 25
    _both(g, c, p, do_post=False)


def test_tech_qp_refine(gpu):
    g, c, p = _case(3000, 6)
    p.qp_admm_on = 1
    st = _both(g, c, p)
    assert st["converged"] == 1


def test_tech_adaptive_search(gpu):
    """The adaptive search moves cells anywhere and carries their edges."""
    g, c, p = _case(3000, 7, util=0.88)
    p.use_adaptive_search = 1
    p.min_iter_to_call_adaptive_search = 20
    p.cooltime_to_call_adaptive_search = 30
    p.ovf_ratio_to_call_adaptive_search = 0.5
    st = _both(g, c, p)
    assert st["converged"] == 1 and st["num_adaptive_calls"] > 0


def test_server_refuses_tech(gpu):
    from admm_legalization_amd.server import LegalizationServer
    g, c, p = _case(500, 8, fences=0)
    with LegalizationServer(max_slots=2, iters_per_sync=4, mem_budget_bytes=1 << 30) as srv:
        with pytest.raises(lg.LegalizerError):
            srv.submit(g, c, p)
