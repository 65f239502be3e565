"""Pins of the wirelength oracle (oracle/ref_wl.py) by what the mathematics
fixes: the reference's pin term is minus the derivative of the WA
wirelength (central differences), two-pin nets tend to |x1 - x2| with unit
forces, the terms of one net sum to zero (translation invariance), weights
scale linearly, HPWL and preconditioning by hand.  No device needed."""
import numpy as np
import pytest

from admm_legalization_amd import synth
from oracle import ref_wl as W


def _random_netlist(seed, cells=40, nets=60, max_deg=12):
    rng = np.random.default_rng(seed)
    deg = rng.integers(2, max_deg + 1, nets)
    ns = np.concatenate([[0], np.cumsum(deg)])
    x = rng.uniform(0, 100, ns[-1])
    y = rng.uniform(0, 60, ns[-1])
    p2c = rng.integers(-1, cells, ns[-1])
    w = rng.uniform(0.5, 2.0, nets)
    return ns, x, y, p2c, w


@pytest.mark.parametrize("g_inv", [0.02, 0.3, 2.0])
def test_pin_term_is_minus_the_wa_derivative(g_inv):
    ns, x, _, _, w = _random_netlist(1)
    nid = np.repeat(np.arange(len(ns) - 1), np.diff(ns))
    g = W.wa_grad_axis(ns, x, g_inv, w)
    h = 1e-5
    for p in range(0, len(x), 7):
        xp, xm = x.copy(), x.copy()
        xp[p] += h
        xm[p] -= h
        n = nid[p]
        d = (W.wa_wirelength(ns, xp, g_inv)[n] - W.wa_wirelength(ns, xm, g_inv)[n]) / (2 * h)
        assert g[p] == pytest.approx(-w[n] * d, rel=1e-6, abs=1e-8)


def test_two_pin_net_tends_to_distance_and_unit_forces():
    ns = np.array([0, 2])
    x = np.array([10.0, 35.0])
    wl = W.wa_wirelength(ns, x, 5.0)[0]
    assert wl == pytest.approx(25.0, rel=1e-12)
    g = W.wa_grad_axis(ns, x, 5.0, np.ones(1))
    np.testing.assert_allclose(g, [1.0, -1.0], atol=1e-12)
    # small gInv: the WA estimate shrinks towards 0 (smooth regime)
    assert 0 < W.wa_wirelength(ns, x, 1e-3)[0] < 1.0


def test_net_terms_sum_to_zero_and_weights_scale():
    ns, x, _, _, w = _random_netlist(2)
    g = W.wa_grad_axis(ns, x, 0.4, w)
    nid = np.repeat(np.arange(len(ns) - 1), np.diff(ns))
    np.testing.assert_allclose(np.bincount(nid, g), 0.0, atol=1e-12)
    g2 = W.wa_grad_axis(ns, x, 0.4, 3.0 * w)
    np.testing.assert_allclose(g2, 3.0 * g, rtol=1e-14)
    # translation of a whole net leaves its terms unchanged
    g3 = W.wa_grad_axis(ns, x + 17.0, 0.4, w)
    np.testing.assert_allclose(g3, g, rtol=1e-8, atol=1e-12)


def test_cell_sums_and_io_pins():
    ns, x, y, p2c, w = _random_netlist(3)
    gx, gy = W.wa_grad(ns, x, y, p2c, 0.5, w, 40)
    px = W.wa_grad_axis(ns, x, 0.5, w)
    for c in (0, 5, 39):
        assert gx[c] == pytest.approx(px[p2c == c].sum(), abs=1e-12)
    # IO pins (-1) contribute to their net but to no cell
    assert gx.shape == (40,)


def test_update_pins_keeps_io_pins():
    x, y = np.array([1.0, 2.0, 3.0]), np.array([4.0, 5.0, 6.0])
    p2c = np.array([1, -1, 0])
    px, py = W.update_pins(x, y, p2c, [0.5, 9.0, -0.5], [0.25, 9.0, 0.0], [10.0, 20.0], [30.0, 40.0])
    np.testing.assert_array_equal(px, [20.5, 2.0, 9.5])
    np.testing.assert_array_equal(py, [40.25, 5.0, 30.0])


def test_hpwl_by_hand_with_the_reference_initial_values():
    ns = np.array([0, 3, 5])
    x = np.array([1.0, 4.0, 2.0, -3.0, -1.0])
    y = np.array([0.0, 2.0, 5.0, 1.0, 1.5])
    # net 0: (4-1) + (5-0); net 1: max starts at 0 -> (0 - -3) + (1.5 - 1)
    assert W.hpwl(ns, x, y) == pytest.approx(3 + 5 + 3 + 0.5)


def test_precondition_by_hand():
    gx, gy = W.precondition(2.0, 1.0, [3.0, 0.0], [0.5, 0.1], [1.0, 2.0], [3.0, 4.0], [0.5, 0.5], [1.0, 1.0],)
    # cell 0: (1 + 2*0.5) / (3 + 1) ; cell 1: max(0 + 0.2, 1) = 1
    np.testing.assert_allclose(gx, [2.0 / 4.0, 3.0 / 1.0])
    np.testing.assert_allclose(gy, [5.0 / 4.0, 6.0 / 1.0])


def test_synthetic_netlist_degrees():
    nl = synth.netlist(np.ones(1000), np.ones(1000), 20000, die=(0, 0, 100, 100), seed=0)
    d = np.diff(nl["net_start"])
    assert d.min() >= 2 and d.max() <= 1000
    assert abs(d.mean() - 3.8) < 0.2
    assert nl["pin2cell"].max() < 1000 and (nl["pin2cell"] == -1).mean() < 0.02
