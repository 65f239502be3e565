"""The global-placement oracle (oracle/ref_gp.py) pinned by what fixes it
independently of the product: the C++ standard's mt19937 values, closed
forms of the lambda / gamma schedules and of the Lipschitz estimate, a
quadratic objective that Nesterov solves in one step, hand-made filler
counts, the statistics of the initial placement, and std::round's halves."""
import math

import numpy as np
import pytest
from scipy import stats

from admm_legalization_amd import synth
from oracle import ref_gp as G

PARAMS = dict(max_opt_iter=2500, max_backtrack_iter=10, init_lambda=1e-4, init_gamma_inv_coef=0.01,
              target_overflow=0.07, min_phi_coef=0.95, max_phi_coef=1.05, min_precond=1.0, init_opt_coef=100.0,
              reference_hpwl=200000.0, min_step_length=1.0, adam_alpha=100.0, adam_beta1=0.9, adam_beta2=0.999,
              adam_epsilon=1e-8)


def test_mt19937_stream_is_the_standard_one():
    """std::mt19937(0) starts 2357136044, 2546248239, 3071714933, and the
    10000th value of a default-seeded (5489) std::mt19937 is 4123659995
    ([rand.predef] in the C++ standard): the filler centres' generator."""
    raw = np.random.RandomState(0).randint(0, 2 ** 32, size=3, dtype=np.uint64)
    assert list(raw) == [2357136044, 2546248239, 3071714933]
    raw = np.random.RandomState(5489).randint(0, 2 ** 32, size=10000, dtype=np.uint64)
    assert raw[-1] == 4123659995


def test_fillers_closed_form():
    """Ten cells of widths 1..10 (shuffled) and heights 2: the [1, 9) slice of
    the sorted widths averages (2 + .. + 9) / 8 = 5.5; the heights are taken
    unsorted from the same slice (all 2 here); die 100 x 100 with a 20 x 10
    fixed block, density 0.9: (10000 - 200) x 0.9 - 110 = 8710, over 11 per
    filler = 791 fillers; centres = mt19937 values mod die side."""
    w = np.array([7, 3, 10, 1, 5, 9, 2, 8, 6, 4], np.float32)
    h = np.full(10, 2, np.float32)
    fw, fh, cnt, fx, fy = G.create_fillers(w, h, (0, 0, 100, 100), 200.0, float((w * h).sum()), 0.9)
    assert fw == 5.5 and fh == 2.0 and cnt == 791
    raw = np.random.RandomState(0).randint(0, 2 ** 32, size=4, dtype=np.uint64)
    assert (fx[0], fy[0], fx[1], fy[1]) == tuple(float(v % 100) for v in raw)
    # the height average follows the unsorted list: a tall cell at index 0
    h2 = h.copy()
    h2[0] = 10
    _, fh2, _, _, _ = G.create_fillers(w, h2, (0, 0, 100, 100), 0.0, float((w * h2).sum()), 1.0)
    assert fh2 == 2.0  # index 0 is outside the [1, 9) slice whatever the widths' order
    h2 = h.copy()
    h2[1] = 10
    _, fh3, _, _, _ = G.create_fillers(w, h2, (0, 0, 100, 100), 0.0, float((w * h2).sum()), 1.0)
    assert fh3 == 3.0
    with pytest.raises(ValueError):
        G.create_fillers(w, h, (0, 0, 10, 10), 0.0, 150.0, 1.0)


def test_gamma_and_lambda_schedules():
    c = 0.01
    assert G.gamma_inv(1.5, c) == np.float32(0.1) * np.float32(c)
    assert G.gamma_inv(0.05, c) == np.float32(10.0) * np.float32(c)
    # overflow 0.55: 20 * 0.45 / 9 - 1 = 0 -> 1 / 10^0 = 1
    assert G.gamma_inv(0.55, c) == pytest.approx(c, rel=1e-6)
    # overflow 1.0: 20 * 0.9 / 9 - 1 = 1 -> 0.1, continuous with the > 1 branch
    assert G.gamma_inv(1.0, c) == pytest.approx(0.1 * c, rel=1e-6)
    # overflow 0.1: 10^-(-1) = 10, continuous with the < 0.1 branch
    assert G.gamma_inv(0.1, c) == pytest.approx(10.0 * c, rel=1e-5)
    ref = 200000.0
    assert G.lambda_coef(100.0, 50.0, ref, 0.95, 1.05) == np.float32(1.05)  # HPWL fell
    assert G.lambda_coef(100.0, 100.0, ref, 0.95, 1.05) == np.float32(1.05)  # p = 0: 1.05 * 1.05^0
    assert G.lambda_coef(0.0, ref, ref, 0.95, 1.05) == pytest.approx(1.0, rel=1e-6)  # p = 1: 1.05^0
    assert G.lambda_coef(0.0, 10 * ref, ref, 0.95, 1.05) == np.float32(0.95)  # clamped below


def test_lipschitz_of_a_quadratic():
    """For grad = -c x the estimate |dx| / |dgrad| is 1 / c whatever the points."""
    rng = np.random.default_rng(0)
    a, b = rng.normal(size=1000), rng.normal(size=1000)
    c = 3.5
    assert G.lipschitz(a, -c * a, b, -c * b) == pytest.approx(1 / c, rel=1e-6)


class _Quadratic(G.Objective):
    """-grad of 0.5 c |x - x*|^2 (a descent direction, as the placer's),
    HPWL = |x - x*|_1 and overflow 0.05 so that the loop stops."""

    def __init__(self, pl, target, c):
        super().__init__(pl)
        self.t, self.c = target, c

    def wirelength(self, var, g_inv):
        return -self.c * (np.asarray(var, np.float64) - self.t), float(np.abs(var - self.t).sum())

    def density(self, var):
        n = len(var) // 2
        return np.ones(len(var)), np.zeros(n), 0.05

    def precondition(self, lam, wl, dens, pc):
        return wl.astype(np.float32)  # the density term left out


def test_nesterov_solves_a_quadratic_in_one_step():
    """initSolver's Lipschitz step on a quadratic is exactly 1 / c, and the
    first forward move from the initial point lands on the minimiser (the
    momentum coefficient of the first iteration is (1 - 1) / a = 0)."""
    d = synth.e2e(300, seed=2)
    pl = G.Placement(d)
    n = pl.n
    rng = np.random.default_rng(1)
    tx = rng.uniform(pl.die[0] + 5000, pl.die[2] - 5000, n)
    ty = rng.uniform(pl.die[1] + 5000, pl.die[3] - 5000, n)
    target = np.concatenate([tx, ty]).astype(np.float32).astype(np.float64)
    x0 = (target + rng.normal(0, 300, 2 * n)).astype(np.float32)
    s = G.Nesterov(pl, PARAMS, x0)
    s.obj = _Quadratic(pl, target, 0.25)
    s.obj.min_precond = 1.0
    s.init_solver()
    assert float(s.step) == pytest.approx(4.0, rel=1e-5)
    s.iterate()
    np.testing.assert_allclose(s.cur_var, target, rtol=0, atol=2e-3 * 300)


def test_initial_placement_statistics():
    """Movable centres ~ N(die centre, 0.05 x side) in x and y, independent;
    fillers mirrored to the die edges; everything inside the die."""
    n, n_mov = 200000, 150000
    die = (0.0, 0.0, 1.0e6, 5.0e5)
    w = np.full(n, 10.0, np.float32)
    h = np.full(n, 8.0, np.float32)
    x, y = G.initial_place(n, n_mov, 7, die, w, h)
    mx, my = x[:n_mov].astype(np.float64), y[:n_mov].astype(np.float64)
    assert abs(mx.mean() - 5e5) < 4 * 5e4 / math.sqrt(n_mov)
    assert abs(my.mean() - 2.5e5) < 4 * 2.5e4 / math.sqrt(n_mov)
    assert mx.std() == pytest.approx(5e4, rel=0.01) and my.std() == pytest.approx(2.5e4, rel=0.01)
    assert stats.kstest((mx - 5e5) / 5e4, "norm").pvalue > 1e-3
    assert abs(np.corrcoef(mx, my)[0, 1]) < 0.02
    fx = x[n_mov:].astype(np.float64)
    assert np.all((fx <= 5e4 * 8) | (fx >= 1e6 - 5e4 * 8))  # mirrored: near the edges
    assert x.min() >= 5 and x.max() <= 1e6 - 5 and y.min() >= 4 and y.max() <= 5e5 - 4
    x2, _ = G.initial_place(n, n_mov, 8, die, w, h)
    assert not np.array_equal(x, x2)


def test_inside_and_export_rounding():
    assert list(G.inside(np.array([1.0, 50.0, 99.0], np.float32), 10.0, 0.0, 100.0)) == [5.0, 50.0, 95.0]
    d = synth.e2e(300, seed=2)
    pl = G.Placement(d)
    n = pl.n
    var = np.zeros(2 * n, np.float32)
    m = pl.n_mov
    var[:m] = pl.w[:m] / 2 + np.float32(10.5)
    var[n:n + m] = pl.h[:m] / 2 + np.float32(2.5)
    lx, ly = G.export_positions(d, pl, var)
    assert np.all(lx[pl.imp["mov"]] == 11) and np.all(ly[pl.imp["mov"]] == 3)
    fixed = np.asarray(d.inst_fixed) != 0
    assert np.array_equal(lx[fixed], d.inst_lx[fixed])


def test_import_rejects_target_density_below_utilisation():
    d = synth.e2e(300, seed=2, util=0.9)
    with pytest.raises(ValueError):
        G.Placement(d, target_density=0.85)


def test_oracle_loop_converges_on_a_small_design():
    """The whole oracle loop on 600 cells: overflow reaches the 0.07 target,
    the HPWL of the last iterations grows as the cells spread (checkConvergence
    then stops it), lambda grows."""
    d = synth.e2e(600, seed=5)
    pl = G.Placement(d)
    x, y = G.initial_place(pl.n, pl.n_mov, 1, pl.die, pl.w, pl.h)
    s = G.Nesterov(pl, PARAMS, np.concatenate([x, y])).solve()
    assert s.done_converged and s.ovf <= 0.07
    tr = np.array(s.trace)
    assert tr[-1, 2] > tr[0, 2] * 100
    assert tr[0, 1] > 0.5
