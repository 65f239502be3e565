"""CPU pins of the density oracle (oracle/ref_density.py) and of the
product's host bin setup (dens_prepare).

The reference holds no fixtures for this path.  The transforms are pinned
by brute force: the step-by-step restatement of the reference's FFT chain
(CudaDct.cu) must equal the closed-form cosine / sine sums at small sizes,
and the closed forms must equal scipy's DCT / DST routines.  The Poisson
solve is pinned by single-mode densities whose potential and field follow
from orthogonality; the scatter by conservation of area.
"""
import numpy as np
import pytest
import scipy.fft as sf

from admm_legalization_amd import density as D
from admm_legalization_amd import synth
from oracle import ref_density as R


# -- transforms ---------------------------------------------------------------
@pytest.mark.parametrize("n", [4, 8, 16, 32])
def test_fft_chain_equals_closed_form(n):
    x = np.random.default_rng(n).standard_normal((n, n))
    scale = np.abs(R.idct2(x)).max()
    np.testing.assert_allclose(R.dct2_chain(x), R.dct2(x), atol=1e-13)
    np.testing.assert_allclose(R.idct2_chain(x), R.idct2(x), atol=1e-12 * scale)
    np.testing.assert_allclose(R.idxst_idct_chain(x), R.idxst_idct(x), atol=1e-12 * scale)
    np.testing.assert_allclose(R.idct_idxst_chain(x), R.idct_idxst(x), atol=1e-12 * scale)


@pytest.mark.parametrize("M,N", [(8, 8), (16, 32), (64, 16)])
def test_closed_forms_equal_scipy(M, N):
    x = np.random.default_rng(M * N).standard_normal((M, N))
    np.testing.assert_allclose(R.dct2(x), sf.dctn(x, type=2) / (M * N), atol=1e-12)
    np.testing.assert_allclose(R.idct2(x), sf.dctn(x, type=3), atol=1e-10)
    # IDXST along x: 2 sum_{j>=1} x_j sin(pi j (2n+1) / 2N) is scipy's DST-III
    # of x shifted down by one index (the last input slot zero)
    col = sf.dct(x, type=3, axis=0)  # cosine synthesis along y
    cs = np.zeros_like(col)
    cs[:, :-1] = col[:, 1:]
    np.testing.assert_allclose(R.idxst_idct(x), sf.dst(cs, type=3, axis=1), atol=1e-10)
    row = sf.dct(x, type=3, axis=1)
    rs = np.zeros_like(row)
    rs[:-1, :] = row[1:, :]
    np.testing.assert_allclose(R.idct_idxst(x), sf.dst(rs, type=3, axis=0), atol=1e-10)


@pytest.mark.parametrize("u,v", [(1, 1), (3, 2), (0, 5)])
def test_poisson_single_mode(u, v):
    """rho = cos(pi u (2m+1)/2M) cos(pi v (2n+1)/2N) has one coefficient
    a[u,v] = c (c = 1 for u, v >= 1, 2 if one index is 0), so
    psi = c_u c_v c rho / (wy^2 + wx^2), and the fields follow likewise."""
    M = N = 16
    m = np.arange(M)[:, None]
    n = np.arange(N)[None, :]
    cu = np.cos(np.pi * u * (2 * m + 1) / (2 * M))
    cv = np.cos(np.pi * v * (2 * n + 1) / (2 * N))
    su = np.sin(np.pi * u * (2 * m + 1) / (2 * M))
    sv = np.sin(np.pi * v * (2 * n + 1) / (2 * N))
    rho = cu * cv
    wx, wy = R.frequencies(M, N)
    a = 4.0 / (M * N) * (M / (1 if u == 0 else 2)) * (N / (1 if v == 0 else 2))
    a /= wy[u] ** 2 + wx[v] ** 2
    wu, wv = (1 if u == 0 else 2), (1 if v == 0 else 2)
    psi, ex, ey = R.poisson(rho)
    np.testing.assert_allclose(psi, wu * wv * a * cu * cv, atol=1e-12)
    np.testing.assert_allclose(ex, wu * 2 * a * wx[v] * cu * sv * (v > 0), atol=1e-12)
    np.testing.assert_allclose(ey, 2 * wv * a * wy[u] * su * cv * (u > 0), atol=1e-12)


def test_poisson_constant_density_has_no_field():
    psi, ex, ey = R.poisson(np.full((32, 32), 0.7))
    for f in (psi, ex, ey):
        assert np.abs(f).max() < 1e-12


def test_poisson_chain_and_closed_form_agree():
    rho = np.random.default_rng(1).random((32, 32))
    for a, b in zip(R.poisson(rho), R.poisson(rho, chain=True)):
        np.testing.assert_allclose(a, b, atol=1e-10 * np.abs(a).max())


# -- bins, density sizes, fixed area ---------------------------------------------
def _sum_f32(a):
    return np.add.accumulate(np.asarray(a, np.float32))[-1]


def test_create_bins_rule_1m_gives_1024():
    """A 1M-cell die at 70 % fill (mean area 4.53) gets 1024 x 1024 bins."""
    g = R.create_bins((0, 0, 2545, 2545), 4.53, 1.0)
    assert (g["num_bin_x"], g["num_bin_y"]) == (1024, 1024)


@pytest.mark.parametrize("die,expect", [((0, 0, 400, 100), (128, 32)), ((0, 0, 100, 250), (64, 128)),
                                        ((0, 0, 64, 64), (32, 32))])
def test_create_bins_aspect_ratio(die, expect):
    """Hand-evaluated createBins cases: 400x100 (ratio 4, ideal 100 bins per
    side) rounds down to 32 x 4; 100x250 (ratio 2, ideal 80) stays at 64 x 2;
    64x64 (ideal exactly 32) gives 32."""
    g = R.create_bins(die, 4.0, 1.0)
    assert (g["num_bin_x"], g["num_bin_y"]) == expect
    assert g["bin_ux"][-1] == np.float32(die[2]) and g["bin_uy"][-1] == np.float32(die[3])


@pytest.mark.parametrize("seed,target", [(0, 1.0), (1, 0.8)])
def test_dens_prepare_matches_oracle(seed, target):
    pl = synth.density_placement(3000, seed=seed, num_fixed=4)
    g = D.prepare(pl["die"], pl["dx"], pl["dy"], fixed_rects=pl["fixed_rects"], target_density=target)
    assert g.sum_movable_area == _sum_f32(pl["dx"] * pl["dy"])
    o = R.create_bins(pl["die"], np.float32(g.sum_movable_area) / np.float32(3000), target)
    assert (o["num_bin_x"], o["num_bin_y"]) == (g.num_bin_x, g.num_bin_y)
    for k in ("bin_lx", "bin_ux", "bin_ly", "bin_uy"):
        np.testing.assert_array_equal(o[k], getattr(g, k), err_msg=k)
    dw, dh, sc = R.density_sizes(pl["dx"], pl["dy"], g.bin_w, g.bin_h)
    np.testing.assert_array_equal(dw, g.cell_dw)
    np.testing.assert_array_equal(dh, g.cell_dh)
    np.testing.assert_array_equal(sc, g.cell_scale)
    fa = R.fixed_area(o, pl["fixed_rects"], target).reshape(-1)
    np.testing.assert_allclose(g.fixed_area, fa, rtol=1e-5, atol=1e-5 * fa.max())


def test_density_sizes_conserve_area():
    dx = np.array([0.5, 1.0, 4.0, 10.0], np.float32)
    dy = np.array([0.5, 6.0, 1.0, 10.0], np.float32)
    dw, dh, sc = R.density_sizes(dx, dy, 2.0, 2.0)
    np.testing.assert_allclose(dw * dh * sc, dx * dy, rtol=1e-6)
    assert (dw >= np.float32(2.0 * np.sqrt(2))).all() and (dh >= np.float32(2.0 * np.sqrt(2))).all()


# -- scatter / gather -------------------------------------------------------------
def test_scatter_conserves_area_inside_die():
    pl = synth.density_placement(2000, seed=4)
    geom = R.bins_from_counts(pl["die"], 32, 32)
    dw, dh, sc = R.density_sizes(pl["dx"], pl["dy"], geom["bin_w"], geom["bin_h"])
    lx, ux = pl["cx"] - dw / 2, pl["cx"] + dw / 2
    ly, uy = pl["cy"] - dh / 2, pl["cy"] + dh / 2
    die = pl["die"]
    inside = (lx >= die[0]) & (ux <= die[2]) & (ly >= die[1]) & (uy <= die[3])
    mov, fil = R.scatter(geom, pl["cx"][inside], pl["cy"][inside], dw[inside], dh[inside], sc[inside])
    assert fil.sum() == 0
    area = dw.astype(np.float64) * dh.astype(np.float64) * sc.astype(np.float64)
    np.testing.assert_allclose(mov.sum(), area[inside].sum(), rtol=1e-12)


def test_scatter_brute_force_tiny():
    geom = R.bins_from_counts((0, 0, 8, 8), 4, 4)
    cx, cy = np.array([1.0, 4.5, 7.9]), np.array([1.0, 3.0, 7.5])
    w, h, s = np.array([1.0, 3.0, 2.0]), np.array([2.0, 1.0, 2.0]), np.array([1.0, 0.5, 1.0])
    mov, _ = R.scatter(geom, cx, cy, w, h, s)
    want = np.zeros((4, 4))
    for i in range(3):
        for by in range(4):
            for bx in range(4):
                ox = max(0.0, min(cx[i] + w[i] / 2, 2 * bx + 2) - max(cx[i] - w[i] / 2, 2 * bx))
                oy = max(0.0, min(cy[i] + h[i] / 2, 2 * by + 2) - max(cy[i] - h[i] / 2, 2 * by))
                want[by, bx] += ox * oy * s[i]
    np.testing.assert_allclose(mov, want, atol=1e-12)


def test_gather_of_uniform_field_is_area_times_field():
    """With E constant and lambda = 1 the gradient of a cell inside the die
    is 0.5 * E * (its density area x scale)."""
    geom = R.bins_from_counts((0, 0, 16, 16), 8, 8)
    cx, cy = np.array([3.0, 9.5]), np.array([4.0, 11.0])
    w, h, s = np.array([3.0, 2.0]), np.array([3.0, 5.0]), np.array([0.5, 1.0])
    ex, ey = np.full((8, 8), 2.0), np.full((8, 8), -1.0)
    gx, gy, pc = R.gather(geom, cx, cy, w, h, s, ex, ey)
    np.testing.assert_allclose(pc, w * h * s)
    np.testing.assert_allclose(gx, 0.5 * 2.0 * w * h * s)
    np.testing.assert_allclose(gy, 0.5 * -1.0 * w * h * s)


# -- committed golden vectors (scipy FP64, tests/golden/make_golden.py) ----------
@pytest.mark.parametrize("n", [16, 64])
def test_oracle_poisson_matches_golden(n):
    import os
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", f"poisson_{n}x{n}.npz"))
    psi, ex, ey = R.poisson(z["rho"].astype(np.float64))
    for mine, want in ((psi, z["psi"]), (ex, z["ex"]), (ey, z["ey"])):
        np.testing.assert_allclose(mine, want, atol=1e-10 * np.abs(want).max())
    psi, ex, ey = R.poisson(z["rho"].astype(np.float64), chain=True)
    for mine, want in ((psi, z["psi"]), (ex, z["ex"]), (ey, z["ey"])):
        np.testing.assert_allclose(mine, want, atol=1e-10 * np.abs(want).max())
