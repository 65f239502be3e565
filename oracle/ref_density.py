"""ORACLE (test infrastructure only): FP64 restatement of the electrostatic
density path of the reference placer (skyplace/src/density_gradient/ and
SkyPlaceDB.cpp).

Imported only by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
leg, as the checker.  The product never imports this module.

Contents
  create_bins            SkyPlaceDB::createBins (SkyPlaceDB.cpp:446-549), FP32 as the reference
  density_sizes          SkyPlaceDB::updateDensitySize (SkyPlaceDB.cpp:821-852)
  fixed_area             SkyPlaceDB::updateFixedOverlapArea (SkyPlaceDB.cpp:909-938)
  scatter                computeDensityContributionOfEachCell (DensityGradient.cu:61-138), FP64
  overflow               computeOverflow + computeVectorSum (DensityGradient.cu:140-156, 360)
  dct2_chain / idct2_chain / idxst_idct_chain / idct_idxst_chain
                         the reference's FFT-based transforms (CudaDct.cu:8-506) restated step
                         by step in FP64 with numpy's FFT in place of cuFFT R2C / C2R
  dct2 / idct2 / idxst_idct / idct_idxst
                         the same transforms as closed-form matrix products (FP64); the
                         tests pin the chains to these by brute force
  poisson                PoissonSolver::solvePoisson (PoissonSolver.cu:224-302)
  gather                 getDensityGradForEachCell (DensityGradient.cu:173-281), FP64

Array conventions: bins are y-major, bin = x + y * num_bin_x
(DensityGradient.cu:130); 2-D maps are numpy arrays of shape
(num_bin_y, num_bin_x).
"""
from __future__ import annotations

import math

import numpy as np

SQRT2 = 1.414213562373095048801  # DENSITY_SCALE (SkyPlaceDB.h:9-10)


# ---------------------------------------------------------------------------
# D1: bins, density sizes, fixed area (host, FP32 like the reference)
# ---------------------------------------------------------------------------
def create_bins(die, avg_std_cell_area: float, target_density: float = 1.0):
    """createBins (SkyPlaceDB.cpp:446-549).  die = (lx, ly, ux, uy).
    Returns dict(num_bin_x, num_bin_y, bin_w, bin_h, bin_lx, bin_ux, bin_ly,
    bin_uy) with the FP32 bin bounds the reference accumulates."""
    f = np.float32
    lx, ly, ux, uy = (f(v) for v in die)
    die_area = f((ux - lx) * (uy - ly))
    ideal_bin_area = f(f(avg_std_cell_area) / f(target_density))
    ideal_bin_count = f(die_area / ideal_bin_area)
    ideal_num_bin = int(math.ceil(float(np.sqrt(ideal_bin_count, dtype=np.float32))))
    dx, dy = f(ux - lx), f(uy - ly)
    aspect = max(float(f(dx / dy)), float(f(dy / dx)))
    ratio = 4 if aspect >= 4.0 else (2 if aspect >= 2.0 else 1)
    nb = 4
    while nb * nb * ratio < ideal_num_bin * ideal_num_bin:
        nb *= 2
        if nb >= 1024:
            break
    e1 = abs(ideal_num_bin * ideal_num_bin - nb * nb * ratio)
    e2 = abs(ideal_num_bin * ideal_num_bin - nb * nb * ratio // 4)
    if e2 < e1:
        nb //= 2
    nbx, nby = (nb * ratio, nb) if dx > dy else (nb, nb * ratio)
    return bins_from_counts(die, nbx, nby)


def bins_from_counts(die, nbx: int, nby: int):
    """Bin bounds for explicit counts, accumulated in FP32 in creation order
    (SkyPlaceDB.cpp:504-546): widths are die/num except that the last column
    (row) ends at the die edge."""
    f = np.float32
    lx0, ly0, ux, uy = (f(v) for v in die)
    bw = f(f(ux - lx0) / f(nbx))
    bh = f(f(uy - ly0) / f(nby))
    bin_lx = np.zeros(nbx, np.float32)
    bin_ux = np.zeros(nbx, np.float32)
    bin_ly = np.zeros(nby, np.float32)
    bin_uy = np.zeros(nby, np.float32)
    x = lx0
    for c in range(nbx):  # every row repeats the same column walk
        w = bw
        if c == nbx - 1 or f(x + bw) > ux:
            w = f(ux - x)
        bin_lx[c], bin_ux[c] = x, f(x + w)
        x = f(x + w)
    y = ly0
    for r in range(nby):
        h = bh
        if r == nby - 1 or f(y + bh) > uy:
            h = f(uy - y)
        bin_ly[r], bin_uy[r] = y, f(y + h)
        y = f(y + h)
    return dict(num_bin_x=nbx, num_bin_y=nby, bin_w=float(bw), bin_h=float(bh), bin_lx=bin_lx,
                bin_ux=bin_ux, bin_ly=bin_ly, bin_uy=bin_uy, die=tuple(float(v) for v in die))


def density_sizes(dx, dy, bin_w: float, bin_h: float):
    """updateDensitySize (SkyPlaceDB.cpp:821-852): cells narrower (shorter)
    than sqrt(2) bins are widened to sqrt(2) bins and their density scaled by
    the area ratio.  FP32 as the reference."""
    f = np.float32
    ld = np.longdouble
    dx = np.asarray(dx, np.float32)
    dy = np.asarray(dy, np.float32)
    # DENSITY_SCALE is a long double literal: threshold and ratio in long double
    s2 = ld("1.414213562373095048801")
    tx = s2 * ld(f(bin_w))
    ty = s2 * ld(f(bin_h))
    wide = dx.astype(ld) > tx
    tall = dy.astype(ld) > ty
    sx = np.where(wide, f(1.0), (dx.astype(ld) / tx).astype(np.float32))
    sy = np.where(tall, f(1.0), (dy.astype(ld) / ty).astype(np.float32))
    dw = np.where(wide, dx, f(tx)).astype(np.float32)
    dh = np.where(tall, dy, f(ty)).astype(np.float32)
    return dw, dh, (sx * sy).astype(np.float32)


def fixed_area(geom, fixed_rects, target_density: float = 1.0):
    """updateFixedOverlapArea (SkyPlaceDB.cpp:909-938) in FP64: overlap area of
    each fixed rectangle (lx, ly, ux, uy) with each bin, times the target
    density; rectangles outside the die are skipped."""
    nbx, nby = geom["num_bin_x"], geom["num_bin_y"]
    out = np.zeros((nby, nbx))
    lx0, ly0, ux0, uy0 = geom["die"]
    blx, bux = geom["bin_lx"].astype(np.float64), geom["bin_ux"].astype(np.float64)
    bly, buy = geom["bin_ly"].astype(np.float64), geom["bin_uy"].astype(np.float64)
    for (lx, ly, ux, uy) in np.asarray(fixed_rects, np.float64).reshape(-1, 4):
        if ux <= lx0 or lx >= ux0 or uy <= ly0 or ly >= uy0:
            continue
        ox = np.maximum(0.0, np.minimum(bux, ux) - np.maximum(blx, lx))
        oy = np.maximum(0.0, np.minimum(buy, uy) - np.maximum(bly, ly))
        out += np.outer(oy, ox) * target_density
    return out


# ---------------------------------------------------------------------------
# D3 / D4: scatter and overflow (FP64)
# ---------------------------------------------------------------------------
def _bin_range(c, d, origin, size, nb):
    """findBinWithDensitySize (DensityGradient.cu:20-59): [floor, ceil) of the
    footprint's bin coordinates, clipped to the grid.  The range is a discrete
    choice of which bins a cell visits, and the reference makes it in FP32
    (float positions, sizes and bin size); it is restated in FP32 here.  The
    bin edges themselves are accumulated in FP32 (createBins,
    SkyPlaceDB.cpp:504-546) and drift from i * size by up to ~1e-2 units on
    a 1024-bin side, so an FP64 range would visit a different set of bins
    for cells whose edge lies within that drift of a bin boundary."""
    f = np.float32
    c32 = np.asarray(c).astype(f)
    h32 = np.asarray(d).astype(f) / f(2)
    lo = np.floor(((c32 - h32) - f(origin)) / f(size)).astype(np.int64)
    hi = np.ceil(((c32 + h32) - f(origin)) / f(size)).astype(np.int64)
    return np.maximum(lo, 0), np.minimum(hi, nb)


def scatter(geom, cx, cy, dw, dh, scale, is_filler=None, is_macro=None, target_density=1.0):
    """movable and filler area per bin: sum over cells of the overlap of the
    density footprint with the bin, times the density scale (times the target
    density for macros)."""
    nbx, nby = geom["num_bin_x"], geom["num_bin_y"]
    cx, cy, dw, dh, scale = (np.asarray(v, np.float64) for v in (cx, cy, dw, dh, scale))
    n = len(cx)
    is_filler = np.zeros(n, bool) if is_filler is None else np.asarray(is_filler, bool)
    is_macro = np.zeros(n, bool) if is_macro is None else np.asarray(is_macro, bool)
    s = np.where(is_macro, target_density * scale, scale)
    blx, bux = geom["bin_lx"].astype(np.float64), geom["bin_ux"].astype(np.float64)
    bly, buy = geom["bin_ly"].astype(np.float64), geom["bin_uy"].astype(np.float64)
    x0, x1 = _bin_range(cx, dw, geom["die"][0], geom["bin_w"], nbx)
    y0, y1 = _bin_range(cy, dh, geom["die"][1], geom["bin_h"], nby)
    mov = np.zeros(nby * nbx)
    fil = np.zeros(nby * nbx)
    span_x = int(max(int((x1 - x0).max(initial=0)), 0))
    span_y = int(max(int((y1 - y0).max(initial=0)), 0))
    lx, ux = cx - dw / 2, cx + dw / 2
    ly, uy = cy - dh / 2, cy + dh / 2
    for i in range(span_x):
        bx = x0 + i
        okx = bx < x1
        bxc = np.minimum(bx, nbx - 1)
        ox = np.maximum(0.0, np.minimum(ux, bux[bxc]) - np.maximum(lx, blx[bxc]))
        for j in range(span_y):
            by = y0 + j
            ok = okx & (by < y1)
            byc = np.minimum(by, nby - 1)
            oy = np.maximum(0.0, np.minimum(uy, buy[byc]) - np.maximum(ly, bly[byc]))
            a = np.where(ok, ox * oy * s, 0.0)
            idx = bxc + byc * nbx
            np.add.at(mov, idx[ok & ~is_filler], a[ok & ~is_filler])
            np.add.at(fil, idx[ok & is_filler], a[ok & is_filler])
    return mov.reshape(nby, nbx), fil.reshape(nby, nbx)


def overflow(mov, fil, fixed, scaled_bin_area, sum_movable_area):
    """computeOverflow (DensityGradient.cu:140-156) and the overflow ratio
    (DensityGradient.cu:360): density = (mov + fixed + filler) / scaled area,
    overflow = sum max(0, mov + fixed - scaled area) / sum movable area."""
    dens = (mov + fixed + fil) / scaled_bin_area
    ovf = np.maximum(0.0, mov + fixed - scaled_bin_area).sum() / sum_movable_area
    return dens, float(ovf)


# ---------------------------------------------------------------------------
# D5: transforms.  M = rows (num_bin_y), N = columns (num_bin_x).
# ---------------------------------------------------------------------------
def _expk(L):
    """exp(-i pi k / (2 L)) (precomputeExpk, PoissonSolver.cu:9-29)."""
    k = np.arange(2 * L + 1)
    return np.exp(-1j * np.pi * k / (2 * L))


def _makhoul_order(L):
    """Makhoul's even/odd reordering: position of x[k] in v (v[j] = x[2j],
    v[L-1-j] = x[2j+1]), CudaDct.cu:8-35."""
    k = np.arange(L)
    return np.where(k % 2 == 0, k // 2, L - (k + 1) // 2)


def dct2_chain(x):
    """dct_2d_fft (CudaDct.cu:140-173): reorder, 2-D real FFT, twiddle
    post-process with 2/(MN) on interior coefficients and 4/(MN) on the
    first row / column (dct2d_postprocess, CudaDct.cu:37-132)."""
    x = np.asarray(x, np.float64)
    M, N = x.shape
    v = np.zeros((M, N))
    v[np.ix_(_makhoul_order(M), _makhoul_order(N))] = x
    V = np.fft.rfft2(v)  # (M, N/2+1), the cuFFT R2C layout
    eM, eN = _expk(M), _expk(N)
    hM, hN = M // 2, N // 2
    two, four = 2.0 / (M * N), 4.0 / (M * N)
    y = np.zeros((M, N))

    def re_mul(a, b):
        return (a * b).real

    def im_mul(a, b):
        return (a * b).imag

    # row 0 and row M/2 (hid == 0 cases)
    y[0, 0] = V[0, 0].real * four
    y[0, hN] = re_mul(eN[hN], V[0, hN]) * four
    y[hM, 0] = eM[hM].real * V[hM, 0].real * four
    y[hM, hN] = eM[hM].real * re_mul(eN[hN], V[hM, hN]) * four
    w = np.arange(1, hN)
    y[0, w] = re_mul(eN[w], V[0, w]) * four
    y[0, N - w] = -im_mul(eN[w], V[0, w]) * four
    y[hM, w] = eM[hM].real * re_mul(eN[w], V[hM, w]) * four
    y[hM, N - w] = -eM[hM].real * im_mul(eN[w], V[hM, w]) * four
    # rows 1..M/2-1 paired with M-h
    h = np.arange(1, hM)[:, None]
    ec = eM[h]
    t1, t2 = V[h, 0], V[M - h, 0]
    up = ec.real * (t1.real + t2.real) + ec.imag * (t2.imag - t1.imag)
    down = -ec.imag * (t1.real + t2.real) + ec.real * (t2.imag - t1.imag)
    y[h, 0] = up * two
    y[M - h, 0] = down * two

    def pair(col):
        a, b = V[h, col], V[M - h, col]
        s, d = a + b, a - b
        up = (ec.real * s.real - ec.imag * d.imag) + 1j * (ec.real * s.imag + ec.imag * d.real)
        down = (-ec.imag * s.real - ec.real * d.imag) + 1j * (-ec.imag * s.imag + ec.real * d.real)
        return up, down

    up, down = pair(np.array([hN]))
    y[h, hN] = re_mul(eN[hN], up) * two
    y[M - h, hN] = re_mul(eN[hN], down) * two
    wc = np.arange(1, hN)[None, :]
    up, down = pair(wc)
    y[h, wc] = re_mul(eN[wc], up) * two
    y[M - h, wc] = re_mul(eN[wc], down) * two
    y[h, N - wc] = -im_mul(eN[wc], up) * two
    y[M - h, N - wc] = -im_mul(eN[wc], down) * two
    return y


def _inverse_tables(M, N):
    """precomputeExpkForInverse (PoissonSolver.cu:31-71), FP64; expkMN_2 keeps
    the reference's offset of N-1 in the M-indexed table."""
    k = np.arange(M)
    eM = np.exp(-1j * np.pi * k / (2 * M))
    mn1 = np.concatenate([eM, np.exp(-1j * np.pi * (k + M) / (2 * M))])
    a = np.pi * (k - (N - 1)) / (2 * M)
    b = np.pi * (k - (N - 1) + M) / (2 * M)
    mn2 = np.concatenate([-np.sin(a) - 1j * np.cos(a), -np.sin(b) - 1j * np.cos(b)])
    eN = _expk(N)
    return eM, eN, mn1, mn2


def idct2_chain(x):
    """idct_2d_fft (CudaDct.cu:326-364): twiddle pre-process into a Hermitian
    half spectrum (idct2d_preprocess, CudaDct.cu:208-318), unnormalised 2-D
    inverse real FFT (cuFFT C2R), inverse Makhoul reordering
    (idct2d_postprocess, CudaDct.cu:175-206)."""
    x = np.asarray(x, np.float64)
    M, N = x.shape
    hM, hN = M // 2, N // 2
    eM, eN, mn1, mn2 = _inverse_tables(M, N)
    out = np.zeros((M, hN + 1), np.complex128)

    def mulconj(a, b):  # conj(a * b)
        return np.conj(a * b)

    out[0, 0] = x[0, 0]
    out[0, hN] = mulconj(eN[hN], x[0, hN] + 1j * x[0, hN])
    out[hM, 0] = mulconj(eM[hM], x[hM, 0] + 1j * x[hM, 0])
    out[hM, hN] = mulconj(mn1[hM + hN], 2j * x[hM, hN])
    w = np.arange(1, hN)
    out[0, w] = mulconj(eN[w], x[0, w] + 1j * x[0, N - w])
    t1, t2 = x[hM, w], x[hM, N - w]
    out[hM, w] = mulconj(mn1[hM + w], (t1 - t2) + 1j * (t1 + t2))
    h = np.arange(1, hM)
    t1, t3 = x[h, 0], x[M - h, 0]
    up = eM[M - h] * (t3 + 1j * t1)
    out[h, 0] = up
    out[M - h, 0] = np.conj(up)
    t1, t3 = x[h, hN], x[M - h, hN]
    out[h, hN] = mulconj(mn1[h + hN], (t1 - t3) + 1j * (t3 + t1))
    out[M - h, hN] = mulconj(mn2[hN - h + (N - 1)], (t3 - t1) + 1j * (t1 + t3))
    hh = h[:, None]
    ww = w[None, :]
    t1, t2 = x[hh, ww], x[hh, N - ww]
    t3, t4 = x[M - hh, ww], x[M - hh, N - ww]
    out[hh, ww] = mulconj(mn1[hh + ww], (t1 - t4) + 1j * (t3 + t2))
    out[M - hh, ww] = mulconj(mn2[ww - hh + (N - 1)], (t3 - t2) + 1j * (t1 + t4))
    # cuFFT C2R is unnormalised; numpy's irfft2 divides by M*N
    z = np.fft.irfft2(out, s=(M, N)) * (M * N)
    y = np.zeros((M, N))
    y[np.ix_(_inv_order(M), _inv_order(N))] = z
    return y


def _inv_order(L):
    """idct2d_postprocess: z[k] goes to y[2k] for k < L/2, to y[2(L-k)-1]
    otherwise."""
    k = np.arange(L)
    return np.where(k < L // 2, 2 * k, 2 * (L - k) - 1)


def idxst_idct_chain(x):
    """idxst_idct (CudaDct.cu:441-506): column reversal with column 0 zeroed,
    idct_2d_fft, sign flip of odd columns."""
    x = np.asarray(x, np.float64)
    M, N = x.shape
    pre = np.zeros((M, N))
    pre[:, 1:] = x[:, N - np.arange(1, N)]
    y = idct2_chain(pre)
    y[:, 1::2] *= -1
    return y


def idct_idxst_chain(x):
    """idct_idxst (CudaDct.cu:366-439): row reversal with row 0 zeroed,
    idct_2d_fft, sign flip of odd rows."""
    x = np.asarray(x, np.float64)
    M, N = x.shape
    pre = np.zeros((M, N))
    pre[1:, :] = x[M - np.arange(1, M), :]
    y = idct2_chain(pre)
    y[1::2, :] *= -1
    return y


# closed forms (matrices); k = frequency, j = sample
def cos_matrix(L):
    """C[k, j] = cos(pi k (2j + 1) / (2L))."""
    k = np.arange(L)[:, None]
    j = np.arange(L)[None, :]
    return np.cos(np.pi * k * (2 * j + 1) / (2 * L))


def sin_matrix(L):
    """S[k, j] = sin(pi k (2j + 1) / (2L))."""
    k = np.arange(L)[:, None]
    j = np.arange(L)[None, :]
    return np.sin(np.pi * k * (2 * j + 1) / (2 * L))


def _inv_weights(L):
    c = np.full(L, 2.0)
    c[0] = 1.0
    return c


def dct2(x):
    """a[u, v] = 4/(MN) sum_m sum_n x[m, n] cos(pi u (2m+1)/2M) cos(pi v (2n+1)/2N)."""
    x = np.asarray(x, np.float64)
    M, N = x.shape
    return 4.0 / (M * N) * cos_matrix(M) @ x @ cos_matrix(N).T


def idct2(a):
    """x[m, n] = sum_u sum_v c_u c_v a[u, v] cos(..m..) cos(..n..), c_0 = 1, c_k = 2."""
    a = np.asarray(a, np.float64)
    M, N = a.shape
    DM = _inv_weights(M)[:, None] * cos_matrix(M)
    DN = _inv_weights(N)[:, None] * cos_matrix(N)
    return DM.T @ a @ DN


def idxst_idct(a):
    """Cosine synthesis along rows (y), sine synthesis along columns (x):
    x[m, n] = sum_u sum_{v>=1} c_u 2 a[u, v] cos(..m..) sin(pi v (2n+1)/2N)."""
    a = np.asarray(a, np.float64)
    M, N = a.shape
    DM = _inv_weights(M)[:, None] * cos_matrix(M)
    SN = 2.0 * sin_matrix(N)
    SN[0] = 0.0
    return DM.T @ a @ SN


def idct_idxst(a):
    """Sine synthesis along rows (y), cosine synthesis along columns (x)."""
    a = np.asarray(a, np.float64)
    M, N = a.shape
    SM = 2.0 * sin_matrix(M)
    SM[0] = 0.0
    DN = _inv_weights(N)[:, None] * cos_matrix(N)
    return SM.T @ a @ DN


def frequencies(M, N):
    """w_u = 2 pi u / N along x, w_v = 2 pi v / M along y
    (divideByWSquare / multiplyW, PoissonSolver.cu:73-118)."""
    wx = 2.0 * np.pi * np.arange(N) / N
    wy = 2.0 * np.pi * np.arange(M) / M
    return wx, wy


def poisson(rho, chain: bool = False):
    """solvePoisson (PoissonSolver.cu:224-302): a = DCT(rho); a /= wx^2 + wy^2
    (a[0, 0] = 0); potential = IDCT(a); Ex = IDXST_IDCT(wx a);
    Ey = IDCT_IDXST(wy a).  chain=True uses the FFT restatements."""
    rho = np.asarray(rho, np.float64)
    M, N = rho.shape
    f = (dct2_chain, idct2_chain, idxst_idct_chain, idct_idxst_chain) if chain else \
        (dct2, idct2, idxst_idct, idct_idxst)
    a = f[0](rho)
    wx, wy = frequencies(M, N)
    den = wy[:, None] ** 2 + wx[None, :] ** 2
    den[0, 0] = 1.0
    a = a / den
    a[0, 0] = 0.0
    psi = f[1](a)
    ex = f[2](a * wx[None, :])
    ey = f[3](a * wy[:, None])
    return psi, ex, ey


# ---------------------------------------------------------------------------
# D6: gather (FP64)
# ---------------------------------------------------------------------------
def gather(geom, cx, cy, dw, dh, scale, ex, ey, lam=None):
    """grad = sum over overlapped bins of lambda * overlap * scale * E * 0.5,
    precond = sum of lambda * overlap * scale (DensityGradient.cu:173-281)."""
    nbx, nby = geom["num_bin_x"], geom["num_bin_y"]
    cx, cy, dw, dh, scale = (np.asarray(v, np.float64) for v in (cx, cy, dw, dh, scale))
    lam = np.ones((nby, nbx)) if lam is None else np.asarray(lam, np.float64)
    exf, eyf, lf = ex.reshape(-1), ey.reshape(-1), lam.reshape(-1)
    blx, bux = geom["bin_lx"].astype(np.float64), geom["bin_ux"].astype(np.float64)
    bly, buy = geom["bin_ly"].astype(np.float64), geom["bin_uy"].astype(np.float64)
    x0, x1 = _bin_range(cx, dw, geom["die"][0], geom["bin_w"], nbx)
    y0, y1 = _bin_range(cy, dh, geom["die"][1], geom["bin_h"], nby)
    n = len(cx)
    gx, gy, pc = np.zeros(n), np.zeros(n), np.zeros(n)
    lx, ux = cx - dw / 2, cx + dw / 2
    ly, uy = cy - dh / 2, cy + dh / 2
    for i in range(int(max(int((x1 - x0).max(initial=0)), 0))):
        bx = x0 + i
        okx = bx < x1
        bxc = np.minimum(bx, nbx - 1)
        ox = np.maximum(0.0, np.minimum(ux, bux[bxc]) - np.maximum(lx, blx[bxc]))
        for j in range(int(max(int((y1 - y0).max(initial=0)), 0))):
            by = y0 + j
            ok = okx & (by < y1)
            byc = np.minimum(by, nby - 1)
            oy = np.maximum(0.0, np.minimum(uy, buy[byc]) - np.maximum(ly, bly[byc]))
            b = bxc + byc * nbx
            a = np.where(ok, ox * oy * scale * lf[b], 0.0)
            gx += a * exf[b] * 0.5
            gy += a * eyf[b] * 0.5
            pc += a
    return gx, gy, pc
