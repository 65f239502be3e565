"""Writes tests/golden/poisson_*.npz: bin-density maps and their potential and
field computed with scipy.fft's DCT / DST routines in FP64 (a library
independent of the oracle's own transforms).

    python tests/golden/make_golden.py

Transforms (M rows = y, N columns = x):
  a   = dctn(rho, type=2) / (M N)                    (reference forward scaling)
  a  /= wy^2 + wx^2, a[0,0] = 0,  w = 2 pi k / bins
  psi = dctn(a, type=3)
  ex  = DST-III along x (input shifted by one) of the DCT-III along y of (wx a)
  ey  = DST-III along y (input shifted by one) of the DCT-III along x of (wy a)
"""
import os

import numpy as np
import scipy.fft as sf

HERE = os.path.dirname(os.path.abspath(__file__))


def poisson_scipy(rho):
    M, N = rho.shape
    a = sf.dctn(rho, type=2) / (M * N)
    wx = 2 * np.pi * np.arange(N) / N
    wy = 2 * np.pi * np.arange(M) / M
    den = wy[:, None] ** 2 + wx[None, :] ** 2
    den[0, 0] = 1.0
    a = a / den
    a[0, 0] = 0.0
    psi = sf.dctn(a, type=3)
    col = sf.dct(a * wx[None, :], type=3, axis=0)
    cs = np.zeros_like(col)
    cs[:, :-1] = col[:, 1:]
    ex = sf.dst(cs, type=3, axis=1)
    row = sf.dct(a * wy[:, None], type=3, axis=1)
    rs = np.zeros_like(row)
    rs[:-1, :] = row[1:, :]
    ey = sf.dst(rs, type=3, axis=0)
    return psi, ex, ey


def main():
    for n, seed in ((16, 1), (64, 2)):
        rng = np.random.default_rng(seed)
        rho = (rng.random((n, n)) * 1.5).astype(np.float32)
        psi, ex, ey = poisson_scipy(rho.astype(np.float64))
        np.savez(os.path.join(HERE, f"poisson_{n}x{n}.npz"), rho=rho, psi=psi, ex=ex, ey=ey)


if __name__ == "__main__":
    main()
