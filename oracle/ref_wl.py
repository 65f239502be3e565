"""ORACLE (test infrastructure only): FP64 restatement of the reference
placer's weighted-average (WA) wirelength gradient and of the target
function's preconditioned total gradient.

Imported only by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
leg, as the checker.  The product never imports this module.

Contents (reference: src/skyline/skyplace/src/)
  update_pins      updatePinCoordinateKernel (wirelength_gradient/WireLengthGradient.cu:187-211):
                   pin = centre of its movable cell + offset; pins of no movable cell keep
                   their coordinates
  net_min_max      computeMinMax (WireLengthGradient.cu:14-45): max starts at 0, min at FLT_MAX
  wa_grad          computeABCKernel + computePinGrad + addPinGrad
                   (WireLengthGradient.cu:47-185): per pin a+- = exp(+-(x - max/min) * gInv),
                   per net b+- = sum a+-, c+- = sum x a+-, per pin
                   g = ((1 - x gInv) b- + gInv c-) a- / b-^2 - ((1 + x gInv) b+ - gInv c+) a+ / b+^2,
                   times the net weight, summed per movable cell
  wa_wirelength    computeWAForEachNet (WireLengthGradient.cu:237-259): W = c+/b+ - c-/b- per
                   net and axis (the function whose derivative pins wa_grad: g = -dW/dx)
  hpwl             computeNetBBox + computeHPWL (WireLengthGradient.cu:213-235, 383-401)
  precondition     jacobianPrecondition (problem_instance/ProblemInstanceKernel.hpp:74-100)

Netlist conventions (WireLengthGradient::initForCUDAKernel,
WireLengthGradient.cu:404-568): pins are numbered net by net, net_start[n] is
the first pin of net n (net_start[num_nets] = num_pins); pin2cell[p] is the
movable cell of pin p or -1.
"""
from __future__ import annotations

import numpy as np

FLT_MAX = float(np.finfo(np.float32).max)


def _net_ids(net_start):
    deg = np.diff(np.asarray(net_start, np.int64))
    return np.repeat(np.arange(deg.size), deg)


def update_pins(pin_x, pin_y, pin2cell, off_x, off_y, cell_cx, cell_cy):
    px, py = np.array(pin_x, np.float64), np.array(pin_y, np.float64)
    m = np.asarray(pin2cell) >= 0
    c = np.asarray(pin2cell)[m]
    px[m] = np.asarray(cell_cx, np.float64)[c] + np.asarray(off_x, np.float64)[m]
    py[m] = np.asarray(cell_cy, np.float64)[c] + np.asarray(off_y, np.float64)[m]
    return px, py


def net_min_max(net_start, v):
    """(max, min) per net with the reference's initial values 0 and FLT_MAX."""
    ns = np.asarray(net_start, np.int64)
    n = ns.size - 1
    nid = _net_ids(ns)
    mx = np.zeros(n)
    mn = np.full(n, FLT_MAX)
    np.maximum.at(mx, nid, v)
    np.minimum.at(mn, nid, v)
    return mx, mn


def _abc(net_start, v, g_inv):
    nid = _net_ids(net_start)
    mx, mn = net_min_max(net_start, v)
    ap = np.exp((v - mx[nid]) * g_inv)
    am = np.exp(-(v - mn[nid]) * g_inv)
    n = len(net_start) - 1
    bp, bm = np.bincount(nid, ap, n), np.bincount(nid, am, n)
    cp, cm = np.bincount(nid, v * ap, n), np.bincount(nid, v * am, n)
    return nid, ap, am, bp, bm, cp, cm


def wa_grad_axis(net_start, v, g_inv, net_weight):
    """Per-pin gradient term of one axis (the reference's pinGrad)."""
    v = np.asarray(v, np.float64)
    nid, ap, am, bp, bm, cp, cm = _abc(net_start, v, g_inv)
    g = ((1.0 - v * g_inv) * bm[nid] + g_inv * cm[nid]) * am / bm[nid] / bm[nid] \
        - ((1.0 + v * g_inv) * bp[nid] - g_inv * cp[nid]) * ap / bp[nid] / bp[nid]
    return g * np.asarray(net_weight, np.float64)[nid]


def wa_grad(net_start, pin_x, pin_y, pin2cell, g_inv, net_weight, num_cells):
    """Cell gradients (x, y) of the WA wirelength as the reference computes them."""
    gx = wa_grad_axis(net_start, pin_x, g_inv, net_weight)
    gy = wa_grad_axis(net_start, pin_y, g_inv, net_weight)
    p2c = np.asarray(pin2cell)
    m = p2c >= 0
    return (np.bincount(p2c[m], gx[m], num_cells), np.bincount(p2c[m], gy[m], num_cells))


def wa_wirelength(net_start, v, g_inv):
    """WA wirelength of every net along one axis."""
    _, _, _, bp, bm, cp, cm = _abc(net_start, np.asarray(v, np.float64), g_inv)
    return cp / bp - cm / bm


def hpwl(net_start, pin_x, pin_y):
    mx, mn = net_min_max(net_start, np.asarray(pin_x, np.float64))
    my, ny = net_min_max(net_start, np.asarray(pin_y, np.float64))
    return float((mx - mn).sum() + (my - ny).sum())


def precondition(lam, min_precond, num_pins, dens_precond, wl_gx, wl_gy, dens_gx, dens_gy):
    """jacobianPrecondition: (wl + lambda dens) / max(#pins + lambda precond, minPrecond).
    The reference converts the per-cell pin count to int first."""
    p = np.floor(np.asarray(num_pins, np.float64)) + lam * np.asarray(dens_precond, np.float64)
    p = np.maximum(p, min_precond)
    return ((np.asarray(wl_gx, np.float64) + lam * np.asarray(dens_gx, np.float64)) / p,
            (np.asarray(wl_gy, np.float64) + lam * np.asarray(dens_gy, np.float64)) / p)
