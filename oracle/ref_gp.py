"""ORACLE (test infrastructure only): restatement of SkyPlace's global
placement as the legalizer's front end -- design import, fillers, random
initial placement, the target function and the Nesterov / Adam loops -- on
top of the FP64 wirelength (ref_wl) and density (ref_density) oracles.

Imported only by tests/ and bench.py's CPU-baseline leg, as the checker.  The
product never imports this module.

Contents (reference: src/skyline/skyplace/src/)
  import_design    SkyPlaceDB::importDB (SkyPlaceDB.cpp:140-297): movable = not
                   fixed, centres = lower-left + size/2 in FP32, fixed area of
                   the instances not outside the die, sumScaledMovableArea
                   (macros x target density); validateTargetDensity (298-317)
  create_fillers   createFillers (SkyPlaceDB.cpp:551-662): mean width and
                   height over the [10 %, 90 %) slice of the width-sorted list
                   (the height list stays unsorted), count = (white space x
                   target density - scaled movable area) / filler area, centres
                   from std::mt19937(0) (numpy's RandomState(0) draws the same
                   32-bit stream)
  initial_place    doRandomInit (initial_place/InitialPlacer.cpp:77-123): normal
                   centres about the die centre, sigma = 0.05 x die side,
                   fillers mirrored (getMirrorX / Y, 735-751), moved inside the
                   die.  The reference draws from one std::default_random_engine
                   shared by 4 OpenMP threads (not reproducible); the product
                   draws sample k of cell i from splitmix64(seed + 2 i + k) by
                   Box-Muller, restated here.
  Objective        TargetFunction::updatePointAndGetGrad (TargetFunction.cu:88-141):
                   WA gradient + lambda x density gradient, Jacobi-preconditioned
                   (ProblemInstanceKernel.hpp:74-100); getInitialGrad (177-239);
                   updateGammaInv (241-256); updateLambda (258-274);
                   updateParameters (276-285); checkConvergence (287-299)
  Nesterov         NesterovSolver::initSolver / solve / backTracking / moveForward /
                   moveBackward / computeLipschitz (nonlinear_solver/NesterovSolver.cu:
                   16-264), clipToChipBoundary (ProblemInstanceKernel.hpp:7-54)
  Adam             AdamSolver (nonlinear_solver/AdamSolver.cu:12-226)
  export_positions SkyPlaceDB::exportDB (SkyPlaceDB.cpp:120-138)

Pins follow GPPin::updatePinLocation (object/GPPin.cpp:61-68): a pin is its
cell's centre plus the master pin's offset (the LgDesign pin_dx / pin_dy).
The oracle's objective is FP64 (FP64 DCT); the loop's scalar rules are the
reference's, with float where the reference keeps a float.
"""
from __future__ import annotations

import math

import numpy as np

from . import ref_density as R
from . import ref_wl as W

F = np.float32
MASK64 = np.uint64(0xFFFFFFFFFFFFFFFF)


# ---------------------------------------------------------------------------
# import, fillers, initial placement
# ---------------------------------------------------------------------------
def import_design(d, target_density: float = 1.0, scaling_factor: float = 1.0):
    """Movable / fixed cells and pins of a Design (SkyPlaceDB::importDB),
    coordinates divided by the scaling factor (GPCell.cpp:45-70)."""
    sf = F(scaling_factor)
    die = tuple(F(F(v) / sf) for v in d.core)
    dx0 = np.asarray(d.inst_dx, np.int64).astype(F)
    dy0 = np.asarray(d.inst_dy, np.int64).astype(F)
    cx = ((np.asarray(d.inst_lx, np.int64).astype(F) + dx0 / F(2)).astype(F) / sf).astype(F)
    cy = ((np.asarray(d.inst_ly, np.int64).astype(F) + dy0 / F(2)).astype(F) / sf).astype(F)
    dx = (dx0 / sf).astype(F)
    dy = (dy0 / sf).astype(F)
    fixed = np.asarray(d.inst_fixed) != 0
    macro = np.asarray(d.inst_macro) != 0
    mov = np.flatnonzero(~fixed)
    fx = np.flatnonzero(fixed)
    rects = np.stack([cx[fx] - dx[fx] / F(2), cy[fx] - dy[fx] / F(2), cx[fx] + dx[fx] / F(2),
                      cy[fx] + dy[fx] / F(2)], 1).astype(F) if len(fx) else np.zeros((0, 4), F)
    sum_fixed = F(0)
    for (lx, ly, ux, uy), a in zip(rects, (dx[fx] * dy[fx]).astype(F)):
        if not (ux <= die[0] or lx >= die[2] or uy <= die[1] or ly >= die[3]):
            sum_fixed = F(sum_fixed + a)
    sum_mov = F(0)
    sum_scaled = F(0)
    td = F(target_density)
    for i in mov:
        a = F(dx[i] * dy[i])
        sum_mov = F(sum_mov + a)
        sum_scaled = F(sum_scaled + (F(a * td) if macro[i] else a))
    mov_of = np.full(d.num_insts, -1, np.int64)
    mov_of[mov] = np.arange(len(mov))
    ns = np.asarray(d.net_start, np.int64)
    pin_inst = np.asarray(d.pin_inst, np.int64)
    off_x = (np.asarray(d.pin_dx, np.int64).astype(F) / sf).astype(F)
    off_y = (np.asarray(d.pin_dy, np.int64).astype(F) / sf).astype(F)
    io = pin_inst < 0
    pi = np.where(io, 0, pin_inst)
    pin2cell = np.where(io, -1, mov_of[pi])
    fixed_pin = (~io) & (pin2cell < 0)
    pin_x = np.where(io, off_x, (cx[pi] + off_x).astype(F)).astype(F)
    pin_y = np.where(io, off_y, (cy[pi] + off_y).astype(F)).astype(F)
    off_x = np.where(io, F(0), off_x).astype(F)
    off_y = np.where(io, F(0), off_y).astype(F)
    num_pins = np.bincount(pin2cell[pin2cell >= 0], minlength=len(mov)).astype(np.float64)
    die_area = F((die[2] - die[0]) * (die[3] - die[1]))
    util = F(sum_mov / F(die_area - sum_fixed))
    if float(td) * 100.0 <= float(util) * 100.0:
        raise ValueError("target density is not above the utilisation")
    return dict(die=die, mov=mov, cx=cx[mov], cy=cy[mov], w=dx[mov], h=dy[mov], macro=macro[mov],
                fixed_rects=rects, sum_fixed=sum_fixed, sum_mov=sum_mov, sum_scaled=sum_scaled,
                net_start=ns, pin2cell=pin2cell, off_x=off_x, off_y=off_y, pin_x=pin_x, pin_y=pin_y,
                fixed_pin=fixed_pin, num_pins=num_pins, td=float(td), sf=float(sf))


def create_fillers(w, h, die, sum_fixed, sum_scaled, target_density):
    """(filler width, filler height, count, centres x, centres y)."""
    n = len(w)
    dxs = np.sort(np.asarray(w, F))
    dys = np.asarray(h, F)
    lo = int(float(F(n)) * 0.10)
    hi = int(float(F(n)) * 0.90)
    if lo == hi:
        lo, hi = 0, n
    sx, sy = F(0), F(0)
    for i in range(lo, hi):
        sx = F(sx + dxs[i])
        sy = F(sy + dys[i])
    fw = F(sx / F(hi - lo))
    fh = F(sy / F(hi - lo))
    avg = F(fw * fh)
    die_area = F((F(die[2]) - F(die[0])) * (F(die[3]) - F(die[1])))
    white = F(die_area - F(sum_fixed))
    area = F(F(white * F(target_density)) - F(sum_scaled))
    if area < 0:
        raise ValueError("filler area is negative")
    count = int(area / avg)
    raw = np.random.RandomState(0).randint(0, 2 ** 32, size=2 * count, dtype=np.uint64)
    dlx, dly = int(die[0]), int(die[1])
    dw, dh = int(F(die[2]) - F(die[0])), int(F(die[3]) - F(die[1]))
    fcx = (raw[0::2] % np.uint64(dw) + np.uint64(dlx)).astype(F)
    fcy = (raw[1::2] % np.uint64(dh) + np.uint64(dly)).astype(F)
    return fw, fh, count, fcx, fcy


def splitmix64(x):
    x = np.asarray(x, np.uint64)
    with np.errstate(over="ignore"):
        z = (x + np.uint64(0x9E3779B97F4A7C15)) & MASK64
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & MASK64
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & MASK64
    return z ^ (z >> np.uint64(31))


def inside(c, size, lo, hi):
    """getXInsideChip (ProblemInstanceKernel.hpp:7-19) in FP32."""
    c = np.asarray(c, F)
    half = (np.asarray(size, F) / F(2)).astype(F)
    lo, hi = np.asarray(lo, F), np.asarray(hi, F)
    r = c.copy()
    r = np.where((c - half).astype(F) < lo, (lo + half).astype(F), r)
    r = np.where((c + half).astype(F) > hi, (hi - half).astype(F), r)
    return r.astype(F)


def initial_place(n, n_mov, seed, die, w, h):
    """The product's doRandomInit draws (see the module header)."""
    i = np.arange(n, dtype=np.uint64)
    s = np.uint64(seed)
    with np.errstate(over="ignore"):
        a = splitmix64(s + np.uint64(2) * i)
        b = splitmix64(s + np.uint64(2) * i + np.uint64(1))
    u1 = ((a >> np.uint64(11)).astype(np.float64) + 1.0) * 2.0 ** -53
    u2 = (b >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    r = np.sqrt(-2.0 * np.log(u1))
    t = 6.283185307179586 * u2
    lx, ly, ux, uy = (float(v) for v in die)
    mx, my = 0.5 * (lx + ux), 0.5 * (ly + uy)
    x = mx + (ux - lx) * 0.05 * (r * np.cos(t))
    y = my + (uy - ly) * 0.05 * (r * np.sin(t))
    fl = np.arange(n) >= n_mov
    x = np.where(fl, np.where(x <= mx, lx + (mx - x), ux - (x - mx)), x)
    y = np.where(fl, np.where(y <= my, ly + (my - y), uy - (y - my)), y)
    return inside(x.astype(F), w, lx, ux), inside(y.astype(F), h, ly, uy)


# ---------------------------------------------------------------------------
# target function
# ---------------------------------------------------------------------------
def gamma_inv(ovf, coef):
    """updateGammaInv (TargetFunction.cu:241-256)."""
    ovf = float(F(ovf))
    if ovf > 1.0:
        g = F(0.1)
    elif ovf < 0.1:
        g = F(10.0)
    else:
        g = F(1.0 / math.pow(10.0, 20 * (ovf - 0.1) / 9.0 - 1.0))
    return F(g * F(coef))


def lambda_coef(prev_hpwl, cur_hpwl, ref_hpwl, min_phi, max_phi):
    """updateLambda's multiplier (TargetFunction.cu:258-274)."""
    p = F(F(F(cur_hpwl) - F(prev_hpwl)) / F(ref_hpwl))
    if p < 0:
        return F(max_phi)
    return max(F(min_phi), F(F(max_phi) * F(np.power(F(max_phi), F(-p)))))


class Placement:
    """The placer's data after SkyPlaceDB::init: cells (movable then fillers),
    bins, density sizes, fixed area, netlist."""

    def __init__(self, d, target_density: float = 1.0, num_bins=None, scaling_factor: float = 1.0):
        imp = import_design(d, target_density, scaling_factor)
        self.imp = imp
        self.die = imp["die"]
        self.n_mov = len(imp["mov"])
        fw, fh, nf, fcx, fcy = create_fillers(imp["w"], imp["h"], self.die, imp["sum_fixed"], imp["sum_scaled"],
                                              target_density)
        self.filler_w, self.filler_h, self.n_fill = fw, fh, nf
        self.n = self.n_mov + nf
        self.w = np.concatenate([imp["w"], np.full(nf, fw, F)]).astype(F)
        self.h = np.concatenate([imp["h"], np.full(nf, fh, F)]).astype(F)
        self.cx0 = np.concatenate([imp["cx"], fcx]).astype(F)
        self.cy0 = np.concatenate([imp["cy"], fcy]).astype(F)
        self.is_filler = np.arange(self.n) >= self.n_mov
        self.is_macro = np.concatenate([imp["macro"], np.zeros(nf, bool)])
        self.td = float(target_density)
        if num_bins:
            self.geom = R.bins_from_counts(self.die, int(num_bins[0]), int(num_bins[1]))
        else:
            avg = F(imp["sum_mov"] / F(self.n_mov))
            self.geom = R.create_bins(self.die, float(avg), self.td)
        g = self.geom
        self.dw, self.dh, self.scale = R.density_sizes(self.w, self.h, g["bin_w"], g["bin_h"])
        self.fixed = R.fixed_area(g, imp["fixed_rects"], self.td)
        nbx, nby = g["num_bin_x"], g["num_bin_y"]
        bw = (g["bin_ux"] - g["bin_lx"]).astype(F)
        bh = (g["bin_uy"] - g["bin_ly"]).astype(F)
        self.scaled = (np.outer(bh, bw).astype(F) * F(self.td)).astype(np.float64).reshape(nby, nbx)
        self.sum_movable_area = float(imp["sum_mov"])
        self.num_pins = np.concatenate([imp["num_pins"], np.zeros(nf)])
        self.pin_x = imp["pin_x"].astype(np.float64)
        self.pin_y = imp["pin_y"].astype(np.float64)


class Objective:
    """updatePointAndGetGrad's two terms and the scalars the loop reads (FP64)."""

    def __init__(self, pl: Placement, min_precond: float = 1.0):
        self.pl = pl
        self.min_precond = min_precond
        self.evals = 0

    def wirelength(self, var, g_inv):
        """(WA gradient [2n], HPWL) with the pins at var."""
        pl, imp = self.pl, self.pl.imp
        n = pl.n
        px, py = W.update_pins(pl.pin_x, pl.pin_y, imp["pin2cell"], imp["off_x"], imp["off_y"],
                               np.asarray(var[:n], np.float64), np.asarray(var[n:], np.float64))
        gx, gy = W.wa_grad(imp["net_start"], px, py, imp["pin2cell"], float(g_inv),
                           np.ones(len(imp["net_start"]) - 1), n)
        return np.concatenate([gx, gy]), W.hpwl(imp["net_start"], px, py)

    def density(self, var):
        """(density gradient [2n], preconditioner [n], overflow) at var."""
        pl = self.pl
        n = pl.n
        cx, cy = np.asarray(var[:n], np.float64), np.asarray(var[n:], np.float64)
        mov, fil = R.scatter(pl.geom, cx, cy, pl.dw, pl.dh, pl.scale, pl.is_filler, pl.is_macro, pl.td)
        dens, ovf = R.overflow(mov, fil, pl.fixed, pl.scaled, pl.sum_movable_area)
        _, ex, ey = R.poisson(dens)
        gx, gy, pc = R.gather(pl.geom, cx, cy, pl.dw, pl.dh, pl.scale, ex, ey)
        return np.concatenate([gx, gy]), pc, float(ovf)

    def precondition(self, lam, wl, dens, pc):
        n = self.pl.n
        gx, gy = W.precondition(float(lam), self.min_precond, self.pl.num_pins, pc, wl[:n], wl[n:], dens[:n],
                                dens[n:])
        return np.concatenate([gx, gy]).astype(F)

    def evaluate(self, var, lam, g_inv):
        """(preconditioned gradient, HPWL, overflow)."""
        wl, hp = self.wirelength(var, g_inv)
        dens, pc, ovf = self.density(var)
        self.evals += 1
        return self.precondition(lam, wl, dens, pc), hp, ovf


def lipschitz(a_var, a_grad, b_var, b_grad):
    """computeLipschitz (NesterovSolver.cu:196-212)."""
    m = len(a_var)
    dx = np.asarray(a_var, np.float64) - np.asarray(b_var, np.float64)
    dg = np.asarray(a_grad, np.float64) - np.asarray(b_grad, np.float64)
    return float(F(math.sqrt(float(dx @ dx) / m) / math.sqrt(float(dg @ dg) / m)))


class _Loop:
    """TargetFunction's state and schedules shared by both solvers."""

    def __init__(self, pl: Placement, params: dict, var0):
        self.pl, self.p = pl, params
        self.obj = Objective(pl, params["min_precond"])
        n = pl.n
        self.size = np.concatenate([pl.w, pl.h]).astype(F)
        self.lo = np.concatenate([np.full(n, pl.die[0], F), np.full(n, pl.die[1], F)])
        self.hi = np.concatenate([np.full(n, pl.die[2], F), np.full(n, pl.die[3], F)])
        self.cur_var = np.asarray(var0, F).copy()
        self.lam = F(params["init_lambda"])
        self.g_inv = F(0)
        self.hpwl = F(0)
        self.prev_hpwl = F(0)
        self.ovf = F(0)
        self.iter = 0
        self.trace = []
        self.last = (0.0, 0.0)

    def clip(self, v):
        return inside(v, self.size, self.lo, self.hi)

    def grad_at(self, var):
        g, hp, ovf = self.obj.evaluate(var, self.lam, self.g_inv)
        self.last = (hp, ovf)
        return g

    def initial_grad(self, var):
        """getInitialGrad: density first (its overflow sets gamma), WA, the
        gradient with lambda = 0, lambda scaled by |wl|_1 / |density|_1."""
        dens, pc, ovf = self.obj.density(var)
        self.ovf = F(ovf)
        self.g_inv = gamma_inv(self.ovf, self.p["init_gamma_inv_coef"])
        wl, hp = self.obj.wirelength(var, self.g_inv)
        self.obj.evals += 1
        self.last = (hp, ovf)
        g = self.obj.precondition(0.0, wl, dens, pc)
        wl_sum, dens_sum = F(np.abs(wl).sum()), F(np.abs(dens).sum())
        self.lam = F(self.lam * F(wl_sum / dens_sum))
        return g

    def update_parameters(self):
        hp, ovf = self.last
        self.prev_hpwl = self.hpwl
        self.hpwl = F(hp)
        self.ovf = F(ovf)
        self.g_inv = gamma_inv(self.ovf, self.p["init_gamma_inv_coef"])
        self.lam = F(self.lam * lambda_coef(self.prev_hpwl, self.hpwl, self.p["reference_hpwl"],
                                            self.p["min_phi_coef"], self.p["max_phi_coef"]))

    def converged(self):
        return self.ovf <= F(self.p["target_overflow"]) and not (self.hpwl < self.prev_hpwl)

    def end_iteration(self, step, bt):
        """checkConvergence, then iterEndCbk; True when the loop ends."""
        if self.converged():
            self.done_converged = True
            return True
        self.update_parameters()
        self.trace.append((float(self.hpwl), float(self.ovf), float(self.lam), float(step), bt))
        self.iter += 1
        return self.iter >= self.p["max_opt_iter"]

    def solve(self):
        self.done_converged = False
        self.init_solver()
        while self.iter < self.p["max_opt_iter"] and not self.iterate():
            pass
        return self


class Nesterov(_Loop):
    """NesterovSolver; state names as the product's gp_get_state."""

    def __init__(self, pl: Placement, params: dict, var0):
        super().__init__(pl, params, var0)
        self.cur_pre = self.cur_var.copy()

    def init_solver(self):
        self.cur_grad = self.initial_grad(self.cur_pre)
        self.prv_pre = self.clip((self.cur_pre - F(self.p["init_opt_coef"]) * self.cur_grad).astype(F))
        self.prv_grad = self.grad_at(self.prv_pre)
        self.step = F(lipschitz(self.prv_pre, self.prv_grad, self.cur_pre, self.cur_grad))
        self.update_parameters()
        self.cur_a = F(1.0)

    def move_forward(self, step, coef):
        nx = (self.cur_pre + F(step) * self.cur_grad).astype(F)
        npre = (nx + F(coef) * (nx - self.cur_var).astype(F)).astype(F)
        return self.clip(nx), self.clip(npre)

    def iterate(self):
        """One loop body (backTracking, updateOneIteration, checkConvergence,
        iterEndCbk); returns True when the loop ends."""
        p = self.p
        prev_a = self.cur_a
        self.cur_a = F(0.5 * (1.0 + math.sqrt(4.0 * float(prev_a) * float(prev_a) + 1.0)))
        coef = F((float(prev_a) - 1.0) / float(self.cur_a))
        ns = F(p["min_step_length"])
        bt = 0
        while bt < p["max_backtrack_iter"]:
            nvar, npre = self.move_forward(self.step, coef)
            ngrad = self.grad_at(npre)
            ns = F(lipschitz(self.cur_pre, self.cur_grad, npre, ngrad))
            if float(ns) >= float(self.step) * 0.95:
                break
            if ns < F(p["min_step_length"]):
                ns = F(p["min_step_length"])
                break
            bt += 1
        self.step = ns
        self.prv_pre, self.prv_grad = self.cur_pre, self.cur_grad
        self.cur_pre, self.cur_grad, self.cur_var = npre, ngrad, nvar
        return self.end_iteration(self.step, min(bt + 1, p["max_backtrack_iter"]))


class Adam(_Loop):
    """AdamSolver (AdamSolver.cu:50-213); epsilon as the product's parameter."""

    def init_solver(self):
        p = self.p
        self.alpha = F(p["adam_alpha"])
        self.b1k, self.b2k = F(p["adam_beta1"]), F(p["adam_beta2"])
        self.cur_grad = self.initial_grad(self.cur_var)
        self.update_parameters()
        self.m = np.zeros_like(self.cur_var)
        self.v = np.zeros_like(self.cur_var)

    def iterate(self):
        p = self.p
        g = self.grad_at(self.cur_var).astype(F)
        b1, b2 = F(p["adam_beta1"]), F(p["adam_beta2"])
        m1 = (b1 * self.m + F(-(F(1) - b1)) * g).astype(F)
        v1 = (b2 * self.v + ((F(1) - b2) * g).astype(F) * g).astype(F)
        c1, c2 = F(1.0 / (1.0 - float(self.b1k))), F(1.0 / (1.0 - float(self.b2k)))
        d = ((c1 * m1).astype(F) / (np.sqrt(np.maximum(F(0), (c2 * v1).astype(F))) + F(p["adam_epsilon"]))).astype(F)
        self.cur_var = self.clip((self.cur_var - self.alpha * d).astype(F))
        self.m, self.v = m1, v1
        step = self.alpha
        self.alpha = F(self.alpha * F(0.997))
        self.b1k = F(self.b1k * b1)
        self.b2k = F(self.b2k * b2)
        return self.end_iteration(step, 1)


def params_dict(gp) -> dict:
    """The fields of a product GpParams (ctypes) as a dict."""
    return {k: getattr(gp, k) for k, _ in gp._fields_}


def export_positions(d, pl: Placement, cur_var):
    """exportDB: lower-left corners of the movable instances, rounded."""
    n = pl.n
    lx = np.asarray(d.inst_lx, np.int32).copy()
    ly = np.asarray(d.inst_ly, np.int32).copy()
    m = pl.n_mov
    llx = (np.asarray(cur_var[:m], F) - (pl.w[:m] / F(2)).astype(F)).astype(F)
    lly = (np.asarray(cur_var[n:n + m], F) - (pl.h[:m] / F(2)).astype(F)).astype(F)
    # std::round of lx * scaling factor (FP32): halves away from zero
    sf = F(pl.imp["sf"])
    rnd = lambda v: (np.sign(v) * np.floor(np.abs((v * sf).astype(F).astype(np.float64)) + 0.5)).astype(np.int32)
    lx[pl.imp["mov"]] = rnd(llx)
    ly[pl.imp["mov"]] = rnd(lly)
    return lx, ly
