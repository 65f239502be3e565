"""Seeded synthetic placements for the legalization workloads.

There is no benchmark data on the box, so every workload is generated here:
site width ``sw`` and row height ``rh`` in dbu (default 100 and 800: a row is
8 sites tall, and the QP refine's 1-dbu termination tolerance is a hundredth
of a site as on real designs), rows alternating N / FS (so the
VDD rail alternates, reference Grid.cpp:337-341), fixed rectangular macros,
optional rectangular fence regions, and movable cells whose global-placement
positions are a mixture of a uniform spread and Gaussian hotspots.  Each cell
carries a ``ground_at_bottom`` bit (drawn 50/50) that decides which rows an
even-height cell may sit on.

Named workloads (sizes as stated in the bench/test configuration):

* ``plain_20k``   - 20k single-row cells, widths U{1..8} sites, 250 rows,
  75 % utilisation (row length derived from that: 490 sites), 10 macros
  20-60 sites x 3-10 rows, 4 Gaussian hotspots (sigma = 6 rows) at ~1.4x
  local density.  20k cells of mean width 4.5 on 250 x 2000 sites would be
  18 % utilisation, so the row is shortened to keep 75 %.
* ``mixed_1m``    - 1M cells, heights {1:86 %, 2:10 %, 3:3 %, 4:1 %},
  1600 rows x 12000 sites, 85 % utilisation (width range calibrated to it),
  3 fences each holding 5 % of the cells at the same utilisation, 200 macros.
* ``mixed(n)``    - the same recipe scaled to ``n`` cells (tests, requests).
"""
from __future__ import annotations

import math

import numpy as np

from .design import ORIENT_FS, ORIENT_N, Design


def _rows(x_size: int, y_size: int, sw: int, rh: int):
    ys = np.arange(y_size, dtype=np.int32)
    return dict(
        row_lx=np.zeros(y_size, np.int32),
        row_ly=(ys * rh).astype(np.int32),
        row_dx=np.full(y_size, x_size * sw, np.int32),
        row_dy=np.full(y_size, rh, np.int32),
        row_site_width=np.full(y_size, sw, np.int32),
        row_orient=np.where(ys % 2 == 0, ORIENT_N, ORIENT_FS).astype(np.int32),
    )


def _place_rects(rng, x_size, y_size, n, wr, hr, occupied, margin=2, tries=200000):
    """Non-overlapping grid-aligned rectangles (in sites/rows) avoiding `occupied`."""
    rects = []
    occ = occupied
    t = 0
    while len(rects) < n and t < tries:
        t += 1
        w = int(rng.integers(wr[0], wr[1] + 1))
        h = int(rng.integers(hr[0], hr[1] + 1))
        if w >= x_size or h >= y_size:
            continue
        x = int(rng.integers(0, x_size - w + 1))
        y = int(rng.integers(0, y_size - h + 1))
        x0, y0 = max(0, x - margin), max(0, y - margin)
        if occ[x0:x + w + margin, y0:y + h + margin].any():
            continue
        occ[x:x + w, y:y + h] = True
        rects.append((x, y, w, h))
    return rects


def make_design(num_cells: int, x_size: int, y_size: int, *, sw: int = 100, rh: int = 800,
                height_mix=None, width_range=(1, 8), num_macros: int = 0,
                macro_w=(20, 60), macro_h=(3, 10), fence_frac: float = 0.0,
                num_fences: int = 0, fence_util: float = 0.85, hotspots: int = 0,
                hotspot_sigma_rows: float = 6.0, hotspot_frac: float = 0.0,
                seed: int = 0, name: str = "custom") -> Design:
    rng = np.random.default_rng(seed)
    height_mix = height_mix or {1: 1.0}
    hs = np.array(sorted(height_mix), np.int32)
    hp = np.array([height_mix[h] for h in hs], np.float64)
    hp /= hp.sum()

    heights = rng.choice(hs, size=num_cells, p=hp).astype(np.int32)
    widths = rng.integers(width_range[0], width_range[1] + 1, size=num_cells).astype(np.int32)
    gnd = rng.integers(0, 2, size=num_cells).astype(np.uint8)
    group = np.full(num_cells, -1, np.int32)

    occupied = np.zeros((x_size, y_size), bool)

    # fences first (big rectangles), sized for their share of cells
    fences = []
    if num_fences > 0 and fence_frac > 0:
        per = int(round(num_cells * fence_frac))
        idx = rng.permutation(num_cells)
        for f in range(num_fences):
            members = idx[f * per:(f + 1) * per]
            area_sites = float((widths[members] * heights[members]).sum())
            # same sites:rows aspect as the mixed-height die (12000:1600)
            side = math.sqrt(area_sites / fence_util / 7.5)
            fw = max(8, int(math.ceil(side * 7.5)))
            fh = max(2, int(math.ceil(area_sites / fence_util / fw)))
            r = _place_rects(rng, x_size, y_size, 1, (fw, fw), (fh, fh), occupied, margin=4)
            if not r:
                raise RuntimeError("cannot place fence")
            fences.append(r[0])
            group[members] = f
    macros = _place_rects(rng, x_size, y_size, num_macros, macro_w, macro_h, occupied, margin=1)

    # -- positions: lower-left corners in dbu ------------------------------
    W, H = x_size * sw, y_size * rh
    blocked = np.zeros((x_size, y_size), bool)
    for (x, y, w, h) in fences + macros:
        blocked[x:x + w, y:y + h] = True

    def_idx = np.nonzero(group < 0)[0]
    n_def = def_idx.shape[0]
    centers = [(rng.uniform(0.15, 0.85) * W, rng.uniform(0.15, 0.85) * H) for _ in range(hotspots)]
    sig = hotspot_sigma_rows * rh
    is_hot = rng.random(n_def) < (hotspot_frac if hotspots else 0.0)
    # prefix sums of the blocked map: reject footprints that touch a fence or macro
    cs = np.zeros((x_size + 1, y_size + 1), np.int64)
    cs[1:, 1:] = np.cumsum(np.cumsum(blocked, 0), 1)
    wd, hd = widths[def_idx], heights[def_idx]
    px = np.empty(n_def, np.int64)
    py = np.empty(n_def, np.int64)
    todo = np.arange(n_def)
    while todo.size:
        hot = is_hot[todo]
        w, h = wd[todo], hd[todo]
        ux = rng.uniform(0, 1, size=todo.size) * (W - w * sw)
        uy = rng.uniform(0, 1, size=todo.size) * (H - h * rh)
        if hotspots:
            k = rng.integers(0, hotspots, size=todo.size)
            ccx = np.array([c[0] for c in centers])[k] - w * sw / 2.0
            ccy = np.array([c[1] for c in centers])[k] - h * rh / 2.0
            ux = np.where(hot, rng.normal(ccx, sig), ux)
            uy = np.where(hot, rng.normal(ccy, sig), uy)
        lxs = np.clip(np.rint(ux), 0, W - w * sw).astype(np.int64)
        lys = np.clip(np.rint(uy), 0, H - h * rh).astype(np.int64)
        x0, y0 = lxs // sw, lys // rh
        x1 = np.minimum(x_size, -(-(lxs + w * sw) // sw))
        y1 = np.minimum(y_size, -(-(lys + h * rh) // rh))
        hit = cs[x1, y1] - cs[x0, y1] - cs[x1, y0] + cs[x0, y0]
        ok = hit == 0
        px[todo[ok]] = lxs[ok]
        py[todo[ok]] = lys[ok]
        todo = todo[~ok]
    lx = np.zeros(num_cells, np.int64)
    ly = np.zeros(num_cells, np.int64)
    lx[def_idx] = px
    ly[def_idx] = py
    for f, (x, y, w, h) in enumerate(fences):
        m = np.nonzero(group == f)[0]
        lx[m] = np.rint(rng.uniform(0, 1, size=m.size) * ((x + w) * sw - widths[m] * sw - x * sw) + x * sw)
        ly[m] = np.rint(rng.uniform(0, 1, size=m.size) * ((y + h) * rh - heights[m] * rh - y * rh) + y * rh)
    lx = lx.astype(np.int32)
    ly = ly.astype(np.int32)

    # -- instances: movable cells first, then fixed macros ------------------
    nm = len(macros)
    inst_lx = np.concatenate([lx, np.array([m[0] * sw for m in macros], np.int32)]).astype(np.int32)
    inst_ly = np.concatenate([ly, np.array([m[1] * rh for m in macros], np.int32)]).astype(np.int32)
    inst_dx = np.concatenate([widths * sw, np.array([m[2] * sw for m in macros], np.int32)]).astype(np.int32)
    inst_dy = np.concatenate([heights * rh, np.array([m[3] * rh for m in macros], np.int32)]).astype(np.int32)
    inst_fixed = np.concatenate([np.zeros(num_cells, np.uint8), np.ones(nm, np.uint8)])
    inst_macro = np.concatenate([np.zeros(num_cells, np.uint8), np.ones(nm, np.uint8)])
    inst_group = np.concatenate([group, np.full(nm, -1, np.int32)]).astype(np.int32)
    inst_gnd = np.concatenate([gnd, np.ones(nm, np.uint8)])

    box = np.array([(f, x * sw, y * rh, (x + w) * sw, (y + h) * rh)
                    for f, (x, y, w, h) in enumerate(fences)], np.int32).reshape(-1, 5)
    free_sites = x_size * y_size - sum(m[2] * m[3] for m in macros)
    util = float((widths * heights).sum()) / max(free_sites, 1)
    return Design(
        core=(0, 0, W, H), **_rows(x_size, y_size, sw, rh),
        inst_lx=inst_lx, inst_ly=inst_ly, inst_dx=inst_dx, inst_dy=inst_dy,
        inst_fixed=inst_fixed, inst_macro=inst_macro, inst_group=inst_group,
        inst_ground_at_bottom=inst_gnd, num_groups=len(fences),
        box_group=box[:, 0].copy(), box_lx=box[:, 1].copy(), box_ly=box[:, 2].copy(),
        box_ux=box[:, 3].copy(), box_uy=box[:, 4].copy(),
        meta=dict(name=name, seed=seed, num_cells=num_cells, x_size=x_size, y_size=y_size,
                  utilization=util, num_macros=nm, num_fences=len(fences)),
    )


MIXED_HEIGHTS = {1: 0.86, 2: 0.10, 3: 0.03, 4: 0.01}


def plain_20k(seed: int = 0) -> Design:
    """20k single-row cells on 250 rows x 490 sites at 75 % utilisation."""
    return make_design(20000, 490, 250, width_range=(1, 8), num_macros=10,
                       macro_w=(20, 60), macro_h=(3, 10), hotspots=4, hotspot_sigma_rows=6.0,
                       hotspot_frac=0.057, seed=seed, name="plain_20k")


def mixed(num_cells: int, util: float = 0.85, seed: int = 0, num_fences: int = 3,
          fence_frac: float = 0.05, macros_per_mcell: float = 200.0,
          hotspots: int = 0, name: str = "mixed") -> Design:
    """The mixed-height recipe at `num_cells` cells: the die keeps the
    1600:12000 (rows:sites) aspect of the 1M-cell case and is sized so that
    movable area / free area = `util` with widths U{4..23}."""
    wr = (4, 23)
    mean_area = (sum(wr) / 2.0) * sum(h * p for h, p in MIXED_HEIGHTS.items())
    n_mac = int(round(macros_per_mcell * num_cells / 1e6))
    mac_area = n_mac * 40 * 6.5
    total = num_cells * mean_area / util + mac_area
    # sites:rows = 12000:1600 in site units -> x = 7.5 * y
    y = max(12, int(round(math.sqrt(total / 7.5))))
    x = max(60, int(math.ceil(total / y)))
    return make_design(num_cells, x, y, height_mix=MIXED_HEIGHTS, width_range=wr,
                       num_macros=n_mac, fence_frac=fence_frac if num_fences else 0.0,
                       num_fences=num_fences, fence_util=util, hotspots=hotspots,
                       hotspot_frac=0.05 if hotspots else 0.0, seed=seed, name=name)


def mixed_1m(seed: int = 0) -> Design:
    """1M mixed-height cells, 1600 rows x 12000 sites, 85 % utilisation."""
    wr = _calibrate_width(1_000_000, 12000 * 1600 - 200 * 260, 0.85)
    return make_design(1_000_000, 12000, 1600, height_mix=MIXED_HEIGHTS,
                       width_range=wr, num_macros=200, fence_frac=0.05, num_fences=3,
                       fence_util=0.85, seed=seed, name="mixed_1m")


def stream_spec(num: int = 2000, seed: int = 0, total: int = 2000):
    """(cells, utilisation, design seed) of the first `num` requests of the
    seeded `total`-request legalization stream: cell counts log-uniform on
    [500, 200k], utilisation uniform on [0.6, 0.9]."""
    rng = np.random.default_rng(seed)
    n = np.exp(rng.uniform(np.log(500), np.log(200_000), total)).astype(np.int64)
    u = rng.uniform(0.6, 0.9, total)
    return [(int(n[i]), float(u[i]), 1000 + i) for i in range(min(num, total))]


def stream_request(n: int, util: float, seed: int) -> Design:
    """One request of the stream: an ECO window (below 20k cells, no fence)
    or a sub-design (3 fences) of the mixed-height recipe."""
    return mixed(n, util=util, seed=seed, num_fences=3 if n >= 20000 else 0, name=f"req{seed}")


def with_wide_cells(d: Design, widths=(260, 270, 280, 290), seed: int = 1) -> Design:
    """A copy of `d` in which a few single-row movable cells are wider than a
    partition (x_hint = 250 sites): such cells belong to no partition, so the
    ADMM steps never move them."""
    rng = np.random.default_rng(seed)
    mv = np.flatnonzero(d.movable_mask() & (d.inst_dy == d.row_height) & (d.inst_group < 0))
    pick = rng.choice(mv, len(widths), replace=False)
    sw = d.site_width
    e = d.copy_positions(d.inst_lx, d.inst_ly)
    e.inst_dx = d.inst_dx.copy()
    e.inst_lx = d.inst_lx.copy()
    for i, w in zip(pick, widths):
        e.inst_dx[i] = w * sw
        e.inst_lx[i] = min(int(e.inst_lx[i]), int(d.core[2]) - w * sw)
    e.meta = dict(d.meta, wide_cells=[int(i) for i in pick], name=d.meta.get("name", "design") + "_wide")
    return e


def with_netlist(d: Design, nets_per_cell: float = 1.2, io_frac: float = 0.02, seed: int = 0,
                 names: bool = True) -> Design:
    """A copy of `d` with a random netlist: nets of 2-8 pins (geometric) on
    the movable cells and fixed macros, pins at integer offsets inside their
    instances, a few IO pins on the core boundary, and instance names
    "o<index>"."""
    rng = np.random.default_rng(seed)
    e = d.copy_positions(d.inst_lx, d.inst_ly)
    cand = np.flatnonzero(d.movable_mask() | (d.inst_macro == 1))
    nn = max(1, int(nets_per_cell * int(d.movable_mask().sum())))
    deg = np.minimum(2 + rng.geometric(0.45, nn) - 1, 8)
    start = np.zeros(nn + 1, np.int64)
    start[1:] = np.cumsum(deg)
    npins = int(start[-1])
    inst = rng.choice(cand, npins).astype(np.int32)
    io = rng.random(npins) < io_frac
    dx = (rng.random(npins) * d.inst_dx[inst]).astype(np.int32)
    dy = (rng.random(npins) * d.inst_dy[inst]).astype(np.int32)
    lx, ly, ux, uy = d.core
    side = rng.integers(0, 2, npins)
    dx = np.where(io, np.where(side == 0, lx, ux), dx).astype(np.int32)
    dy = np.where(io, rng.integers(ly, uy, npins), dy).astype(np.int32)
    e.net_start = start.astype(np.int32)
    e.pin_inst = np.where(io, -1, inst).astype(np.int32)
    e.pin_dx, e.pin_dy = dx, dy
    if names:
        e.inst_name = [f"o{i}" for i in range(d.num_insts)]
    return e


def _morton(x, y):
    """Interleaved bits of 16-bit x and y (Z-order key)."""
    def spread(v):
        v = v.astype(np.uint64) & np.uint64(0xFFFF)
        v = (v | (v << np.uint64(8))) & np.uint64(0x00FF00FF)
        v = (v | (v << np.uint64(4))) & np.uint64(0x0F0F0F0F)
        v = (v | (v << np.uint64(2))) & np.uint64(0x33333333)
        return (v | (v << np.uint64(1))) & np.uint64(0x55555555)
    return spread(x) | (spread(y) << np.uint64(1))


def with_power_law_netlist(d: Design, nets_per_cell: float = 1.05, min_deg: int = 2, max_deg: int = 1000,
                           mean_deg: float = 3.8, io_frac: float = 0.005, local: bool = True,
                           seed: int = 0) -> Design:
    """A copy of `d` with a random netlist whose net degrees follow a power
    law on [min_deg, max_deg] with mean `mean_deg`: pins on movable cells and
    fixed macros at integer offsets inside them, a fraction `io_frac` of IO
    pins on the core boundary.  With `local` a net's cells are drawn from a
    window of 8 x degree (at least 64) consecutive instances in Z order of
    the design's positions, so that nets are short in a good placement (as
    in real netlists); otherwise from all instances uniformly."""
    rng = np.random.default_rng(seed)
    e = d.copy_positions(d.inst_lx, d.inst_ly)
    cand = np.flatnonzero(d.movable_mask() | (d.inst_macro == 1))
    nn = max(1, int(round(nets_per_cell * int(d.movable_mask().sum()))))
    a = _power_law_alpha(min_deg, max_deg, mean_deg)
    dd = np.arange(min_deg, max_deg + 1)
    p = dd.astype(np.float64) ** -a
    deg = rng.choice(dd, size=nn, p=p / p.sum()).astype(np.int64)
    start = np.zeros(nn + 1, np.int64)
    start[1:] = np.cumsum(deg)
    npins = int(start[-1])
    if local:
        lx, ly, ux, uy = d.core
        cxs = (d.inst_lx[cand] + d.inst_dx[cand] / 2.0 - lx) / max(ux - lx, 1) * 65535
        cys = (d.inst_ly[cand] + d.inst_dy[cand] / 2.0 - ly) / max(uy - ly, 1) * 65535
        order = cand[np.argsort(_morton(np.clip(cxs, 0, 65535), np.clip(cys, 0, 65535)), kind="stable")]
        win = np.maximum(8 * deg, 64)
        anchor = (rng.random(nn) * np.maximum(cand.size - win, 1)).astype(np.int64)
        net_of_pin = np.repeat(np.arange(nn), deg)
        pos = anchor[net_of_pin] + (rng.random(npins) * win[net_of_pin]).astype(np.int64)
        inst = order[np.minimum(pos, cand.size - 1)].astype(np.int32)
    else:
        inst = cand[rng.integers(0, cand.size, npins)].astype(np.int32)
    io = rng.random(npins) < io_frac
    dx = (rng.random(npins) * d.inst_dx[inst]).astype(np.int32)
    dy = (rng.random(npins) * d.inst_dy[inst]).astype(np.int32)
    lx, ly, ux, uy = d.core
    side = rng.integers(0, 2, npins)
    dx = np.where(io, np.where(side == 0, lx, ux), dx).astype(np.int32)
    dy = np.where(io, rng.integers(ly, uy, npins), dy).astype(np.int32)
    e.net_start = start.astype(np.int32)
    e.pin_inst = np.where(io, -1, inst).astype(np.int32)
    e.pin_dx, e.pin_dy = dx, dy
    e.meta = dict(d.meta, nets=nn, pins=npins, degree_alpha=a)
    return e


def e2e(num_cells: int, util: float = 0.9, seed: int = 0, macros_per_mcell: float = 150.0,
        nets_per_cell: float = 1.05) -> Design:
    """The end-to-end placement recipe at `num_cells` cells: the mixed-height
    cells at `util` with fixed macros (150 per million cells: 300 at 2M) and
    no fences, and a power-law netlist (degrees 2-1000, mean 3.8, 1.05 nets
    per cell: 2.1M nets at 2M cells)."""
    d = mixed(num_cells, util=util, seed=seed, num_fences=0, macros_per_mcell=macros_per_mcell,
              name=f"e2e_{num_cells}")
    return with_power_law_netlist(d, nets_per_cell=nets_per_cell, seed=seed + 1)


def e2e_placer_params(d: Design, bins: int = 4096):
    """Global-placer settings of the end-to-end recipe: BF16 transforms on
    bins x bins, coordinates in sites (scaling factor = site width), the
    lambda schedule's reference HPWL at 1 % of the design's own HPWL, target
    density 0.95 (5 points above the utilisation) and an overflow target for
    the bin size: 0.37 at 4096 x 4096 bins, where a bin is a quarter of a
    cell (the loop then stops at the placement that 1024 x 1024 bins measure
    at overflow 0.17), else 0.2."""
    from . import density as D
    from . import gp
    p = gp.default_params()
    p.precision = D.PREC_BF16
    p.num_bin_x = p.num_bin_y = int(bins)
    p.scaling_factor = float(d.site_width)
    p.reference_hpwl_ratio = 0.01
    p.target_density = 0.95
    p.target_overflow = 0.37 if bins >= 4096 else 0.2
    return p


def _calibrate_width(n, free_sites, util):
    mean_h = sum(h * p for h, p in MIXED_HEIGHTS.items())
    target_w = util * free_sites / n / mean_h
    lo = 4
    hi = max(lo, int(round(2 * target_w - lo)))
    return (lo, hi)


def tiny(num_cells: int = 40, x_size: int = 48, y_size: int = 12, seed: int = 0,
         heights=None, num_macros: int = 1, fences: int = 0) -> Design:
    """Small dense case for exhaustive checks."""
    return make_design(num_cells, x_size, y_size,
                       height_mix=heights or {1: 0.8, 2: 0.2}, width_range=(1, 4),
                       num_macros=num_macros, macro_w=(3, 5), macro_h=(2, 3),
                       num_fences=fences, fence_frac=0.1 if fences else 0.0,
                       hotspots=1, hotspot_frac=0.3, hotspot_sigma_rows=2.0, seed=seed,
                       name="tiny")


def with_tech(grid: dict, cells: dict, seed: int = 0, masters_per_size: int = 3, stripe_period: int = 24,
              stripe_width: int = 2, rules=(0, 0, 2, 3, 1)):
    """ICCAD 2017-style technology on device inputs (grid / cells dicts of
    lg.device_inputs), in place: vertical layer-1 power stripes of
    `stripe_width` sites every `stripe_period` sites over all rows, short
    layer-2 pieces on every 7th row; `masters_per_size` masters for each cell
    size with edge types 0-2 and one to three signal-pin boxes on layers 0
    and 1 (flattened x * h + y in the cell); the edge spacing rule table
    `rules` (index a + b: spacing in sites between edge types a and b)."""
    rng = np.random.default_rng(seed)
    xs, ys = int(grid["x_size"]), int(grid["y_size"])
    pm = np.full((xs, ys), -1, np.int32)
    for x0 in range(int(rng.integers(0, stripe_period)), xs, stripe_period):
        pm[x0:x0 + stripe_width, :] = 1
    for y0 in range(3, ys, 7):
        for _ in range(max(1, xs // 400)):
            x0 = int(rng.integers(0, max(xs - 12, 1)))
            pm[x0:x0 + 12, y0] = 2
    w = np.asarray(cells["w_grid"], np.int64)
    h = np.asarray(cells["h_grid"], np.int64)
    sizes, inv = np.unique(np.stack([w, h], 1), axis=0, return_inverse=True)
    inv = inv.ravel()
    M = len(sizes) * masters_per_size
    mstart, pb, pe, pl = [0], [], [], []
    m_led = rng.integers(0, 3, M).astype(np.int32)
    m_red = rng.integers(0, 3, M).astype(np.int32)
    for s, (sw_, sh_) in enumerate(sizes):
        for _ in range(masters_per_size):
            for _ in range(int(rng.integers(1, 4))):
                x1, x2 = sorted(int(v) for v in rng.integers(0, sw_, 2))
                y1, y2 = sorted(int(v) for v in rng.integers(0, sh_, 2))
                pb.append(x1 * int(sh_) + y1)
                pe.append(x2 * int(sh_) + y2)
                pl.append(int(rng.integers(0, 2)))
            mstart.append(len(pb))
    macro = (inv * masters_per_size + rng.integers(0, masters_per_size, len(w))).astype(np.int32)
    grid["power_layer"] = pm.ravel()
    grid["edge_rules"] = np.asarray(rules, np.int32)
    cells.update(ledge_type=m_led[macro], redge_type=m_red[macro], macro_id=macro,
                 macro_pin_start=np.array(mstart, np.int32), pin_bgn=np.array(pb, np.int32),
                 pin_end=np.array(pe, np.int32), pin_layer=np.array(pl, np.int32))
    return grid, cells


# ---------------------------------------------------------------------------
# Electrostatic density workloads (global-placement snapshots)
# ---------------------------------------------------------------------------
def density_placement(num_cells: int = 1_000_000, *, fill: float = 0.7, clusters: int = 8,
                      cluster_frac: float = 0.5, cluster_sigma: float = 0.04, num_fixed: int = 0,
                      macro_frac: float = 0.0, filler_frac: float = 0.0, seed: int = 0) -> dict:
    """Square cells with area ~ LogNormal(ln 4, 0.5) (site units) on a square
    die sized for `fill` movable area / die area; centres half uniform, half
    drawn from `clusters` Gaussian clusters (sigma = cluster_sigma x die side).
    Optional fixed rectangles, movable macros (x8 side) and filler cells."""
    rng = np.random.default_rng(seed)
    area = rng.lognormal(np.log(4.0), 0.5, num_cells)
    side = np.sqrt(area)
    is_macro = np.zeros(num_cells, np.int32)
    if macro_frac > 0:
        m = rng.random(num_cells) < macro_frac
        is_macro[m] = 1
        side[m] *= 8.0
    is_filler = (rng.random(num_cells) < filler_frac).astype(np.int32)
    die_side = float(math.sqrt(float((side ** 2).sum()) / fill))
    n_cl = int(num_cells * cluster_frac)
    cx = rng.random(num_cells) * die_side
    cy = rng.random(num_cells) * die_side
    if clusters > 0 and n_cl > 0:
        centres = rng.random((clusters, 2)) * die_side * 0.8 + 0.1 * die_side
        which = rng.integers(0, clusters, n_cl)
        cx[:n_cl] = centres[which, 0] + rng.normal(0, cluster_sigma * die_side, n_cl)
        cy[:n_cl] = centres[which, 1] + rng.normal(0, cluster_sigma * die_side, n_cl)
    cx = np.clip(cx, side / 2, die_side - side / 2)
    cy = np.clip(cy, side / 2, die_side - side / 2)
    fixed = np.zeros((num_fixed, 4), np.float32)
    for f in range(num_fixed):
        w, h = rng.uniform(0.02, 0.08, 2) * die_side
        x, y = rng.random(2) * (die_side - np.array([w, h]))
        fixed[f] = (x, y, x + w, y + h)
    return dict(die=(0.0, 0.0, die_side, die_side), dx=side.astype(np.float32), dy=side.astype(np.float32),
                cx=cx.astype(np.float32), cy=cy.astype(np.float32), is_macro=is_macro, is_filler=is_filler,
                fixed_rects=fixed, name=f"density_{num_cells}")


# ---------------------------------------------------------------------------
# Netlists for the wirelength gradient (global placement)
# ---------------------------------------------------------------------------
def _power_law_alpha(lo: int, hi: int, mean: float) -> float:
    """Exponent a of P(d) ~ d^-a on {lo..hi} whose mean is `mean` (bisection)."""
    d = np.arange(lo, hi + 1, dtype=np.float64)

    def m(a):
        w = d ** -a
        return float((d * w).sum() / w.sum())

    a0, a1 = 0.0, 10.0
    for _ in range(100):
        am = 0.5 * (a0 + a1)
        if m(am) > mean:
            a0 = am
        else:
            a1 = am
    return 0.5 * (a0 + a1)


def netlist(cell_w, cell_h, num_nets: int, *, die=(0.0, 0.0, 1.0, 1.0), min_deg: int = 2,
            max_deg: int = 1000, mean_deg: float = 3.8, io_frac: float = 0.005, seed: int = 0) -> dict:
    """Random netlist over cells of sizes (cell_w, cell_h): net degrees from a
    power law on [min_deg, max_deg] with mean `mean_deg`; each pin sits on a
    uniformly drawn cell at a uniform offset inside it (offset = pin centre -
    cell centre), or -- with probability io_frac -- is an IO pin of no
    movable cell at a uniform position on the die.  Pins are numbered net by
    net (the reference's layout)."""
    rng = np.random.default_rng(seed)
    n_cells = len(cell_w)
    a = _power_law_alpha(min_deg, max_deg, mean_deg)
    d = np.arange(min_deg, max_deg + 1)
    p = d.astype(np.float64) ** -a
    deg = rng.choice(d, size=num_nets, p=p / p.sum()).astype(np.int64)
    net_start = np.zeros(num_nets + 1, np.int64)
    np.cumsum(deg, out=net_start[1:])
    n_pins = int(net_start[-1])
    pin2cell = rng.integers(0, n_cells, n_pins).astype(np.int32)
    io = rng.random(n_pins) < io_frac
    pin2cell[io] = -1
    cw = np.asarray(cell_w, np.float64)
    ch = np.asarray(cell_h, np.float64)
    c = np.where(io, 0, pin2cell)
    off_x = ((rng.random(n_pins) - 0.5) * cw[c]).astype(np.float32)
    off_y = ((rng.random(n_pins) - 0.5) * ch[c]).astype(np.float32)
    off_x[io] = 0.0
    off_y[io] = 0.0
    pin_x = (die[0] + rng.random(n_pins) * (die[2] - die[0])).astype(np.float32)
    pin_y = (die[1] + rng.random(n_pins) * (die[3] - die[1])).astype(np.float32)
    return dict(net_start=net_start.astype(np.int32), pin2cell=pin2cell, pin_offset_x=off_x,
                pin_offset_y=off_y, pin_x=pin_x, pin_y=pin_y,
                net_weight=np.ones(num_nets, np.float32), num_cells=n_cells, alpha=a)
