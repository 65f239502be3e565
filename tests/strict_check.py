"""Independent strict legality counts (numpy) and planted violations, shared
by the oracle's CPU tests and the device's GPU tests.

The counts follow the placement filters of the reference's
computePenaltySum (AdmmKernels.hpp:355-373) and countRowOrientVioKernel
(DatabaseKernels.hpp:468-504): a footprint must lie inside the grid, on valid
sites, of the cell's own fence group; an odd-height cell is flipped exactly
when its ground pin side disagrees with the row's rail; an even-height cell
(never flipped) must sit on a row whose rail matches.
"""
import numpy as np


def numpy_strict_counts(grid, cells, lx, ly, rot):
    xs, ys = int(grid["x_size"]), int(grid["y_size"])
    valid = np.asarray(grid["valid"]).astype(bool)
    group = np.asarray(grid["group"])
    vdd = np.asarray(grid["vdd_up"])
    usage = np.zeros(xs * ys, np.int64)
    out = dict(total_ovf=0, num_row_orient_vio=0, num_site_vio=0, num_fence_vio=0, num_core_vio=0,
               num_rail_vio=0)
    for i in range(len(lx)):
        w, h = int(cells["w_grid"][i]), int(cells["h_grid"][i])
        x0, y0 = int(lx[i]), int(ly[i])
        xx, yy = np.meshgrid(np.arange(x0, x0 + w), np.arange(y0, y0 + h), indexing="ij")
        inside = (xx >= 0) & (xx < xs) & (yy >= 0) & (yy < ys)
        out["num_core_vio"] += int(not inside.all())
        pid = (xx[inside] * ys + yy[inside]).ravel()
        usage[pid] += 1
        site_bad = ~valid[pid]
        out["num_site_vio"] += int(site_bad.any())
        out["num_fence_vio"] += int((group[pid][~site_bad] != int(cells["group"][i])).any())
        row_vdd = int(vdd[y0]) if 0 <= y0 < ys else -1
        gnd = int(cells["ground_at_bottom"][i])
        if h % 2 == 1:
            out["num_row_orient_vio"] += int((rot[i] == 0 and gnd != row_vdd) or (rot[i] == 1 and gnd == row_vdd))
        else:
            bad = int(gnd != row_vdd)
            out["num_row_orient_vio"] += bad
            out["num_rail_vio"] += bad
    out["total_ovf"] = int(np.maximum(usage - 1, 0).sum())
    return out


def plant_violations(grid, cells, seed=0):
    """A copy of `cells` whose preplaced positions (dbu) put a few cells on
    invalid sites, in a foreign fence group and (even heights) on the wrong
    rail.  Returns (cells, {kind: [cell ids]})."""
    rng = np.random.default_rng(seed)
    xs, ys = int(grid["x_size"]), int(grid["y_size"])
    sw, rh = int(grid["site_width"]), int(grid["row_height"])
    cx, cy = int(grid["core_lx"]), int(grid["core_ly"])
    valid = np.asarray(grid["valid"]).reshape(xs, ys).astype(bool)
    group = np.asarray(grid["group"]).reshape(xs, ys)
    vdd = np.asarray(grid["vdd_up"])
    c = {k: np.array(v, copy=True) for k, v in cells.items()}
    w, h, g, gnd = c["w_grid"], c["h_grid"], c["group"], c["ground_at_bottom"]
    planted = {"site": [], "fence": [], "rail": []}

    def put(i, x, y):
        c["lx_dbu"][i] = cx + x * sw
        c["ly_dbu"][i] = cy + y * rh

    single = np.flatnonzero((h == 1) & (g == -1))
    # on an invalid pixel (a macro or blockage): its lower-left corner there
    bad = np.argwhere(~valid[: xs - 8, : ys - 1])
    for i, (x, y) in zip(single[:3], bad[rng.choice(len(bad), 3, replace=False)]):
        if int(w[i]) <= xs - x:
            put(i, int(x), int(y))
            planted["site"].append(int(i))
    # default-region cells on a valid fence pixel
    fence = np.argwhere(valid[: xs - 8, :] & (group[: xs - 8, :] >= 0))
    if len(fence):
        for i, (x, y) in zip(single[3:6], fence[rng.choice(len(fence), 3, replace=False)]):
            put(i, int(x), int(y))
            planted["fence"].append(int(i))
    # even-height cells one row off their rail
    for i in np.flatnonzero((h == 2) & (g == -1))[:3]:
        y = (int(c["ly_dbu"][i]) - cy) // rh
        y2 = y + 1 if y + 1 + 2 <= ys else y - 1
        if vdd[y2] != gnd[i]:
            put(i, (int(c["lx_dbu"][i]) - cx) // sw, y2)
            planted["rail"].append(int(i))
    return c, planted
