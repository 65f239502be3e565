"""Writes a synthetic Design as LEF + DEF text (test input for the LEF/DEF
reader), with the Design the reader should reconstruct from it.

Independent of the reader: the masters, pins, nets, fences and (optionally)
the ICCAD 2017 technology (edge types, signal-pin layers, power stripes in
SPECIALNETS, the cell edge spacing table) are laid out here, and the
expected values follow the LEF/DEF semantics directly (micron -> dbu as
int(micron * dbu), a pin's position as the integer average of its port's
corners, an IO pin's as the centre of its box).
"""
from __future__ import annotations

import numpy as np

from admm_legalization_amd.design import Design


def _um(v, dbu):
    return f"{v / dbu:.4f}"


def _dbu(s, dbu):
    return int(float(s) * float(dbu))


def write(d: Design, lef_path: str, def_path: str, *, pins_per_master: int = 3, seed: int = 0,
          tech: bool = False, stripe_period: int = 20):
    rng = np.random.default_rng(seed)
    dbu = int(d.dbu_per_micron)
    sw, rh = d.site_width, d.row_height
    n = d.num_insts
    # one master per (width, height, macro, ground-at-bottom, variant); two
    # variants per size so that masters of one size differ in pins and edges
    keys, master_of = {}, np.zeros(n, np.int64)
    for i in range(n):
        k = (int(d.inst_dx[i]), int(d.inst_dy[i]), int(d.inst_macro[i]), int(d.inst_ground_at_bottom[i]),
             int(rng.integers(0, 2)) if not d.inst_macro[i] else 0)
        master_of[i] = keys.setdefault(k, len(keys))
    masters = list(keys)
    M = len(masters)
    ledge = rng.integers(0, 3, M) if tech else np.zeros(M, np.int64)
    redge = rng.integers(0, 3, M) if tech else np.zeros(M, np.int64)
    lef = [f"VERSION 5.8 ;\nBUSBITCHARS \"[]\" ;\nDIVIDERCHAR \"/\" ;\n",
           f"UNITS\n  DATABASE MICRONS {dbu} ;\nEND UNITS\n",
           "PROPERTYDEFINITIONS\n  LIBRARY LEF58_CELLEDGESPACINGTABLE STRING ;\n"
           "  MACRO LEF58_EDGETYPE STRING ;\nEND PROPERTYDEFINITIONS\n"]
    rules = [(1, 1, 0.2), (1, 2, 0.3), (2, 2, 0.1)]
    if tech:
        tbl = " ".join(f"EDGETYPE {a} {b} {s}" for a, b, s in rules)
        lef.append(f"PROPERTY LEF58_CELLEDGESPACINGTABLE \"CELLEDGESPACINGTABLE {tbl} ;\" ;\n")
    for name, typ in (("metal1", "ROUTING"), ("via1", "CUT"), ("metal2", "ROUTING"), ("via2", "CUT"),
                      ("metal3", "ROUTING")):
        lef.append(f"LAYER {name}\n  TYPE {typ} ;\n  SPACING 0.1 ;\nEND {name}\n")
    lef.append("VIA via12 DEFAULT\n  LAYER metal1 ;\n  RECT -0.05 -0.05 0.05 0.05 ;\nEND via12\n")
    lef.append(f"SITE core\n  CLASS CORE ;\n  SIZE {_um(sw, dbu)} BY {_um(rh, dbu)} ;\nEND core\n")
    pin_avg = []     # per master: list of (name, avg_x, avg_y)
    sig = []         # per master: (layer, lx, ly, ux, uy) of each signal pin (dbu)
    for m, (w, h, mac, gnd, var) in enumerate(masters):
        name = f"M{m}"
        body = [f"MACRO {name}\n  CLASS {'BLOCK' if mac else 'CORE'} ;\n  ORIGIN 0 0 ;\n",
                f"  SIZE {_um(w, dbu)} BY {_um(h, dbu)} ;\n  SYMMETRY X Y ;\n  SITE core ;\n"]
        if tech and not mac:
            body.append(f"  PROPERTY LEF58_EDGETYPE \"EDGETYPE LEFT {ledge[m]} ; EDGETYPE RIGHT {redge[m]} ;\" ;\n")
        avgs, sp = [], []
        if not mac:
            # rails: the ground pin at the bottom when gnd, else the power pin
            lo, hi = (0, 60), (h - 60, h)
            for pname, use, (y0, y1) in (("VSS", "GROUND", lo if gnd else hi), ("VDD", "POWER", hi if gnd else lo)):
                body.append(f"  PIN {pname}\n    DIRECTION INOUT ;\n    USE {use} ;\n    PORT\n      LAYER metal1 ;\n"
                            f"      RECT 0 {_um(y0, dbu)} {_um(w, dbu)} {_um(y1, dbu)} ;\n    END\n  END {pname}\n")
            for k in range(pins_per_master):
                # a pin box on whole 10-dbu steps inside the cell
                x0 = int(rng.integers(0, max(w // 10 - 2, 1))) * 10
                y0 = int(rng.integers(1, max(h // 10 - 3, 2))) * 10
                x1 = min(x0 + 10 * int(rng.integers(1, 5)), w)
                y1 = min(y0 + 10 * int(rng.integers(1, 8)), h)
                layer = int(rng.integers(0, 2)) if tech else 0
                lname = ("metal1", "metal2")[layer]
                body.append(f"  PIN A{k}\n    DIRECTION INPUT ;\n    USE SIGNAL ;\n    PORT\n      LAYER {lname} ;\n"
                            f"      RECT {_um(x0, dbu)} {_um(y0, dbu)} {_um(x1, dbu)} {_um(y1, dbu)} ;\n    END\n"
                            f"  END A{k}\n")
                cx = [_dbu(_um(v, dbu), dbu) for v in (x0, x1, x1, x0)]
                cy = [_dbu(_um(v, dbu), dbu) for v in (y0, y0, y1, y1)]
                avgs.append((f"A{k}", sum(cx) // 4, sum(cy) // 4))
                sp.append((layer, min(cx), min(cy), max(cx), max(cy)))
        else:
            body.append(f"  OBS\n    LAYER metal1 ;\n    RECT 0 0 {_um(w, dbu)} {_um(h, dbu)} ;\n  END\n")
        body.append(f"END {name}\n")
        lef.append("".join(body))
        pin_avg.append(avgs)
        sig.append(sp)
    lef.append("END LIBRARY\n")
    with open(lef_path, "w") as f:
        f.write("".join(lef))

    names = []
    gr = np.asarray(d.inst_group)
    for i in range(n):  # group 0 by pattern ("g0/*"), the others by name
        names.append(f"g0/c{i}" if gr[i] == 0 else f"c{i}")
    core = d.core
    out = [f"VERSION 5.8 ;\nDIVIDERCHAR \"/\" ;\nBUSBITCHARS \"[]\" ;\nDESIGN synth ;\n",
           f"UNITS DISTANCE MICRONS {dbu} ;\n",
           f"DIEAREA ( {core[0]} {core[1]} ) ( {core[2]} {core[3]} ) ;\n"]
    for r in range(len(d.row_lx)):
        o = ("N", "S", "W", "E", "FN", "FS", "FW", "FE")[int(d.row_orient[r])]
        nx = int(d.row_dx[r]) // sw
        out.append(f"ROW row{r} core {int(d.row_lx[r])} {int(d.row_ly[r])} {o} DO {nx} BY 1 STEP {sw} 0 ;\n")
    out.append("TRACKS X 0 DO 10 STEP 100 LAYER metal1 ;\n")
    out.append("VIAS 1 ;\n- myvia\n + RECT metal1 ( -10 -10 ) ( 10 10 ) ;\nEND VIAS\n")
    out.append(f"COMPONENTS {n} ;\n")
    for i in range(n):
        st = "FIXED" if d.inst_fixed[i] else "PLACED"
        out.append(f"- {names[i]} M{master_of[i]} + SOURCE NETLIST + {st} ( {int(d.inst_lx[i])} "
                   f"{int(d.inst_ly[i])} ) N ;\n")
    out.append("END COMPONENTS\n")
    # nets: every pin of the design's nets on one of its master's signal pins
    ns = np.asarray(d.net_start)
    io_pins, nets_txt = [], []
    exp_inst, exp_dx, exp_dy, exp_start = [], [], [], [0]
    for k in range(len(ns) - 1):
        parts = []
        for p in range(int(ns[k]), int(ns[k + 1])):
            inst = int(d.pin_inst[p])
            if inst < 0 or d.inst_macro[inst] or not pin_avg[master_of[inst]]:
                x, y = (int(d.pin_dx[p]), int(d.pin_dy[p])) if inst < 0 else (int(d.inst_lx[inst]), int(d.inst_ly[inst]))
                pname = f"io{len(io_pins)}"
                io_pins.append((pname, f"n{k}", x, y))
                parts.append(f"( PIN {pname} )")
                exp_inst.append(-1)
                exp_dx.append((x - 20 + x + 20) // 2)
                exp_dy.append((y - 20 + y + 20) // 2)
            else:
                avgs = pin_avg[master_of[inst]]
                pn, ax, ay = avgs[p % len(avgs)]
                parts.append(f"( {names[inst]} {pn} )")
                exp_inst.append(inst)
                exp_dx.append(ax)
                exp_dy.append(ay)
        exp_start.append(len(exp_inst))
        nets_txt.append(f"- n{k} " + " ".join(parts) + " + USE SIGNAL ;\n")
    out.append(f"PINS {len(io_pins)} ;\n")
    for pname, net, x, y in io_pins:
        out.append(f"- {pname} + NET {net} + DIRECTION INPUT + USE SIGNAL\n  + LAYER metal2 ( -20 -20 ) ( 20 20 )\n"
                   f"  + PLACED ( {x} {y} ) N ;\n")
    out.append("END PINS\n")
    out.append("BLOCKAGES 1 ;\n- PLACEMENT RECT ( 0 0 ) ( 10 10 ) ;\nEND BLOCKAGES\n")
    # fences: a region per group (box order kept), the group's members
    ng = int(d.num_groups)
    bg = np.asarray(d.box_group)
    out.append(f"REGIONS {ng} ;\n")
    for g in range(ng):
        boxes = " ".join(f"( {int(d.box_lx[b])} {int(d.box_ly[b])} ) ( {int(d.box_ux[b])} {int(d.box_uy[b])} )"
                         for b in range(len(bg)) if bg[b] == g)
        out.append(f"- reg{g} {boxes} + TYPE FENCE ;\n")
    out.append("END REGIONS\n")
    wires = []
    if tech:
        out.append("SPECIALNETS 2 ;\n")
        x = core[0] + int(rng.integers(0, stripe_period)) * sw
        segs = []
        while x < core[2]:
            segs.append(f"NEW metal2 {2 * sw} ( {x} {core[1]} ) ( * {core[3]} )")
            wires.append((1, x, core[1], x, core[3], 2 * sw))
            x += stripe_period * sw
        for r in range(0, len(d.row_ly), 9):  # short metal3 pieces along some rows
            y = int(d.row_ly[r])
            x0 = core[0] + int(rng.integers(0, max((core[2] - core[0]) // sw - 20, 1))) * sw
            segs.append(f"NEW metal3 {rh // 4} ( {x0} {y} ) ( {x0 + 15 * sw} * )")
            wires.append((2, x0, y, x0 + 15 * sw, y, rh // 4))
        segs.append(f"NEW metal1 {rh // 8} ( {core[0]} {core[1]} ) ( {core[2]} * )")  # ignored layer
        wires.append((0, core[0], core[1], core[2], core[1], rh // 8))
        body = "\n  ".join(segs)[4:]  # the first wire is + ROUTED, not NEW
        out.append(f"- VDD ( * VDD ) + ROUTED {body} + USE POWER ;\n")
        out.append("- VSS ( * VSS ) + USE GROUND ;\n")
        out.append("END SPECIALNETS\n")
    out.append(f"NETS {len(nets_txt)} ;\n")
    out.extend(nets_txt)
    out.append("END NETS\n")
    out.append(f"GROUPS {ng} ;\n")
    for g in range(ng):
        mem = "g0/*" if g == 0 else " ".join(names[i] for i in range(n) if gr[i] == g)
        out.append(f"- grp{g} {mem} + REGION reg{g} ;\n")
    out.append("END GROUPS\nEND DESIGN\n")
    with open(def_path, "w") as f:
        f.write("".join(out))

    # what the reader should give
    gnd = np.where(d.inst_macro != 0, 1, d.inst_ground_at_bottom).astype(np.uint8)
    exp = Design(core=tuple(int(v) for v in core), row_lx=d.row_lx, row_ly=d.row_ly, row_dx=d.row_dx,
                 row_dy=d.row_dy, row_site_width=d.row_site_width, row_orient=d.row_orient,
                 inst_lx=d.inst_lx, inst_ly=d.inst_ly, inst_dx=d.inst_dx, inst_dy=d.inst_dy,
                 inst_fixed=d.inst_fixed, inst_macro=d.inst_macro, inst_group=d.inst_group,
                 inst_ground_at_bottom=gnd, num_groups=ng, box_group=d.box_group, box_lx=d.box_lx,
                 box_ly=d.box_ly, box_ux=d.box_ux, box_uy=d.box_uy,
                 net_start=np.array(exp_start, np.int32), pin_inst=np.array(exp_inst, np.int32),
                 pin_dx=np.array(exp_dx, np.int32), pin_dy=np.array(exp_dy, np.int32), inst_name=names,
                 dbu_per_micron=dbu)
    info = dict(master_of=master_of, masters=masters, ledge=ledge, redge=redge, sig=sig, wires=wires,
                rules=rules if tech else [])
    return exp, info
