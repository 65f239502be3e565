"""Digests of legalization results, shared by the fixture writer
(make_config_golden.py) and the GPU tests that compare against it.

A digest is the SHA-256 of the movable cells' legalized lower-left corners
(dbu) and flip flags, as int32 arrays in movable-instance order (the order in
which the legalizer imports instances, reference DatabaseIO.cpp:100-149).
Equal digests mean bit-identical positions.
"""
import hashlib

import numpy as np


def position_digest(lx_dbu, ly_dbu, flipped) -> str:
    h = hashlib.sha256()
    for a in (lx_dbu, ly_dbu, flipped):
        h.update(np.ascontiguousarray(a, dtype=np.int32).tobytes())
    return h.hexdigest()


def array_digest(a, dtype=np.int32) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype=dtype).tobytes()).hexdigest()


def grid_to_dbu(grid: dict, lx_grid, ly_grid):
    """Grid positions of the device layer (sites, rows) -> dbu."""
    lx = np.asarray(lx_grid, np.int64) * int(grid["site_width"]) + int(grid["core_lx"])
    ly = np.asarray(ly_grid, np.int64) * int(grid["row_height"]) + int(grid["core_ly"])
    return lx.astype(np.int32), ly.astype(np.int32)
