"""Python front of the legalizer's request server (lg_srv_* in
include/skyline_lg.h).

`LegalizationServer` keeps a continuous-batching loop running on the GPU:
`submit()` queues a request (the device-level inputs of one legalization: a
grid and its preplaced cells, as `lg.device_inputs` returns them) and returns
a ticket at once; `wait(ticket)` blocks until that request is legalized and
returns its positions and statistics.  Requests that arrive while others run
are admitted as slots and device memory free up and run in the same batched
kernel launches; each gets exactly the result `lg.DeviceLegalizer` gives it
alone.  No counterpart in the reference (one design per process there).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .design import LgParams, LgStats, LgVioStat
from .lg import LegalizerError, _check, _p, lib, marshal_device_inputs

QUEUED, RUNNING, DONE, FAILED = 0, 1, 2, 3


class LgSrvConfig(C.Structure):
    _fields_ = [("max_slots", C.c_int32), ("iters_per_sync", C.c_int32),
                ("mem_budget_bytes", C.c_int64), ("pool_chunk_bytes", C.c_int64),
                ("use_graphs", C.c_int32), ("short_cells", C.c_int32), ("max_overtakes", C.c_int32),
                ("reserved", C.c_int32)]


class LgSrvResult(C.Structure):
    _fields_ = [("out_lx", C.POINTER(C.c_int32)), ("out_ly", C.POINTER(C.c_int32)),
                ("out_rot", C.POINTER(C.c_int32)), ("num_cells", C.c_int32), ("status", C.c_int32),
                ("stats", LgStats), ("queued_s", C.c_double), ("run_s", C.c_double),
                ("latency_s", C.c_double)]


class LgSrvCounters(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("failed", C.c_int64),
                ("batches", C.c_int64), ("iterations", C.c_int64), ("peak_slots", C.c_int32),
                ("active", C.c_int32), ("queued", C.c_int32), ("pool_reserved_bytes", C.c_int64),
                ("pool_peak_bytes", C.c_int64), ("graph_captures", C.c_int64)]


def _bind():
    L = lib()
    if not getattr(L, "_srv_bound", False):
        L.lg_srv_default_config.argtypes = [C.POINTER(LgSrvConfig)]
        L.lg_srv_create.argtypes = [C.POINTER(LgSrvConfig), C.POINTER(C.c_void_p)]
        L.lg_srv_submit.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(LgParams),
                                    C.POINTER(C.c_int64)]
        L.lg_srv_poll.argtypes = [C.c_void_p, C.c_int64, C.POINTER(C.c_int32)]
        L.lg_srv_wait.argtypes = [C.c_void_p, C.c_int64, C.POINTER(LgSrvResult)]
        L.lg_srv_counters.argtypes = [C.c_void_p, C.POINTER(LgSrvCounters)]
        L.lg_srv_profile.argtypes = [C.c_void_p, C.c_int]
        L.lg_srv_profile_read.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int64),
                                          C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.lg_srv_destroy.argtypes = [C.c_void_p]
        L._srv_bound = True
    return L


def default_config() -> LgSrvConfig:
    c = LgSrvConfig()
    _bind().lg_srv_default_config(C.byref(c))
    return c


class LegalizationServer:
    """Continuous-batching legalization service on the current GPU."""

    def __init__(self, max_slots: int | None = None, iters_per_sync: int | None = None,
                 mem_budget_bytes: int | None = None, pool_chunk_bytes: int | None = None,
                 use_graphs: bool | None = None, short_cells: int | None = None,
                 max_overtakes: int | None = None):
        cfg = default_config()
        if short_cells is not None:
            cfg.short_cells = int(short_cells)
        if max_overtakes is not None:
            cfg.max_overtakes = int(max_overtakes)
        if use_graphs is not None:
            cfg.use_graphs = 1 if use_graphs else 0
        if max_slots is not None:
            cfg.max_slots = int(max_slots)
        if iters_per_sync is not None:
            cfg.iters_per_sync = int(iters_per_sync)
        if mem_budget_bytes is not None:
            cfg.mem_budget_bytes = int(mem_budget_bytes)
        if pool_chunk_bytes is not None:
            cfg.pool_chunk_bytes = int(pool_chunk_bytes)
        self.config = cfg
        h = C.c_void_p()
        _check(_bind().lg_srv_create(C.byref(cfg), C.byref(h)))
        self.h = h
        self._sizes = {}

    def submit(self, grid: dict, cells: dict, params: LgParams) -> int:
        """Queues one legalization; returns its ticket.  The arrays are copied
        before the call returns."""
        g, c, keep = marshal_device_inputs(grid, cells)
        t = C.c_int64()
        _check(_bind().lg_srv_submit(self.h, C.byref(g), C.byref(c), C.byref(params), C.byref(t)))
        del keep
        self._sizes[t.value] = int(c.num_cells)
        return t.value

    def poll(self, ticket: int) -> int:
        s = C.c_int32()
        _check(_bind().lg_srv_poll(self.h, int(ticket), C.byref(s)))
        return s.value

    def wait(self, ticket: int) -> dict:
        """Blocks until the request is done; returns lx, ly, rot (grid
        units, per cell), the statistics and the request's timings.  Raises
        LegalizerError if the request failed."""
        n = self._sizes.pop(int(ticket), 0)
        lx, ly, rot = (np.zeros(max(n, 1), np.int32) for _ in range(3))
        r = LgSrvResult()
        r.out_lx, r.out_ly, r.out_rot = _p(lx), _p(ly), _p(rot)
        _check(_bind().lg_srv_wait(self.h, int(ticket), C.byref(r)))
        st = {k: getattr(r.stats, k) for k, _ in LgStats._fields_ if k != "vio"}
        st["vio"] = {k: getattr(r.stats.vio, k) for k, _ in LgVioStat._fields_}
        return dict(lx=lx[:n], ly=ly[:n], rot=rot[:n], stats=st, queued_s=r.queued_s,
                    run_s=r.run_s, latency_s=r.latency_s)

    def counters(self) -> dict:
        c = LgSrvCounters()
        _check(_bind().lg_srv_counters(self.h, C.byref(c)))
        return {k: getattr(c, k) for k, _ in LgSrvCounters._fields_}

    def profile(self, enable: bool = True):
        """Times the batched partition-kernel launches of every later batch."""
        _check(_bind().lg_srv_profile(self.h, int(bool(enable))))

    def profile_read(self) -> dict:
        """ms: summed partition-launch time; ms_union: the time during which
        any lane's partition launches ran."""
        ms, n, b, u = C.c_double(), C.c_int64(), C.c_double(), C.c_double()
        _check(_bind().lg_srv_profile_read(self.h, C.byref(ms), C.byref(n), C.byref(b), C.byref(u)))
        return dict(ms=ms.value, iterations=n.value, bytes=b.value, ms_union=u.value)

    def close(self):
        if self.h:
            _bind().lg_srv_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


__all__ = ["LegalizationServer", "LgSrvConfig", "LgSrvResult", "LgSrvCounters", "default_config",
           "LegalizerError", "QUEUED", "RUNNING", "DONE", "FAILED"]
