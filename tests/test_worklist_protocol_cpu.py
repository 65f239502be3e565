"""Host emulation of a persistent-grid form of the primal launch: a fixed
grid of workgroups, each taking entries of its class's work list through an
atomic counter, fed by the work lists the scan kernel writes (scan_body in
admm_legalization_amd/csrc/lg_kernels.hip) over a run of iterations that
includes adaptive-search iterations (every partition listed in class 0 with
no cells, class 1 empty) and iterations with no active partition (both lists
empty).

Workgroups are Python generators that yield at every access to shared
state; a seeded scheduler interleaves them with at most `resident` of them
on the device at once (a workgroup starts when a resident one exits), which
is how a grid larger than the chip's occupancy runs, or how a grid sized to
the occupancy runs while another kernel holds part of the chip.

Three ways to make the counter start at zero in every launch are checked:
  * "producer": the scan kernel zeroes it when it writes the list counts (the
    stream orders that before the launch);
  * "consumer": the last workgroup to finish zeroes it (a finish counter);
  * a grid barrier after the list is drained (the workgroups wait for each
    other).
Only the first is correct under every interleaving and residency: after one
workgroup leaves early, the consumer's finish count is one short for good,
so later launches start past the head of their lists or reset the counter
in the middle of a launch (partitions skipped, or taken twice at once); the
barrier never completes when fewer workgroups than the grid are resident.  The device code has no persistent launch; one workgroup per
listed partition takes its entry by block index.
"""
import numpy as np
import pytest

K_NPAD_SMALL = 512  # the four-wave class bound (kNpadSmall)


def scan_lists(counts, act, adaptive, skip_ok):
    """The two work lists of one iteration (scan_body's listing rule): a
    partition is listed unless the skip test holds and its activity flag is
    clear; class 1 takes partitions above K_NPAD_SMALL cells except in the
    adaptive-search iteration, whose primal step orders no cell."""
    run = np.ones(len(counts), bool) if (adaptive or not skip_ok) else act != 0
    big = (~np.bool_(adaptive)) & (counts > K_NPAD_SMALL)
    return [list(np.flatnonzero(run & ~big)), list(np.flatnonzero(run & big))]


class Device:
    """Shared state of one class: list, count, take counter, finish counter,
    barrier counter."""

    def __init__(self):
        self.work, self.n = [], 0
        self.take = 0
        self.finished = 0
        self.arrived = 0


def workgroup(dev, grid, protocol, taken, early_exit_at=None):
    """One persistent workgroup; yields at each shared access."""
    while True:
        yield
        idx = dev.take
        dev.take += 1  # atomicAdd(&take, 1)
        yield
        if idx >= dev.n:
            break
        if early_exit_at is not None and idx == early_exit_at:
            return  # e.g. an early `return` of the per-partition body kept in the loop
        taken.append(dev.work[idx])
    if protocol == "consumer":
        yield
        dev.finished += 1
        if dev.finished == grid:  # the last one out resets both counters
            dev.take = 0
            dev.finished = 0
    elif protocol == "barrier":
        yield
        dev.arrived += 1
        while dev.arrived < grid:  # grid-wide barrier
            yield
        dev.take = 0


def run_launch(dev, grid, resident, protocol, rng, early_exit_at=None, max_steps=200000):
    """Runs one launch; returns (entries taken, finished without deadlock)."""
    taken = []
    pending = [workgroup(dev, grid, protocol, taken, early_exit_at) for _ in range(grid)]
    live = []
    steps = 0
    while pending or live:
        while pending and len(live) < resident:
            live.append(pending.pop(0))
        i = int(rng.integers(len(live)))
        try:
            next(live[i])
        except StopIteration:
            live.pop(i)
        steps += 1
        if steps > max_steps:
            return taken, False  # every resident workgroup is waiting
    return taken, True


def iterations(rng, K=40, n_iter=12):
    """Work lists of a run: perturbation iterations (all listed), normal ones
    with random activity, two adaptive-search iterations and an idle one."""
    counts = rng.integers(0, 900, K)
    out = []
    for it in range(n_iter):
        adaptive = it in (5, 6)
        skip_ok = it >= 2 and not adaptive
        act = rng.integers(0, 2, K) if it != 8 else np.zeros(K, int)
        out.append(scan_lists(counts, act, adaptive, skip_ok))
    return out


def test_listing_rule_of_the_adaptive_and_idle_iterations():
    rng = np.random.default_rng(0)
    its = iterations(rng)
    assert its[5][1] == [] and len(its[5][0]) == 40  # adaptive: all in class 0, class 1 empty
    assert its[8] == [[], []]  # no active partition
    assert its[0][0] and its[0][1]


@pytest.mark.parametrize("resident", [1, 3, 8])
def test_producer_reset_takes_every_entry_once(resident):
    rng = np.random.default_rng(resident)
    devs = [Device(), Device()]
    for lists in iterations(rng):
        for cls in (0, 1):
            d = devs[cls]
            d.work, d.n, d.take = lists[cls], len(lists[cls]), 0  # written by the scan kernel
            taken, ok = run_launch(d, grid=8, resident=resident, protocol="producer", rng=rng)
            assert ok
            assert sorted(taken) == sorted(lists[cls])  # each listed partition exactly once


def test_consumer_reset_is_lost_after_one_early_exit():
    rng = np.random.default_rng(1)
    d = Device()
    its = iterations(rng)
    got = []
    for n, lists in enumerate(its):
        d.work, d.n = lists[0], len(lists[0])
        # one workgroup leaves early in the adaptive-search iteration
        taken, ok = run_launch(d, grid=8, resident=8, protocol="consumer", rng=rng,
                               early_exit_at=0 if n == 5 else None)
        assert ok
        got.append(sorted(taken))
    assert all(g == sorted(lst[0]) for g, lst in zip(got[:5], its[:5]))
    # from then on the finish counter is one short: a later launch starts
    # with its take counter past the head of its list (those partitions take
    # no primal step), and the counter is reset in the middle of a launch
    # when the stale count reaches the grid (partitions taken twice, by two
    # workgroups at once) -- no error is raised either way
    wrong = [n for n in range(6, len(its)) if got[n] != sorted(its[n][0])]
    assert wrong
    assert any(len(set(got[n])) < len(got[n]) for n in wrong)


def test_consumer_reset_without_early_exit_is_correct():
    rng = np.random.default_rng(2)
    d = Device()
    for lists in iterations(rng):
        d.work, d.n = lists[0], len(lists[0])
        taken, ok = run_launch(d, grid=8, resident=3, protocol="consumer", rng=rng)
        assert ok and sorted(taken) == sorted(lists[0])


@pytest.mark.parametrize("resident,completes", [(8, True), (7, False), (1, False)])
def test_grid_barrier_needs_every_workgroup_resident(resident, completes):
    rng = np.random.default_rng(3)
    d = Device()
    lists = iterations(rng)[0]
    d.work, d.n = lists[0], len(lists[0])
    taken, ok = run_launch(d, grid=8, resident=resident, protocol="barrier", rng=rng, max_steps=20000)
    assert ok == completes
    if completes:
        assert sorted(taken) == sorted(lists[0])
