"""CPU checks of the dataflow schedule's work queues (mode 3 of the uneven
batch; fftamd_describe_flow, host-only).

The dataflow launch completes without a grid barrier because of how its
queues are laid out: every tile of a transform sits in one queue, its
second-pass tiles after all of its first-pass tiles (here: after every
first-pass tile of the queue), and each queue is taken in order.  These
tests check that layout on the 448-transform workload and on small mixes,
and replay the launch with a few workers (one, three, five, nine; chunked
tickets; a worker moving on to the other queues when its own is drained; a
worker counting its finished first-pass tiles only when it publishes them --
before a second-pass tile, on leaving a queue, or at 64 pending runs -- and a
second-pass tile blocking its worker until its transform's first-pass tiles
are counted) to show that every tile finishes whatever the number of resident
workgroups.
"""
import ctypes
from collections import defaultdict

import numpy as np
import pytest

QUEUES = 8
TILES = {32: (4, 4), 64: (32, 32), 128: (256, 128)}  # first-pass, second-pass tiles per transform


def describe_flow(fl, sides):
    lib = fl.load()
    dims = (ctypes.c_uint * (3 * len(sides)))(*[n for n in sides for _ in range(3)])
    qs = (ctypes.c_uint * (QUEUES + 1))()
    n = lib.fftamd_describe_flow(len(sides), dims, qs, None, 0)
    assert n >= 0, fl.last_error()
    buf = (ctypes.c_uint * (3 * n))()
    assert lib.fftamd_describe_flow(len(sides), dims, qs, buf, 3 * n) == n
    return list(qs), np.frombuffer(buf, dtype=np.uint32).reshape(n, 3).astype(np.int64)


def batch_448():
    from fftfpga_amd import datagen
    return datagen.batch_order({32: 256, 64: 160, 128: 32}, 2)


CASES = {
    "448": batch_448,
    "one_32": lambda: [32],
    "one_128": lambda: [128],
    "mix": lambda: [64, 32, 128, 32, 64, 128, 32, 32, 64, 128, 32],
    "many_32": lambda: [32] * 40,
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_every_tile_once_in_one_queue_second_pass_last(fl, case):
    sides = CASES[case]()
    qs, items = describe_flow(fl, sides)
    assert qs[0] == 0 and qs[-1] == len(items) and all(a <= b for a, b in zip(qs, qs[1:]))
    assert len(items) == sum(sum(TILES[n]) for n in sides)
    seen = set()
    queue_of = {}
    for q in range(QUEUES):
        last_first = {}  # transform -> position of its last first-pass tile
        first_second = {}
        counts = defaultdict(int)
        for pos in range(qs[q], qs[q + 1]):
            t, p, sub = (int(v) for v in items[pos])
            assert 0 <= t < len(sides) and p in (0, 1)
            assert 0 <= sub < TILES[sides[t]][p]
            assert (t, p, sub) not in seen
            seen.add((t, p, sub))
            assert queue_of.setdefault(t, q) == q, "a transform's tiles span two queues"
            counts[(t, p)] += 1
            if p == 0:
                assert t not in first_second, "a first-pass tile after a second-pass tile of its transform"
                last_first[t] = pos
            else:
                first_second.setdefault(t, pos)
        last_any_first = max(last_first.values(), default=-1)
        for t, pos in first_second.items():
            assert counts[(t, 0)] == TILES[sides[t]][0]
            assert pos > last_any_first, "second-pass tiles follow every first-pass tile of the queue"
    assert len(seen) == len(items)


@pytest.mark.parametrize("case", sorted(CASES))
def test_queues_balanced_by_bytes(fl, case):
    sides = CASES[case]()
    qs, items = describe_flow(fl, sides)
    load = np.zeros(QUEUES)
    for q in range(QUEUES):
        ts = {int(t) for t in items[qs[q]:qs[q + 1], 0]}
        load[q] = sum(sides[t] ** 3 for t in ts)
    assert load.max() - load.min() <= max(sides) ** 3


PEND = 64  # pending runs a workgroup publishes at the latest


def replay(sides, qs, items, workers, chunk):
    """The kernel's loop with `workers` resident workgroups, one tile per
    step each.  Returns the steps taken, or None when no worker can move."""
    need = {t: TILES[n][0] for t, n in enumerate(sides)}
    done = defaultdict(int)
    head = [0] * QUEUES
    state = [[0, [], []] for _ in range(workers)]  # queue offset, held tiles, pending runs
    finished = [False] * workers
    remaining = len(items)
    steps = 0

    def publish(pend):
        for t, c in pend:
            done[t] += c
        pend.clear()

    while remaining:
        moved = False
        for w in range(workers):
            if finished[w]:
                continue
            qi, held, pend = state[w]
            while not held and qi < QUEUES:  # take a ticket (own queue first, then the others)
                q = (w % QUEUES + qi) % QUEUES
                count = qs[q + 1] - qs[q]
                if head[q] >= count:
                    publish(pend)  # before leaving the queue
                    qi += 1
                    continue
                t0 = head[q]
                head[q] = min(count, t0 + chunk)
                held = [qs[q] + t for t in range(t0, head[q])]
            state[w] = [qi, held, pend]
            if not held:
                finished[w] = True
                moved = True
                continue
            t, p, _ = (int(v) for v in items[held[0]])
            if p == 1:
                if pend:
                    publish(pend)  # before it waits
                    moved = True
                if done[t] < need[t]:
                    continue  # waits on its transform's first pass
            else:
                if pend and pend[-1][0] == t:
                    pend[-1][1] += 1
                else:
                    pend.append([t, 1])
                if len(pend) == PEND:
                    publish(pend)
            held.pop(0)
            remaining -= 1
            moved = True
        steps += 1
        if not moved:
            return None
    for w in range(workers):
        publish(state[w][2])
    assert all(done[t] == need[t] for t in need)
    return steps


@pytest.mark.parametrize("case", ["one_32", "one_128", "mix", "many_32"])
@pytest.mark.parametrize("workers,chunk", [(1, 1), (3, 2), (9, 4)])
def test_finishes_with_any_number_of_workers(fl, case, workers, chunk):
    sides = CASES[case]()
    qs, items = describe_flow(fl, sides)
    assert replay(sides, qs, items, workers, chunk) is not None


def test_448_finishes_with_few_workers(fl):
    sides = batch_448()
    qs, items = describe_flow(fl, sides)
    assert replay(sides, qs, items, 5, 4) is not None


def test_other_shapes_refused(fl):
    lib = fl.load()
    dims = (ctypes.c_uint * 6)(32, 32, 32, 64, 64, 32)
    qs = (ctypes.c_uint * (QUEUES + 1))()
    assert lib.fftamd_describe_flow(2, dims, qs, None, 0) == -1
    assert "64" in fl.last_error() or "transform" in fl.last_error()
