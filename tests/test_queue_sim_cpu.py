"""CPU simulation of the batched executor's device work queues (no GPU).

The persistent kernels of a batch claim items from one atomic counter per
queue and map each claim to (job, tile) with the library's queue mapping;
the same function is exported for the host as mgx_queue_map.  Here
workgroups claim in random interleavings through that exported copy, and the
invariants the reference's host job scheduler was meant to provide
(mat-gpu/job.cu:57-81 dependency edges, :183-336 runner) are checked: every
tile of every job is taken exactly once, empty jobs are never handed out,
every workgroup leaves after one failing claim (the counter ends at items +
grid), and a later stage starts only after every tile of the earlier one.
"""
import ctypes
import random

import pytest

i64, c_int = ctypes.c_int64, ctypes.c_int


@pytest.fixture(scope="module")
def qmap():
    import mgx
    L = mgx.lib()

    def f(prefix_arr, njobs, item):
        j, t = c_int(), i64()
        r = L.mgx_queue_map(prefix_arr, njobs, item, ctypes.byref(j), ctypes.byref(t))
        return (j.value, t.value) if r == 0 else None
    return f


def prefix_of(counts):
    p = [0]
    for c in counts:
        p.append(p[-1] + c)
    return (i64 * len(p))(*p), p


def simulate(qmap, counts, grid, rng, clock):
    """One persistent launch: `grid` workgroups claim in random order.
    Returns {(job, tile): time}, the final counter and the number of
    workgroups that left the loop."""
    arr, p = prefix_of(counts)
    counter = 0
    live = list(range(grid))
    taken = {}
    left = 0
    while live:
        w = rng.choice(live)
        item = counter
        counter += 1
        m = qmap(arr, len(counts), item)
        clock[0] += 1
        if m is None:
            live.remove(w)
            left += 1
            continue
        assert m not in taken, ("claimed twice", m)
        taken[m] = clock[0]
    return taken, counter, left, p


@pytest.mark.parametrize("seed", range(8))
def test_every_tile_exactly_once(qmap, seed):
    rng = random.Random(seed)
    njobs = rng.randint(1, 40)
    counts = [rng.choice([0, 0, 1, 2, rng.randint(1, 60)]) for _ in range(njobs)]
    grid = rng.choice([1, 3, 17, 64, 300])
    taken, counter, left, p = simulate(qmap, counts, grid, rng, [0])
    want = {(j, t) for j, c in enumerate(counts) for t in range(c)}
    assert set(taken) == want
    assert counter == p[-1] + grid and left == grid
    assert all(counts[j] > 0 for j, _ in taken)


def test_empty_queue_and_bounds(qmap):
    arr, p = prefix_of([0, 0, 0])
    assert qmap(arr, 3, 0) is None
    arr, p = prefix_of([5])
    assert [qmap(arr, 1, i) for i in range(6)] == [(0, t) for t in range(5)] + [None]
    assert qmap(arr, 1, -1) is None
    assert qmap(arr, 0, 0) is None   # no jobs at all


def test_int64_offsets(qmap):
    """Tile ranges beyond 2^32 (the counters and prefixes are 64-bit): items
    at every range boundary map into the right job."""
    counts = [3, 0, 1 << 33, 7, 0, (1 << 40) + 5, 1]
    arr, p = prefix_of(counts)
    for j, c in enumerate(counts):
        if c == 0:
            continue
        for item in (p[j], p[j] + c // 2, p[j + 1] - 1):
            assert qmap(arr, len(counts), item) == (j, item - p[j])
    assert qmap(arr, len(counts), p[-1]) is None


def test_stages_in_launch_order(qmap):
    """A batch launches its stages one after another (reduce, residue GEMM,
    reconstruct, carry for the multi-modular products): every tile of a
    stage is claimed before any tile of the next, for every product."""
    rng = random.Random(5)
    products = 12
    stages = []
    for s in range(4):   # tiles per product at each stage
        stages.append([rng.randint(0, 30) if s == 0 else rng.randint(1, 30) for _ in range(products)])
    clock = [0]
    times = []
    for counts in stages:
        taken, counter, left, p = simulate(qmap, counts, rng.choice([4, 16, 64]), rng, clock)
        assert left > 0 and counter == p[-1] + left
        first, last = {}, {}
        for (j, t), tm in taken.items():
            first[j] = min(first.get(j, tm), tm)
            last[j] = max(last.get(j, tm), tm)
        times.append((first, last))
    for s in range(3):
        for j in range(products):
            if j in times[s][1] and j in times[s + 1][0]:
                assert times[s][1][j] < times[s + 1][0][j]
