"""CPU simulation of the batched executor's device work queues (no GPU).

The persistent kernels of a batch claim items from eight counters per queue
(contiguous item ranges, one per XCD; a workgroup starts at its XCD's range
and moves on when a claim there fails) and map each claim to (job, tile)
with the library's queue mapping; both functions are exported for the host
(mgx_queue_claim, the same claim step as the device code with plain
increments in place of atomics, and mgx_queue_map).  Here workgroups make
single claim attempts in random interleavings through those exported copies,
and the invariants the reference's host job scheduler was meant to provide
(mat-gpu/job.cu:57-81 dependency edges, :183-336 runner) are checked: every
tile of every job is taken exactly once, empty jobs are never handed out,
every workgroup leaves after exactly one failing claim per range (counter x
ends at the size of range x + grid), and a later stage starts only after
every tile of the earlier one.
"""
import ctypes
import random

import pytest

i64, c_int = ctypes.c_int64, ctypes.c_int


RANGES = 8


@pytest.fixture(scope="module")
def qmap():
    import mgx
    L = mgx.lib()

    def f(prefix_arr, njobs, item):
        j, t = c_int(), i64()
        r = L.mgx_queue_map(prefix_arr, njobs, item, ctypes.byref(j), ctypes.byref(t))
        return (j.value, t.value) if r == 0 else None
    return f


@pytest.fixture(scope="module")
def qclaim():
    import mgx
    L = mgx.lib()

    def f(counters, total, state):
        """One claim attempt of a workgroup whose state is [cur, moved]."""
        cur, moved = c_int(state[0]), c_int(state[1])
        r = L.mgx_queue_claim(counters, total, ctypes.byref(cur), ctypes.byref(moved))
        state[0], state[1] = cur.value, moved.value
        return r
    return f


def prefix_of(counts):
    p = [0]
    for c in counts:
        p.append(p[-1] + c)
    return (i64 * len(p))(*p), p


def simulate(qmap, qclaim, counts, grid, rng, clock):
    """One persistent launch: `grid` workgroups (workgroup w starting at
    range w % 8) make claim attempts in random order.  Returns
    {(job, tile): time}, the final counters, the number of workgroups that
    left the loop and the prefix."""
    arr, p = prefix_of(counts)
    total = p[-1]
    counters = (ctypes.c_uint64 * RANGES)()
    state = {w: [w % RANGES, 0] for w in range(grid)}
    live = list(range(grid))
    taken = {}
    left = 0
    while live:
        w = rng.choice(live)
        item = qclaim(counters, total, state[w])
        clock[0] += 1
        if item == -2:
            live.remove(w)
            left += 1
            continue
        if item == -1:
            continue
        assert 0 <= item < total
        m = qmap(arr, len(counts), item)
        assert m is not None
        assert m not in taken, ("claimed twice", m)
        taken[m] = clock[0]
    return taken, list(counters), left, p


def range_sizes(total):
    per = (total + RANGES - 1) // RANGES
    return [max(0, min((x + 1) * per, total) - x * per) for x in range(RANGES)]


@pytest.mark.parametrize("seed", range(8))
def test_every_tile_exactly_once(qmap, qclaim, seed):
    rng = random.Random(seed)
    njobs = rng.randint(1, 40)
    counts = [rng.choice([0, 0, 1, 2, rng.randint(1, 60)]) for _ in range(njobs)]
    grid = rng.choice([1, 3, 17, 64, 300])
    taken, counters, left, p = simulate(qmap, qclaim, counts, grid, rng, [0])
    want = {(j, t) for j, c in enumerate(counts) for t in range(c)}
    assert set(taken) == want
    assert counters == [n + grid for n in range_sizes(p[-1])] and left == grid
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


def test_stages_in_launch_order(qmap, qclaim):
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
        taken, counters, left, p = simulate(qmap, qclaim, counts, rng.choice([4, 16, 64]), rng, clock)
        assert left > 0 and counters == [n + left for n in range_sizes(p[-1])]
        first, last = {}, {}
        for (j, t), tm in taken.items():
            first[j] = min(first.get(j, tm), tm)
            last[j] = max(last.get(j, tm), tm)
        times.append((first, last))
    for s in range(3):
        for j in range(products):
            if j in times[s][1] and j in times[s + 1][0]:
                assert times[s][1][j] < times[s + 1][0][j]


def test_claims_start_in_own_range(qclaim):
    """Workgroups claim neighbouring items of their own XCD's range before
    moving on, and a drained range hands out nothing more."""
    total = 800
    counters = (ctypes.c_uint64 * RANGES)()
    per = total // RANGES
    for x in range(RANGES):
        st = [x, 0]
        assert [qclaim(counters, total, st) for _ in range(3)] == [x * per, x * per + 1, x * per + 2]
    st = [3, 0]
    got = [qclaim(counters, total, st) for _ in range(per - 3)]
    assert got == list(range(3 * per + 3, 4 * per))
    assert qclaim(counters, total, st) == -1 and st == [4, 1]   # range 3 drained: on to range 4
    assert qclaim(counters, total, st) == 4 * per + 3
    done = [7, RANGES]
    assert qclaim(counters, total, done) == -2
    empty = (ctypes.c_uint64 * RANGES)()
    st = [5, 0]
    assert [qclaim(empty, 0, st) for _ in range(RANGES + 1)] == [-1] * RANGES + [-2]
