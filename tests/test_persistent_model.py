"""CPU model of the persistent batch launch's protocol (pm_kernels.hip:
k_batch_persistent, grid_barrier) before it runs on a GPU.

The launch is G blocks, one per compute unit.  Every block runs the same
sequence: a grid-stride loop over the conversion items, barrier 0, a
block-stride loop over the tile items, barrier 1, a grid-stride loop over the
CRT groups.  A barrier adds 1 to its counter and waits (thread 0) until the
counter reaches G, the shared error word is set, or its bound of wall-clock
ticks / spins runs out -- then it sets the error word itself.  This model runs
that protocol as a discrete-time simulation with a limited number of resident
blocks (a block not yet resident starts when a resident one exits), and checks:
  * with every block resident, both barriers complete, no error is set and
    every item is processed exactly once;
  * with fewer resident blocks than G, the launch still ends in bounded time
    (no block waits forever) and the error word is set, so the host reports a
    timeout instead of hanging;
  * the grid-stride partitions cover each phase's items exactly once for the
    work-list sizes of the benchmark's batches.
"""
import pytest


def partition(count, blocks, per_block_lanes=1):
    """Items each (block, lane) of a grid-stride loop takes."""
    out = {}
    stride = blocks * per_block_lanes
    for b in range(blocks):
        for w in range(per_block_lanes):
            out[(b, w)] = list(range(b * per_block_lanes + w, count, stride))
    return out


def simulate(G, resident, work, limit):
    """work[phase][block] = time units of that block's items in the phase.
    Returns (end time, error set, barrier arrivals)."""
    counters = [0, 0]
    error = False
    state = {b: {"phase": 0, "left": work[0][b], "waiting": None, "waited": 0, "done": False} for b in range(G)}
    running = list(range(min(resident, G)))
    queue = list(range(min(resident, G), G))
    t = 0
    while running or queue:
        t += 1
        assert t < 10_000_000, "the model did not end"
        for b in list(running):
            s = state[b]
            if s["waiting"] is not None:  # thread 0 at barrier s["waiting"]
                k = s["waiting"]
                if counters[k] >= G or error:
                    s["waiting"] = None
                    s["phase"] += 1
                    s["left"] = work[s["phase"]][b]
                    continue
                s["waited"] += 1
                if s["waited"] > limit:
                    error = True
                    s["waiting"] = None
                    s["phase"] += 1
                    s["left"] = work[s["phase"]][b]
                continue
            if s["left"] > 0:
                s["left"] -= 1
                continue
            if s["phase"] < 2:  # arrive at the barrier after this phase
                counters[s["phase"]] += 1
                s["waiting"] = s["phase"]
                s["waited"] = 0
                continue
            s["done"] = True  # after the CRT phase the block exits
            running.remove(b)
            if queue:
                running.append(queue.pop(0))
    return t, error, counters


def phase_work(G, conv_items, tile_items, crt_items, waves=8):
    """Per-block time of each phase: the largest item count a lane of the
    block takes (grid-stride over waves for conversion and CRT, block-stride
    for tiles)."""
    conv = partition(conv_items * 16, G, waves)  # 16 slots per conversion item
    tile = partition(tile_items, G)
    crt = partition(crt_items * 16, G, waves)  # 16 groups of 64 per CRT item
    w = [[0] * G for _ in range(3)]
    for b in range(G):
        w[0][b] = max(len(conv[(b, x)]) for x in range(waves))
        w[1][b] = 4 * len(tile[(b, 0)])
        w[2][b] = max(len(crt[(b, x)]) for x in range(waves))
    return w


# 128 x N = 2^15 with two fused outer stages: 2 x 16 conversion items per
# product, 59 primes x 4 tiles of 2^13 two to an item, 32 CRT items per product
HEADLINE = (128 * 2 * 16, 128 * 59 * 4 // 2, 128 * 32)


@pytest.mark.parametrize("G", [8, 64, 256])
def test_all_resident_completes_without_error(G):
    conv, tiles, crt = [max(1, x * G // 256) for x in HEADLINE]
    work = phase_work(G, conv, tiles, crt)
    t, error, counters = simulate(G, G, work, limit=10 ** 6)
    assert not error and counters == [G, G]


@pytest.mark.parametrize("G,resident", [(16, 15), (64, 32), (8, 1)])
def test_missing_residency_times_out_instead_of_hanging(G, resident):
    work = phase_work(G, 64, 64, 64)
    t, error, counters = simulate(G, resident, work, limit=200)
    assert error, "a barrier that cannot complete must set the error word"
    assert counters[0] == G  # every block still arrived, in bounded time


@pytest.mark.parametrize("count,G,lanes", [(HEADLINE[0] * 16, 256, 8), (HEADLINE[1], 256, 1),
                                           (HEADLINE[2] * 16, 256, 8), (37, 256, 8), (1, 256, 1)])
def test_partitions_cover_each_item_once(count, G, lanes):
    seen = []
    for items in partition(count, G, lanes).values():
        seen += items
    assert sorted(seen) == list(range(count))
