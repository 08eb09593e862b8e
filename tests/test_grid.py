"""The work grid on the CPU: the product's native grid (x2_grid_* in the device library) and the oracle's restatement
(oracle/oracle_grid.py) held to the reference's own scheduler invariants (xpu2/interactive/async_grid_test.cc:19-190:
every (pixel, sample) claimed exactly once under concurrent workers of both devices, appends while claiming, the
checkerboard and the cellular rules, the progressive ranges and the frontier, important cells and the rights of
passage, the ownership flips), and against each other call for call: driven by the same sequence of calls on one
thread, both pick the same cells and hand out the same chunks. The reference's grid includes async_integrator.h,
which cannot be built here (NanoVDB is absent), so these invariants and the restatement are what pin it."""
import threading
import time

import numpy as np
import pytest

import oracle_grid as O
from x2amd import grid as G

CPU, GPU, NO = G.CPU, G.GPU, G.NO_CELL


def tile_pixels(tile, w, h):
    tx = (w + 7) // 8
    x0, y0 = (tile % tx) * 8, (tile // tx) * 8
    return [y * w + x for y in range(y0, min(y0 + 8, h)) for x in range(x0, min(x0 + 8, w))]


def exactly_once(w, h, spp, opts, workers):
    """async_grid_test.cc:19-61: CPU workers claim one chunk at a time on their own threads while the GPU's eight
    cellular workers claim 97 at a time on this one."""
    g = G.WorkGrid(w, h, spp, opts)
    claimed = [[] for _ in range(workers + 1)]

    def cpu_worker(k):
        cell = g.seed(CPU, k, workers)
        buf = (G.abi.Chunk * 1)()
        while True:
            cell = g.choose(CPU, cell, k)
            if cell == NO:
                break
            _, got = g.claim(cell, 1, CPU, buf)
            if got:
                c = buf[0]
                claimed[k].append((c.tile, c.first, c.count))

    threads = [threading.Thread(target=cpu_worker, args=(k,)) for k in range(workers)]
    for t in threads:
        t.start()
    buf = (G.abi.Chunk * 97)()
    cells = [g.seed(GPU, k, 8) for k in range(8)]
    while not g.empty():
        for k in range(8):
            cell = g.choose(GPU, cells[k], k)
            if cell == NO:
                continue
            cells[k] = cell
            _, got = g.claim(cell, 97, GPU, buf)
            claimed[workers].extend((buf[i].tile, buf[i].first, buf[i].count) for i in range(got))
    for t in threads:
        t.join()
    samples = np.zeros(w * h, np.int64)
    seen = set()
    n = 0
    for lst in claimed:
        for tile, first, count in lst:
            n += 1
            assert (tile, first) not in seen, "no chunk is claimed twice"
            seen.add((tile, first))
            samples[tile_pixels(tile, w, h)] += count
    assert n == g.total_chunks(), "every chunk is claimed"
    assert (samples == spp).all(), "every pixel gets exactly spp samples"
    return claimed


@pytest.mark.parametrize("w,h,spp,opts,workers", [
    (384, 288, 64, G.options(), 12),
    (1920, 1080, 16, G.options(64, 4, 2), 12),
    (100, 37, 9, G.options(32, 4, 2), 5),   # ragged edges, a partial last chunk
    (64, 64, 7, G.options(64, 3, 0), 3),    # CPU only
])
def test_exactly_once(w, h, spp, opts, workers):
    claimed = exactly_once(w, h, spp, opts, workers)
    if opts.mode == 2 and w * h >= 384 * 288:  # a small frame can be gone before the other device's first claim
        assert claimed[-1] and sum(len(c) for c in claimed[:-1]), "both devices claimed"


def test_append_while_claiming():
    """async_grid_test.cc:66-112: passes appended to an empty grid while 12 workers claim, past the queues' 4,096-chunk
    blocks; every appended chunk claimed exactly once."""
    w, h = 384, 288
    tiles_x, tiles = w // 8, (w // 8) * (h // 8)
    g = G.WorkGrid(w, h, 1024, G.options(64, 4, 0), empty=True)
    done = threading.Event()
    claimed = [[] for _ in range(12)]

    def worker(k):
        cell = g.seed(CPU, k, 12)
        buf = (G.abi.Chunk * 1)()
        while True:
            nxt = g.choose(CPU, cell, k)
            if nxt == NO:
                if done.is_set() and g.empty():
                    break
                time.sleep(0.0005)  # nothing appended yet: leave the interpreter to the coordinator
                continue
            cell = nxt
            _, got = g.claim(cell, 1, CPU, buf)
            if got:
                claimed[k].append((buf[0].tile, buf[0].first))

    threads = [threading.Thread(target=worker, args=(k,)) for k in range(12)]
    for t in threads:
        t.start()
    appended = 0
    for p in range(40):
        per_cell = {}
        for tile in range((p * 7) % 3, tiles, 1 + p % 3):
            for off in range(0, 32, 4):
                c = g.cell_of(tile)
                per_cell.setdefault(c, []).append((tile, (p << 16) | off, 4, c))
        for c, lst in per_cell.items():
            g.append(c, lst)
            appended += len(lst)
    done.set()
    for t in threads:
        t.join()
    flat = [c for lst in claimed for c in lst]
    assert len(set(flat)) == len(flat), "no appended chunk is claimed twice"
    assert len(flat) == appended == g.total_chunks(), "every appended chunk is claimed"
    assert tiles_x == 48


def rules(make, claim, choose, owner, remaining, cell_count):
    """async_grid_test.cc:120-145 on either grid: checkerboard seeding, finish the cell, then a same-owner neighbour,
    then steal; the GPU takes its own cells first."""
    g = make()
    assert owner(g, 0) == GPU and owner(g, 1) == CPU and owner(g, 4) == CPU and owner(g, 5) == GPU
    start = 1
    assert choose(g, CPU, start, 0) == start
    claim(g, start, remaining(g, start), CPU)
    nxt = choose(g, CPU, start, 0)
    assert nxt != start and owner(g, nxt) == CPU
    left = True
    while left:
        left = False
        for i in range(cell_count(g)):
            if owner(g, i) == CPU and remaining(g, i):
                claim(g, i, remaining(g, i), CPU)
                left = True
    stolen = choose(g, CPU, start, 0)
    assert stolen != NO and owner(g, stolen) == GPU
    assert owner(g, choose(g, GPU, NO, 0)) == GPU


def test_cellular_rules_product():
    rules(lambda: G.WorkGrid(256, 256, 4, G.options(64, 4, 2)), lambda g, c, n, d: g.claim_list(c, n, d),
          lambda g, d, c, s: g.choose(d, c, s), lambda g, c: g.owner(c), lambda g, c: g.remaining(c),
          lambda g: g.cell_count())


def test_cellular_rules_oracle():
    rules(lambda: O.Grid(256, 256, 4, 64, 4, 2), lambda g, c, n, d: g.claim(c, n, d),
          lambda g, d, c, s: g.choose(d, c, s), lambda g, c: g.owner[c], lambda g, c: g.remaining(c),
          lambda g: g.cell_count())


def test_progressive_ranges_and_frontier():
    """async_grid_test.cc:147-158, on both: ranges 1, 1, 2, 4; a cell hands out one range up to the frontier, which
    waits for the other cell."""
    for g, claim, claimable, ranges in [
        (G.WorkGrid(128, 64, 16, G.options(64, 4, 0)), lambda g, c, n: len(g.claim_list(c, n, CPU)),
         lambda g, c: g.claimable(c, CPU), lambda g: g.ranges()),
        (O.Grid(128, 64, 16, 64, 4, 0), lambda g, c, n: len(g.claim(c, n, CPU)), lambda g, c: g.claimable(c, CPU),
         lambda g: g.ranges),
    ]:
        r = ranges(g)
        assert len(r) == 6 and r[2] == (2, 2) and r[3][1] == 4
        assert claim(g, 0, 1000) == 64
        assert claimable(g, 0) == 0 and claimable(g, 1) == 64
        assert claim(g, 1, 1000) == 64 and g.frontier(CPU) == 1 and claimable(g, 0) == 64


def test_important_cells_and_passage():
    """async_grid_test.cc:159-172, on both: above the mean is important, runs a range ahead and stays its owner's; the
    rest are anyone's; a right of passage only where the CPU would not hold the frame up."""
    gp = G.WorkGrid(256, 64, 16, G.options(64, 4, 2))
    go = O.Grid(256, 64, 16, 64, 4, 2)
    for g, important, reserved in [(gp, gp.important, gp.reserved),
                                   (go, lambda c: go.important_[c], go.reserved)]:
        g.set_scores([1.0, 0.0, 0.0, 0.0])
        assert important(0) and not important(2) and reserved(0) and not reserved(2)
        assert g.claimable(0, GPU) == 128 and g.claimable(2, GPU) == 64
        assert g.choose(CPU, 1, 0) == 1
    gp.set_rates(1e9, 1e3, 1, 0.01)
    gp.gpu_claimed()
    go.set_rates(1e9, 1e3, 1, 0.01)
    assert not gp.passage(CPU, 0) and not go.passage(CPU, 0, since_gpu_claim=0.0)
    gp.set_rates(1e3, 1e9, 1, 0.01)
    go.set_rates(1e3, 1e9, 1, 0.01)
    assert gp.passage(CPU, 0) and go.passage(CPU, 0, since_gpu_claim=0.0)
    # A GPU that has not claimed for a while lets the CPU in (a latency gap), in the oracle with the clock given.
    go.set_rates(1e9, 1e3, 1, 0.01)
    assert go.passage(CPU, 0, since_gpu_claim=0.06)


def test_ownership_flips():
    """async_grid_test.cc:174-186, on both: the GPU renders everything for 12 frames; cells flip to it, at most 8 a
    frame; product and oracle agree on every owner."""
    own, oown = G.Ownership(), O.Ownership()
    for frame in range(12):
        owners = own.owners(256)
        oo = oown.owners(256)
        assert (owners is None) == (oo is None) and (owners is None or list(owners) == oo)
        g = G.WorkGrid(1024, 1024, 4, G.options(64, 4, 2), owners=owners)
        o = O.Grid(1024, 1024, 4, 64, 4, 2, owners=oo)
        for i in range(256):
            g.rendered(i, GPU, 100)
            o.rendered(i, GPU, 100)
        own.finished(g)
        oown.finished(o)
    after = own.owners(256)
    cpu_cells = int((after == CPU).sum())
    assert 128 - 8 * 5 <= cpu_cells < 128
    assert list(after) == oown.owners(256)


def test_ownership_matches_oracle_under_mixed_shares():
    """Random per-cell shares over 120 frames (flips both ways, the cooldown and the freeze): the same owners."""
    rng = np.random.default_rng(3)
    own, oown = G.Ownership(), O.Ownership()
    for frame in range(120):
        owners, oo = own.owners(64), oown.owners(64)
        assert (owners is None) == (oo is None) and (owners is None or list(owners) == oo), frame
        g = G.WorkGrid(512, 512, 4, G.options(64, 4, 2), owners=owners)
        o = O.Grid(512, 512, 4, 64, 4, 2, owners=oo)
        bias = rng.random(64)
        for i in range(64):
            gpu = int(rng.integers(0, 1000) * bias[i])
            cpu = int(rng.integers(0, 1000) * (1 - bias[i]))
            for grid in (g, o):
                grid.rendered(i, GPU, gpu)
                grid.rendered(i, CPU, cpu)
        own.finished(g)
        oown.finished(o)


@pytest.mark.parametrize("seed", range(6))
def test_calls_match_oracle(seed):
    """The same random sequence of seeds, choices, claims, scores, suitabilities, rates and appends on one thread:
    the product and the oracle pick the same cells and hand out the same chunks, to the grid's last chunk."""
    rng = np.random.default_rng(seed)
    w, h = [(300, 200), (257, 129), (64, 64), (520, 130), (200, 300), (128, 520)][seed]
    spp = int(rng.integers(3, 40))
    cell, spc, mode = [32, 64, 16][seed % 3], int(rng.integers(1, 6)), [2, 2, 0, 1, 2, 2][seed]
    owners = rng.integers(0, 2, ((w + cell - 1) // cell) * ((h + cell - 1) // cell)).astype(np.uint8) if seed == 4 else None
    vp = [5, 3, w - 7, h - 2] if seed == 1 else None
    first, count = (3, 9) if seed == 5 else (0, 0)
    gp = G.WorkGrid(w, h, spp, G.options(cell, spc, mode, seed != 3, first, count), owners=owners, viewport=vp)
    go = O.Grid(w, h, spp, cell, spc, mode, seed != 3, first, count, owners=None if owners is None else list(owners),
                viewport=vp)
    n = gp.cell_count()
    assert n == go.cell_count() and gp.total_chunks() == go.total
    assert gp.ranges() == go.ranges
    assert [gp.owner(i) for i in range(n)] == go.owner
    workers = {d: [gp.seed(d, k, 6) for k in range(6)] for d in (CPU, GPU)}
    assert workers == {d: [go.seed(d, k, 6) for k in range(6)] for d in (CPU, GPU)}
    claimed = 0
    steps = 0
    while not gp.empty():
        steps += 1
        assert steps < 200000
        r = rng.random()
        if r < 0.01:
            scores = [float(v) if rng.random() > 0.1 else -1.0 for v in rng.gamma(1.0, 1.0, n)]
            gp.set_scores(scores)
            go.set_scores(scores)
            continue
        if r < 0.015:
            p = rng.random(n).astype(np.float32)
            gp.set_suitability(p)
            go.gpu_p = [float(v) for v in p]
            continue
        if r < 0.02:  # rates with the GPU never having claimed: passage depends on the queues alone
            rates = (float(rng.choice([0, 1e3, 1e6, 1e9])), float(rng.choice([0, 1e3, 1e6])), int(rng.integers(1, 16)),
                     float(rng.choice([0.001, 0.01, 0.1])))
            gp.set_rates(*rates)
            go.set_rates(*rates)
            continue
        d = int(rng.integers(0, 2))
        k = int(rng.integers(0, 6))
        a = gp.choose(d, workers[d][k], k)
        b = go.choose(d, workers[d][k], k)
        assert a == b, (steps, d, k)
        if a == NO:
            if all(gp.choose(e, NO, 0) == NO for e in (CPU, GPU)):
                gp.set_rates(0, 0, 1, 0.01)  # neither may claim: clear the rates (passage everywhere)
                go.set_rates(0, 0, 1, 0.01)
            continue
        workers[d][k] = a
        want = int(rng.integers(1, 40))
        x = gp.claim_list(a, want, d)
        y = go.claim(a, want, d)
        assert [c[:3] for c in x] == [c[:3] for c in y] and all(c[3] == a for c in x), steps
        claimed += len(x)
        for i in range(n):
            assert gp.claimable(i, d) == go.claimable(i, d)
        assert gp.frontier(CPU) == go.frontier(CPU) and gp.frontier(GPU) == go.frontier(GPU)
    assert claimed == gp.total_chunks()
    assert go.empty()


def test_arguments_fail_cleanly():
    lib = G.abi.load()
    with pytest.raises(RuntimeError):
        G.WorkGrid(64, 64, 4, G.options(60, 4, 2))  # a cell not a multiple of 8
    with pytest.raises(RuntimeError):
        G.WorkGrid(0, 64, 4)
    g = G.WorkGrid(64, 64, 4, G.options(32, 4, 2), empty=True)
    with pytest.raises(RuntimeError):
        g.append(0, [(4, 0, 4, 0)])  # tile 4 lies in cell 1
    assert lib.x2_grid_claim(None, 0, 1, None, 0) == 0 and lib.x2_grid_choose(None, 0, 0, 0) == NO
    assert g.choose(CPU, NO, 0) == NO and g.empty()
    g.append(1, [(4, 0, 4, 1)])
    assert g.choose(CPU, NO, 0) == 1 and g.claim_list(1, 5, CPU) == [(4, 0, 4, 1)]
    g.close()
    assert g.empty()
