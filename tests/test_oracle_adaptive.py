"""Adaptive frames on the CPU: the oracle's restatement (oracle/oracle_adaptive.py) pinned by what the mathematics fixes, and
the product's host trees and pixel error (x2_adaptive_masks, x2_pixel_errors in the device library, which load and run
without a GPU) bit for bit against it.

Pins: the schedule's sizes and coverage; the pixel error's closed forms (equal halves, the two-sample case against a
float64 formula written from the half-images, scaling by 4 doubling it exactly, NaN, no luma); the trees' structure
(splits partition their parent along one axis at MinNodeSize from either end, converged nodes are below the target
and shorter than MaxNodeSize, open leaves cannot converge or split), sampleArea against a brute-force coverage test
of the open leaves, and the regions' layout covering every tile. No reference test holds adaptive vectors, and the
reference's own adaptive headers include async_integrator.h, which cannot be built here (NanoVDB is absent), so no
golden vectors come from the reference for this path."""
import math

import numpy as np
import pytest

import oracle_adaptive as A
from x2amd import lane as L

F = np.float32


def test_schedule():
    assert A.schedule(1) == [(0, 1)]
    assert A.schedule(5) == [(0, 1), (1, 5)]
    assert A.schedule(64) == [(0, 1), (1, 5), (5, 11), (11, 19), (19, 29), (29, 41), (41, 55), (55, 64)]
    for n in [2, 7, 16, 100, 1024, 4096]:
        p = A.schedule(n)
        assert p[0] == (0, 1) and p[-1][1] == n
        assert all(a[1] == b[0] for a, b in zip(p, p[1:]))
        sizes = [e - s for s, e in p]
        assert all(sizes[i] == 4 + 2 * (i - 1) for i in range(1, len(sizes) - 1))
        assert 0 < sizes[-1] <= 4 + 2 * (len(sizes) - 2) + 2 or len(sizes) <= 2


def test_pixel_error_closed_forms():
    rng = np.random.default_rng(5)
    assert A.pixel_error([1, 1, 1, 1], [0, 0, 0, 0], 1) == F(-1)
    assert A.pixel_error([0, 0, 0, 0], [0, 0, 0, 0], 0) == F(-1)
    assert A.pixel_error([float("nan"), 0, 0, 1], [0, 0, 0, 1], 2) == F(0)
    for _ in range(200):
        a = rng.uniform(0.01, 4, 4)
        b = rng.uniform(0.01, 4, 4)
        # Two samples, index 0 even and index 1 odd: the halves' means are a and b.
        e = A.pixel_error(a + b, b, 2)
        d = np.abs(a - b)
        luma = lambda v: 0.299 * v[0] + 0.587 * v[1] + 0.114 * v[2]  # noqa: E731
        want = luma(d) / math.sqrt(luma((a + b) / 2)) + d[3]
        assert abs(float(e) - want) <= 1e-5 * want
        # Equal halves: no error.
        assert A.pixel_error(2 * a, a, 2) == F(0)
        # Four times the film (exact in every product): exactly twice the error, alpha equal.
        s = rng.uniform(0.1, 4, 4)
        o = rng.uniform(0.1, 2, 4)
        o[3] = s[3] / 2
        n = int(rng.integers(2, 64)) * 2
        e1 = A.pixel_error(s * n, o * n, n)
        assert A.pixel_error(s * n * 4, o * n * 4, n) == F(2) * e1
    # No luma: the alpha difference alone.
    assert A.pixel_error([0, 0, 0, 1.5], [0, 0, 0, 0.25], 2) == F(abs(1.5 - 0.25 * 2))


def random_film(rng, n, spp_max=16):
    counts = rng.integers(0, spp_max, n).astype(np.uint32)
    sums = rng.gamma(1.0, 1.0, (n, 4)) * counts[:, None]
    odd = sums * rng.uniform(0.2, 0.8, (n, 1))
    sums[rng.random(n) < 0.01, 0] = float("nan")
    return sums, odd, counts


def test_pixel_errors_product_matches_oracle():
    rng = np.random.default_rng(11)
    sums, odd, counts = random_film(rng, 3000)
    a = L.pixel_errors(sums, odd, counts)
    b = A.pixel_errors(sums, odd, counts)
    assert (a.view(np.uint32) == b.view(np.uint32)).all()


def error_image(rng, w, h, kind):
    if kind == "noise":
        return rng.gamma(0.6, 1e-3, w * h).astype(np.float32)
    if kind == "hot":
        e = np.full(w * h, 1e-6, np.float32)
        e[(h // 3) * w + w // 2] = 5.0
        return e
    y, x = np.divmod(np.arange(w * h), w)
    if kind == "ramp":  # error growing a thousandfold across the image: trees deep on one side, converged on the other
        return (rng.gamma(0.6, 1.0, w * h) * 10.0 ** (-3 + 3 * x / w)).astype(np.float32)
    if kind == "blob":
        blob = np.exp(-((x - w * 0.3) ** 2 + (y - h * 0.6) ** 2) / (0.02 * w * h))
        return (rng.gamma(0.6, 1.0, w * h) * (1e-3 + blob)).astype(np.float32)
    if kind == "missing":
        e = (rng.gamma(0.6, 1.0, w * h) * 10.0 ** (-3 + 3 * x / w)).astype(np.float32)
        e[rng.integers(0, w * h, 3)] = -1.0
        return e
    raise ValueError(kind)


def check_tree(tree, sat):
    """Structure of a built tree against the split and stopping rules."""
    nodes = tree.nodes
    root = tree.root
    for bounds, children, complete in nodes:
        w, h = F(bounds[2] - bounds[0]), F(bounds[3] - bounds[1])
        length = max(w, h)
        fate, _, _ = A.decide(sat, root, bounds, tree.target)
        if children >= 0:
            a, b = nodes[children][0], nodes[children + 1][0]
            horizontal = w > h
            k = 0 if horizontal else 1
            assert a[k] == bounds[k] and b[k + 2] == bounds[k + 2] and a[k + 2] == b[k]
            assert a[1 - k] == bounds[1 - k] == b[1 - k] and a[3 - k] == bounds[3 - k] == b[3 - k]
            cut = float(a[k + 2]) - float(bounds[k])
            assert cut >= float(A.MIN_NODE) * (1 - 1e-6) and float(bounds[k + 2]) - float(a[k + 2]) >= float(A.MIN_NODE) * (1 - 1e-6)
            assert complete == (nodes[children][2] and nodes[children + 1][2])
            assert fate == A.SPLIT
        elif complete:
            assert fate == A.CONVERGED and length < A.MAX_NODE
        else:
            assert fate == A.UNCONVERGED


def brute_sample_area(tree, tx, ty):
    """The pixels of the tile whose unit square overlaps an open leaf, filled to their bounding rectangle."""
    m = 0
    for bounds, children, complete in tree.nodes:
        if children >= 0 or complete:
            continue
        for q in range(64):
            x, y = tx + q % 8, ty + q // 8
            if x + 1 > bounds[0] and x < bounds[2] and y + 1 > bounds[1] and y < bounds[3]:
                m |= 1 << q
    return A.fill_gaps(m)


@pytest.mark.parametrize("w,h", [(64, 64), (37, 23), (100, 60), (15, 80)])
@pytest.mark.parametrize("kind,target", [("noise", 2e-6), ("hot", 2e-6), ("ramp", 1e-3), ("ramp", 3e-3), ("blob", 1e-3),
                                         ("missing", 1e-3)])
def test_trees_structure_and_product(w, h, kind, target):
    rng = np.random.default_rng(w * 1000 + h)
    errors = error_image(rng, w, h, kind)
    target = F(target)
    regions = A.Regions(w, h, target)
    regions.update(errors)
    for tree in regions.trees:
        if len(tree.nodes) > 1:
            check_tree(tree, A.Sat(errors, w, tree.root))
    tiles_x, tiles_y = (w + 7) // 8, (h + 7) // 8
    masks = regions.masks(w, h)
    for t in range(tiles_x * tiles_y):
        tx, ty = (t % tiles_x) * 8, (t // tiles_x) * 8
        r = min(ty // regions.rh, regions.per - 1) * regions.per + min(tx // regions.rw, regions.per - 1)
        tree = regions.trees[r]
        x0, y0, x1, y1 = tree.root
        assert x0 <= tx < x1 and y0 <= ty < y1 and min(tx + 8, w) <= x1 and min(ty + 8, h) <= y1
        assert int(masks[t]) == brute_sample_area(tree, tx, ty)
        if kind == "missing":
            region = errors.reshape(h, w)[y0:y1, x0:x1]
            if (region < 0).any():
                assert len(tree.nodes) == 1
    product, done = L.adaptive_masks(w, h, float(target), errors)
    assert (product == masks).all(), f"{kind} {w}x{h}"
    assert done == regions.done()


def test_trees_converge_and_stay_open():
    w, h = 64, 48
    quiet = np.zeros(w * h, np.float32)
    masks, done = L.adaptive_masks(w, h, 1e-3, quiet)
    assert done and not masks.any()
    loud = np.full(w * h, 10.0, np.float32)  # above target x SplitThreshold everywhere: nothing splits or converges
    masks, done = L.adaptive_masks(w, h, 1e-3, loud)
    assert not done and (masks == np.uint64(0xFFFFFFFFFFFFFFFF)).all()
    r = A.Regions(w, h, F(1e-3))
    r.update(loud)
    assert all(len(t.nodes) == 1 for t in r.trees)


@pytest.mark.parametrize("w,h", [(64, 64), (37, 23), (100, 60)])
@pytest.mark.parametrize("budget", [False, True])
def test_cellular_step_matches_oracle(w, h, budget):
    """One generation of the cellular rules: the product's host step equals the oracle's, and the rules hold."""
    rng = np.random.default_rng(w * 7 + h + (1000 if budget else 0))
    errors = error_image(rng, w, h, "missing")
    tiles_x, tiles_y = (w + 7) // 8, (h + 7) // 8
    cfg = type("C", (), {"width": w, "height": h, "viewport": [2, 1, w - 3, h]})
    inside = A.in_image(cfg)
    inside[rng.integers(0, len(inside), 2)] = 0  # a few tiles outside the render region
    alive = (rng.random(len(inside)) < 0.7).astype(np.uint8) & (inside != 0)
    planned = rng.integers(1, 40, w * h).astype(np.uint32) if budget else None
    left = float(rng.integers(1000, 60000)) if budget else 0.0
    target = 1e-3
    o = A.Cellular(w, h, target)
    o.dead_forever(inside)
    o.alive = [int(v) for v in alive]
    live = o.update(errors, (planned, left, 64, 0.0) if budget else None)
    nxt, ratios, plive = L.cellular_step(w, h, target, alive, inside, errors, planned, left, 64, 0.0)
    assert plive == live and list(nxt) == o.alive
    assert (ratios.view(np.uint32) == np.array(o.ratio, np.float32).view(np.uint32)).all()
    assert 0 < live < len(inside)
    for t in range(len(inside)):
        r = float(o.ratio[t])
        if not int(inside[t]):
            assert not nxt[t]
        elif r >= 1:
            assert nxt[t]
        elif r < 0.5:
            assert not nxt[t]
