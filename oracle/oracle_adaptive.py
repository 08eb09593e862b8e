"""The oracle for adaptive frames — test infrastructure only (imported by tests/ and nothing shipped).

A restatement in numpy float32 arithmetic, operation for operation, of MoonRay's adaptive sampling as the reference
gives it: the pass schedule (xpu2/interactive/async_adaptive.h:27-38), the pixel error (async_integrator.h:3102-3125),
the region error trees (async_adaptive.h:42-97 Tree, :99-145 Regions; async_adaptive_core.h:29-179 for the node
arithmetic, the split rule and sampleArea) and the broker's adaptive passes (async_render.h:1219-1440 generate:
passes below the minimum sample every pixel, later passes the pixels the trees leave open, each pixel from its own
count). The trees are built recursively, depth first, as the reference's host builds them.

Every float32 operation rounds on its own (numpy float32 scalars), as the reference's mulRn/addRn do; sums over
errors are in float64, as the reference's summed-area table is."""
import math

import numpy as np

F = np.float32
MIN_NODE = F(1.27201964951407)  # sqrt of the golden ratio (async_adaptive_core.h:29)
MAX_NODE = F(16.0)
SPLIT_THRESHOLD = F(256.0)
NO_SPLIT = F(-1.0)
IMAGE_AREA = F(F(F(815.0) - F(0.0)) * F(F(1943.0) - F(-24.0))) / F(4.0)  # async_adaptive_core.h:54
CONVERGED, SPLIT, UNCONVERGED = 1, 2, 0
NO_SLOT = 0xFFFFFFFF
PASS_SHIFT = 16


def schedule(max_samples):
    """MoonRay's batch passes: one sample, then 4, 6, 8, ... more, the rest last (async_adaptive.h:27-38)."""
    passes = [(0, 1)]
    current, remaining, size = 1, max_samples - 1, 4
    while remaining > size and len(passes) < 999:
        passes.append((current, current + size))
        remaining -= size
        current += size
        size += 2
    if remaining:
        passes.append((current, current + remaining))
    return passes


def pixel_error(total, odd, samples):
    """async_integrator.h:3109-3125."""
    n = F(samples)
    n_odd = F(samples // 2)
    n_even = F(n - n_odd)
    if n_odd <= 0 or n_even <= 0:
        return F(-1.0)
    diff = [F(0)] * 4
    avg = [F(0)] * 3
    for c in range(4):
        t, o = F(total[c]), F(odd[c])
        if t != t or o != o:
            return F(0.0)
        diff[c] = F(abs(F(F(F(t * n_odd) + F(-F(o * n))) / F(n_odd * n_even))))
        if c < 3:
            avg[c] = F(t / n)

    def luma(v):
        return F(F(F(F(0.299) * v[0]) + F(F(0.587) * v[1])) + F(F(0.114) * v[2]))

    lum_diff, lum_avg = luma(diff), luma(avg)
    if lum_avg <= 0:
        return diff[3]
    return F(F(lum_diff / F(np.sqrt(lum_avg))) + diff[3])


def pixel_errors(sums, odd, counts):
    return np.array([pixel_error(sums[i], odd[i], int(counts[i])) for i in range(len(counts))], np.float32)


def round_larger(b):
    return (int(math.floor(b[0])), int(math.floor(b[1])), int(math.ceil(b[2])), int(math.ceil(b[3])))


def area_error(error, area):
    return F(F(error) / F(np.sqrt(F(F(area) * IMAGE_AREA))))


def split_helper(location, length):
    if F(length * location) >= MIN_NODE:
        return location
    if length >= F(MIN_NODE * F(2)):
        return F(MIN_NODE / length)
    return NO_SPLIT


def split_location(location, length):
    if location <= F(0.5):
        return split_helper(location, length)
    h = split_helper(F(F(1.0) - location), length)
    return NO_SPLIT if h == NO_SPLIT else F(F(1.0) - h)


class Sat:
    """A region's summed-area table (float64), (w + 1) x (h + 1), the first row and column zero."""

    def __init__(self, errors, width, root):
        x0, y0, x1, y1 = root
        self.x0, self.y0 = x0, y0
        w, h = x1 - x0, y1 - y0
        self.t = np.zeros((h + 1, w + 1), np.float64)
        for y in range(h):
            row = 0.0
            for x in range(w):
                row += float(errors[(y0 + y) * width + x0 + x])
                self.t[y + 1, x + 1] = self.t[y, x + 1] + row

    def rect(self, ax, ay, bx, by):
        at = lambda x, y: self.t[y - self.y0, x - self.x0]  # noqa: E731
        return ((at(bx, by) - at(ax, by)) - at(bx, ay)) + at(ax, ay)


def decide(sat, root, b, target):
    """A node's fate and, when it splits, its halves (async_adaptive_core.h:145-164)."""
    w, h = F(b[2] - b[0]), F(b[3] - b[1])
    horizontal = w > h
    length = w if horizontal else h
    ib = round_larger(b)
    ib = (max(ib[0], root[0]), max(ib[1], root[1]), min(ib[2], root[2]), min(ib[3], root[3]))
    n = (ib[2] - ib[0]) if horizontal else (ib[3] - ib[1])

    def m(i):
        if horizontal:
            return F(sat.rect(ib[0], ib[1], ib[0] + i + 1, ib[3]))
        return F(sat.rect(ib[0], ib[1], ib[2], ib[1] + i + 1))

    total = m(n - 1) if n > 0 else F(0)
    error = area_error(total, F(w * h))
    if error < target and length < MAX_NODE:
        return CONVERGED, None, None
    if error < F(target * SPLIT_THRESHOLD) and length > F(MIN_NODE * F(2)):
        # midpoint (async_adaptive_core.h:110-128)
        half = F(total * F(0.5))
        i = n // 2
        if total != 0:
            lo, hi = 0, n
            while lo < hi:
                mid = (lo + hi) // 2
                if m(mid) > half:
                    hi = mid
                else:
                    lo = mid + 1
            i = lo
        i = min(max(i, 1), n - 1)
        lower, higher = m(i - 1), m(i)
        delta = F(higher - lower)
        offset = F(F(half - lower) / delta) if delta > 0 else F(0.5)
        location = F(F(F(i) + offset) / F(n))
        offset = split_location(location, length)
        if offset == NO_SPLIT:
            return UNCONVERGED, None, None
        left, right = list(b), list(b)
        if horizontal:
            left[2] = right[0] = F(b[0] + F(offset * w))
        else:
            left[3] = right[1] = F(b[1] + F(offset * h))
        return SPLIT, tuple(left), tuple(right)
    return UNCONVERGED, None, None


class Tree:
    """One region's tree, rebuilt recursively from the errors (async_adaptive.h:42-97)."""

    def __init__(self, root, target):
        self.root = root
        self.target = F(target)
        self.nodes = [[tuple(F(v) for v in root), -1, 0]]

    def update(self, errors, width):
        x0, y0, x1, y1 = self.root
        self.nodes = [[tuple(F(v) for v in self.root), -1, 0]]
        region = np.array([[errors[y * width + x] for x in range(x0, x1)] for y in range(y0, y1)])
        if (region < 0).any():  # not every pixel has both halves: sample everything
            return
        sat = Sat(errors, width, self.root)

        def visit(k):
            fate, left, right = decide(sat, self.root, self.nodes[k][0], self.target)
            if fate == CONVERGED:
                self.nodes[k][2] = 1
            if fate != SPLIT:
                return
            first = len(self.nodes)
            self.nodes.append([left, -1, 0])
            self.nodes.append([right, -1, 0])
            self.nodes[k][1] = first
            visit(first)
            visit(first + 1)
            if self.nodes[first][2] and self.nodes[first + 1][2]:
                self.nodes[k][2] = 1

        visit(0)

    def done(self):
        return self.nodes[0][2] != 0

    def sample_area(self, tx, ty):
        """The tile's open pixels, filled to their bounding rectangle (async_adaptive_core.h:135-179)."""
        tile = (F(tx), F(ty), F(tx + 8), F(ty + 8))
        stack, mask = [0], 0
        while stack:
            bounds, children, complete = self.nodes[stack.pop()]
            if complete:
                continue
            i = (max(bounds[0], tile[0]), max(bounds[1], tile[1]), min(bounds[2], tile[2]), min(bounds[3], tile[3]))
            if i[2] <= i[0] or i[3] <= i[1]:
                continue
            if children >= 0:
                stack += [children, children + 1]
                continue
            r = round_larger(i)
            for y in range(max(r[1], ty), min(r[3], ty + 8)):
                for x in range(max(r[0], tx), min(r[2], tx + 8)):
                    mask |= 1 << ((y - ty) * 8 + (x - tx))
        return fill_gaps(mask)


def fill_gaps(m):
    if not m:
        return 0
    bits = [b for b in range(64) if m >> b & 1]
    rows, cols = [b // 8 for b in bits], [b % 8 for b in bits]
    out = 0
    for y in range(min(rows), max(rows) + 1):
        for x in range(min(cols), max(cols) + 1):
            out |= 1 << (y * 8 + x)
    return out


class Regions:
    """Up to 4 x 4 overlapping regions of whole tiles widened by 8 pixels (async_adaptive.h:99-145)."""

    def __init__(self, width, height, target):
        tile, overlap, desired = 8, 8, 4
        self.width = width
        if tile + overlap >= width or tile + overlap >= height:
            self.per, self.rw, self.rh = 1, width, height
        else:
            self.per = max(1, min(desired, min(width // (tile + overlap), height // (tile + overlap))))
            up = lambda n: n if n % tile == 0 else n + tile - n % tile  # noqa: E731
            self.rw = up(1 + (width - 1) // self.per)
            self.rh = up(1 + (height - 1) // self.per)
        self.trees = []
        for i in range(self.per * self.per):
            rx, ry = i % self.per, i // self.per
            b = (rx * self.rw - overlap, ry * self.rh - overlap, (rx + 1) * self.rw + overlap, (ry + 1) * self.rh + overlap)
            b = (max(b[0], 0), max(b[1], 0), min(b[2], width), min(b[3], height))
            self.trees.append(Tree(b, target))

    def update(self, errors):
        for t in self.trees:
            t.update(errors, self.width)

    def done(self):
        return all(t.done() for t in self.trees)

    def sample_area(self, tx, ty):
        r = min(ty // self.rh, self.per - 1) * self.per + min(tx // self.rw, self.per - 1)
        return self.trees[r].sample_area(tx, ty)

    def masks(self, width, height):
        tiles_x, tiles_y = (width + 7) // 8, (height + 7) // 8
        return np.array([self.sample_area((t % tiles_x) * 8, (t // tiles_x) * 8) for t in range(tiles_x * tiles_y)],
                        np.uint64)


def in_image(config):
    w, h, vp = config.width, config.height, config.viewport
    tiles_x, tiles_y = (w + 7) // 8, (h + 7) // 8
    out = np.zeros(tiles_x * tiles_y, np.uint64)
    for t in range(tiles_x * tiles_y):
        m = 0
        for q in range(64):
            x, y = (t % tiles_x) * 8 + q % 8, (t // tiles_x) * 8 + q // 8
            if x < w and y < h and vp[0] <= x < vp[2] and vp[1] <= y < vp[3]:
                m |= 1 << q
        out[t] = m
    return out


def adaptive_frame(oracle_scene, config, min_samples, max_samples, target, threads=None):
    """The broker's adaptive frame on the oracle's integrator: (sums, odd, counts, per-pass masks)."""
    w, h = config.width, config.height
    tiles_x = (w + 7) // 8
    inside = in_image(config)
    regions = Regions(w, h, target)
    sums = np.zeros((w * h, 4), np.float64)
    odd = np.zeros((w * h, 4), np.float64)
    counts = np.zeros(w * h, np.uint32)
    history = []
    for start, end in schedule(max_samples):
        uniform = start < min_samples
        size = end - start
        if not uniform:
            regions.update(pixel_errors(sums, odd, counts))
        masks = inside.copy()
        if not uniform:
            masks &= regions.masks(w, h)
        first = np.zeros(w * h, np.uint32)
        count = np.zeros(w * h, np.uint32)
        for t in np.nonzero(masks)[0]:
            m = int(masks[t])
            for q in range(64):
                if m >> q & 1:
                    p = ((t // tiles_x) * 8 + q // 8) * w + (t % tiles_x) * 8 + q % 8
                    first[p] = start if uniform else counts[p]
                    count[p] = size
        if not count.any():
            break
        history.append(masks)
        oracle_scene.render_pixels(config, first, count, sums, odd, counts, threads)
    return sums, odd, counts, history


class Cellular:
    """Cellular generations of the image's tiles (async_adaptive.h:147-317): a tile lives on its own error (ratio >= 1)
    or, down to SURVIVE with 3 live neighbours of 8, and comes back from BIRTH with 5. With a budget the target is the
    one the samples left would bring every tile to (budgetTarget, bisected in log space, float64 throughout)."""
    SURVIVE, BIRTH = F(0.5), F(0.75)

    def __init__(self, width, height, target):
        self.w, self.h = width, height
        self.tx, self.ty = (width + 7) // 8, (height + 7) // 8
        self.target = F(target)
        n = self.tx * self.ty
        self.alive = [1] * n
        self.dead = []
        self.ratio = [F(0)] * n
        self.error = [F(0)] * n
        self.pixels = [0.0] * n
        self.samples = [0.0] * n

    def dead_forever(self, in_image):
        self.dead = [0] * len(self.alive)
        for t in range(len(self.alive)):
            if not int(in_image[t]):
                self.dead[t] = 1
                self.alive[t] = 0

    def budget_target(self, planned, remaining, most):
        lo, hi = 3.4e38, 0.0
        live = [t for t in range(len(self.error)) if not (self.dead and self.dead[t]) and self.error[t] > 0 and self.samples[t] > 0]
        for t in live:
            lo, hi = min(lo, float(self.error[t])), max(hi, float(self.error[t]))
        if remaining <= 0 or hi <= 0:
            return F(3.4e38)

        def cost(target):
            c = 0.0
            for t in live:
                n = self.samples[t]
                r = float(self.error[t]) / target
                c += self.pixels[t] * (min(max(n * r * r, n), float(most)) - min(n, float(most)))
            return c

        low, high = lo * 1e-3, hi
        if cost(low) <= remaining:
            return F(low)
        for _ in range(48):
            middle = math.sqrt(low * high)
            if cost(middle) > remaining:
                low = middle
            else:
                high = middle
        return F(high)

    def update(self, errors, budget=None):
        """budget: (planned per pixel, samples left, most, floor) or None. Returns the live tiles."""
        for ty in range(self.ty):
            for tx in range(self.tx):
                t = ty * self.tx + tx
                total, most, rendered, n, unknown = 0.0, 0.0, 0.0, 0, False
                for y in range(ty * 8, min(self.h, ty * 8 + 8)):
                    for x in range(tx * 8, min(self.w, tx * 8 + 8)):
                        i = y * self.w + x
                        e = errors[i]
                        if e < 0:
                            unknown = True
                        else:
                            total += float(e)
                        n += 1
                        if budget is not None:
                            s = float(budget[0][i])
                            most = max(most, s)
                            rendered += 1 if s > 0 else 0
                self.error[t] = F(-1) if unknown else area_error(F(total), F(n))
                self.pixels[t] = rendered
                self.samples[t] = most
        target = self.target
        if budget is not None:
            target = max(self.budget_target(budget[0], budget[1], budget[2]), F(budget[3]))
        self.ratio = [F(3.4e38) if e < 0 else F(e / target) for e in self.error]
        nxt, live = [0] * len(self.alive), 0
        for ty in range(self.ty):
            for tx in range(self.tx):
                t = ty * self.tx + tx
                k = sum(1 for dy in (-1, 0, 1) for dx in (-1, 0, 1)
                        if (dx or dy) and 0 <= tx + dx < self.tx and 0 <= ty + dy < self.ty
                        and self.alive[(ty + dy) * self.tx + tx + dx])
                r = self.ratio[t]
                if self.alive[t]:
                    stays = r >= 1 or (r >= self.SURVIVE and k >= 3)
                else:
                    stays = r >= 1 or (r >= self.BIRTH and k >= 5)
                if self.dead and self.dead[t]:
                    stays = False
                nxt[t] = 1 if stays else 0
                live += nxt[t]
        self.alive = nxt
        return live


def cellular_frame(oracle_scene, config, min_samples, max_samples, target, sample_budget=0, noise_floor=0.0,
                   threads=None):
    """A frame by cellular generations (async_render.h:1362-1440, 1561-1563), optionally by a sample budget: (sums,
    odd, counts, per-pass masks). Ties among the tiles a budget keeps are broken by tile index."""
    w, h = config.width, config.height
    tiles_x = (w + 7) // 8
    inside = in_image(config)
    cells = Cellular(w, h, target)
    cells.dead_forever(inside)
    sums = np.zeros((w * h, 4), np.float64)
    odd = np.zeros((w * h, 4), np.float64)
    counts = np.zeros(w * h, np.uint32)
    planned = np.zeros(w * h, np.uint32)
    spent, history = 0.0, []
    for start, end in schedule(max_samples):
        uniform = start < min_samples
        size = end - start
        if not uniform:
            if sample_budget and spent >= sample_budget:
                break
            budget = (planned, max(0.0, sample_budget - spent), max_samples, noise_floor) if sample_budget else None
            if cells.update(pixel_errors(sums, odd, counts), budget) == 0:
                break
        dropped = None
        if not uniform and sample_budget:
            live = [(cells.ratio[t], t) for t in range(len(inside)) if int(inside[t]) and cells.alive[t]]
            wanted = sum(bin(int(inside[t])).count("1") * size for _, t in live)
            left = max(0.0, sample_budget - spent)
            if wanted > left:
                live.sort(key=lambda x: -float(x[0]))  # stable: ties keep tile order
                dropped = [1] * len(inside)
                taken = 0.0
                for _, t in live:
                    cost = bin(int(inside[t])).count("1") * size
                    if taken + cost > left:
                        break
                    taken += cost
                    dropped[t] = 0
        masks = inside.copy()
        if not uniform:
            for t in range(len(masks)):
                if not (cells.alive[t] and (dropped is None or not dropped[t])):
                    masks[t] = 0
        first = np.zeros(w * h, np.uint32)
        count = np.zeros(w * h, np.uint32)
        for t in np.nonzero(masks)[0]:
            m = int(masks[t])
            for q in range(64):
                if m >> q & 1:
                    p = ((t // tiles_x) * 8 + q // 8) * w + (t % tiles_x) * 8 + q % 8
                    first[p] = counts[p]
                    count[p] = size
        if not count.any():
            break
        planned += count
        spent += float(count.sum())
        history.append(masks)
        oracle_scene.render_pixels(config, first, count, sums, odd, counts, threads)
    return sums, odd, counts, history
