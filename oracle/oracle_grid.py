"""The oracle for the work grid — test infrastructure only (imported by tests/ and nothing shipped).

A single-threaded restatement in plain Python of the reference's work grid and cellular rules
(xpu2/interactive/async_grid.h): the progressive sample ranges and each cell's chunks in range-then-tile order
(:131-177), the claim limit and the devices' frontiers (:372-395), claims and what a device may claim (:231-246), the
important cells (:250-270), the rights of passage (:279-296, with the clock made an argument), the worker's choice of
cell (:325-359) and starting cell (:362-368), and the owners across frames (:402-437). One deliberate difference
from the reference, shared with the product: a range never holds more sample indices than a chunk
(min(size, samples_per_chunk)), so a grid with samples_per_chunk 1 hands out ranges of 1.

Because a worker's choices depend only on the queues' cursors, the product's grid driven by the same sequence of
calls on one thread must make the same choices and hand out the same chunks."""

TILE = 8
TILE_PIXELS = 64
CPU, GPU = 0, 1
NO_CELL = 0xFFFFFFFF


def spread(a, b):
    """The tie-break mix of (cell, salt) (async_grid.h:424-430), in 32-bit arithmetic."""
    m = 0xFFFFFFFF
    h = ((a * 0x9E3779B1) & m) ^ ((b + 0x7F4A7C15 + ((a << 6) & m) + (a >> 2)) & m)
    h ^= h >> 16
    h = (h * 0x85EBCA6B) & m
    h ^= h >> 13
    return h


def f32(x):
    import numpy as np
    return float(np.float32(x))


class Grid:
    def __init__(self, width, height, spp, cell_size=64, samples_per_chunk=4, mode=2, progressive=True,
                 sample_first=0, sample_count=0, owners=None, empty=False, viewport=None):
        assert width and height and spp and cell_size and cell_size % TILE == 0 and samples_per_chunk
        self.w, self.h, self.spp = width, height, spp
        self.cell_size, self.spc, self.mode = cell_size, samples_per_chunk, mode
        self.viewport = list(viewport) if viewport is not None else [0, 0, 0x7FFFFFFF, 0x7FFFFFFF]
        self.cx = (width + cell_size - 1) // cell_size
        self.cy = (height + cell_size - 1) // cell_size
        self.tx = (width + TILE - 1) // TILE
        self.ty = (height + TILE - 1) // TILE
        per = cell_size // TILE
        n = self.cx * self.cy
        if owners is not None and len(owners) != n:
            owners = None
        self.ranges = []
        if not empty:
            lo = sample_first if sample_count else 0
            hi = min(spp, sample_first + sample_count) if sample_count else spp
            first = lo
            while first < hi:
                done = first - lo
                size = samples_per_chunk if not progressive else min(samples_per_chunk, 1 if done < 2 else 2 if done < 4
                                                                     else samples_per_chunk)
                self.ranges.append((first, min(size, hi - first)))
                first += self.ranges[-1][1]
        self.owner, self.queue, self.taken, self.tiles = [], [], [], []
        self.gpu_p = [0.5] * n
        self.score = [0.0] * n
        self.important_ = [False] * n
        self.samples = [[0, 0] for _ in range(n)]
        self.total = 0
        for cy in range(self.cy):
            for cx in range(self.cx):
                index = cy * self.cx + cx
                if mode == 0:
                    o = CPU
                elif mode == 1:
                    o = GPU
                elif owners is not None:
                    o = GPU if owners[index] else CPU
                else:
                    o = GPU if (cx + cy) % 2 == 0 else CPU
                self.owner.append(o)
                chunks = []
                if not empty:
                    for first, count in self.ranges:
                        for ty in range(cy * per, min(self.ty, (cy + 1) * per)):
                            for tx in range(cx * per, min(self.tx, (cx + 1) * per)):
                                t = ty * self.tx + tx
                                if self.tile_pixels(t):
                                    chunks.append((t, first, count, index))
                self.tiles.append(len(chunks) // len(self.ranges) if self.ranges else 0)
                self.queue.append(chunks)
                self.taken.append(0)
                self.total += len(chunks)
        self.closed = False
        self.scored = False
        self.rates = (0.0, 0.0, 1, 0.01)
        self.frontier_ = [0, 0]
        self.refresh()

    def cell_count(self):
        return self.cx * self.cy

    def tile_pixels(self, t):
        x, y = (t % self.tx) * TILE, (t // self.tx) * TILE
        vp = self.viewport
        x0, y0 = max(x, vp[0]), max(y, vp[1])
        x1, y1 = min(x + TILE, self.w, vp[2]), min(y + TILE, self.h, vp[3])
        return (x1 - x0) * (y1 - y0) if x1 > x0 and y1 > y0 else 0

    def cell_of(self, t):
        per = self.cell_size // TILE
        return ((t // self.tx) // per) * self.cx + (t % self.tx) // per

    def remaining(self, i):
        return 0 if self.closed else max(0, len(self.queue[i]) - self.taken[i])

    def empty(self):
        return all(self.remaining(i) == 0 for i in range(self.cell_count()))

    def append(self, i, chunks):
        self.queue[i].extend(chunks)
        self.total += len(chunks)

    def reserved(self, i):
        return not self.scored or self.important_[i]

    def limit(self, i, device):
        if not self.ranges:
            return 0xFFFFFFFF
        lead = 1 if self.scored and self.important_[i] else 0
        up_to = min(self.frontier_[device] + lead + 1, len(self.ranges))
        return min(up_to * self.tiles[i], 0xFFFFFFFF)

    def refresh(self):
        if not self.ranges:
            return
        least = [None, None, None]
        for i in range(self.cell_count()):
            if not self.tiles[i] or len(self.queue[i]) - self.taken[i] <= 0:
                continue
            at = self.taken[i] // self.tiles[i]
            open_ = not self.reserved(i)
            for dev in (CPU, GPU):
                if open_ or self.owner[i] == dev:
                    least[dev] = at if least[dev] is None else min(least[dev], at)
            least[2] = at if least[2] is None else min(least[2], at)
        for dev in (CPU, GPU):
            v = least[dev] if least[dev] is not None else least[2]
            self.frontier_[dev] = min(v if v is not None else 0xFFFFFFFF, len(self.ranges))

    def frontier(self, device):
        return self.frontier_[device]

    def claimable(self, i, device):
        if self.closed:
            return 0
        end = min(len(self.queue[i]), self.limit(i, device))
        return max(0, end - self.taken[i])

    def claim(self, i, n, device):
        if self.closed:
            return []
        end = min(len(self.queue[i]), self.limit(i, device))
        start = self.taken[i]
        if start >= end:
            return []
        stop = min(end, start + n)
        self.taken[i] = stop
        got = self.queue[i][start:stop]
        if got and self.ranges:
            self.refresh()
        return got

    def set_scores(self, scores):
        if len(scores) != self.cell_count():
            return
        known = [s for s in scores if s >= 0]
        if not known:
            return
        mean = f32(sum(float(s) for s in known) / len(known))
        for i, s in enumerate(scores):
            self.score[i] = f32(max(f32(s), 0.0))
            self.important_[i] = f32(s) > mean
        self.scored = True
        self.refresh()

    def set_rates(self, gpu, cpu_worker, workers, window):
        self.rates = (gpu, cpu_worker, workers, window)

    def reserved_chunks(self, device):
        return sum(self.remaining(i) for i in range(self.cell_count()) if self.owner[i] == device and self.reserved(i))

    def passage(self, device, i, since_gpu_claim=None):
        """since_gpu_claim: seconds since the GPU's last claim (None: it has not claimed)."""
        gpu, cpu, workers, window = self.rates
        if gpu <= 0 or cpu <= 0:
            return True
        chunk = float(self.spc) * TILE_PIXELS
        if device == CPU:
            if since_gpu_claim is not None and since_gpu_claim > max(4 * window, 0.05):
                return True
            return chunk / cpu < 0.5 * float(self.reserved_chunks(GPU)) * chunk / gpu
        if f32(self.gpu_p[i]) < f32(0.2):
            return False
        return float(self.reserved_chunks(CPU)) * chunk / (cpu * max(1, workers)) > 4 * window

    def neighbours(self, i):
        x, y = i % self.cx, i // self.cx
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                nx, ny = x + dx, y + dy
                if (dx or dy) and 0 <= nx < self.cx and 0 <= ny < self.cy:
                    yield ny * self.cx + nx

    def choose(self, device, current, salt):
        n = self.cell_count()

        def tier(i):
            return 1 if not self.reserved(i) else 0 if self.owner[i] == device else 2

        cpu_passage = []

        def passes(i):
            if device == GPU:
                return self.passage(device, i)
            if not cpu_passage:
                cpu_passage.append(self.passage(device, i))
            return cpu_passage[0]

        if current < n and self.claimable(current, device) and (tier(current) != 2 or passes(current)):
            return current
        for t in range(3):
            best = [NO_CELL, 0.0]

            def consider(i):
                if tier(i) != t:
                    return
                work = self.claimable(i, device)
                if not work or (t == 2 and not passes(i)):
                    return
                key = f32(self.score[i] + f32(f32(1e-9) * f32(work))) if t == 1 else f32(work)
                if best[0] == NO_CELL or key > best[1] or (key == best[1] and spread(i, salt) > spread(best[0], salt)):
                    best[0], best[1] = i, key

            if current < n:
                for j in self.neighbours(current):
                    consider(j)
                if best[0] != NO_CELL:
                    return best[0]
            for j in range(n):
                consider(j)
            if best[0] != NO_CELL:
                return best[0]
        return NO_CELL

    def seed(self, device, worker, workers):
        own = [i for i in range(self.cell_count()) if self.owner[i] == device]
        if not own:
            return self.choose(device, NO_CELL, worker)
        return own[(worker * len(own)) // max(1, workers)]

    def rendered(self, i, device, samples):
        self.samples[i][device] += samples


class Ownership:
    """Owners across frames (async_grid.h:402-437)."""
    WINDOW = 8

    def __init__(self):
        self.owners_ = []
        self.history = []
        self.last_flip = []
        self.frame = 0
        self.frozen_until = 0

    def owners(self, cells):
        return list(self.owners_) if len(self.owners_) == cells else None

    def finished(self, grid):
        n = grid.cell_count()
        if len(self.owners_) != n:
            self.owners_ = list(grid.owner)
            self.history = [0.0] * (n * self.WINDOW)
            self.last_flip = [-1000000] * n
        for i in range(n):
            cpu, gpu = float(grid.samples[i][CPU]), float(grid.samples[i][GPU])
            self.history[i * self.WINDOW + self.frame % self.WINDOW] = f32(gpu / (cpu + gpu)) if cpu + gpu > 0 else 0.5
        self.frame += 1
        if self.frame < self.frozen_until or self.frame < self.WINDOW:
            return
        flips = []
        for i in range(n):
            if self.frame - self.last_flip[i] < 16:
                continue
            share = f32(0.0)
            for k in range(self.WINDOW):
                share = f32(share + self.history[i * self.WINDOW + k])
            share = f32(share / self.WINDOW)
            other = share if self.owners_[i] == CPU else f32(1 - share)
            if other > f32(0.6):
                flips.append((i, other))
        flips.sort(key=lambda f: -f[1])  # stable: equal shares keep cell order
        most = min(8, max(1, n // 10))
        for i, _ in flips[:most]:
            if self.frame - self.last_flip[i] < 32:
                self.frozen_until = self.frame + 64
                return
            self.owners_[i] = GPU if self.owners_[i] == CPU else CPU
            self.last_flip[i] = self.frame
