"""The work grid and the cell owners across frames (xpu2/interactive/async_grid.h Grid, Ownership), over the device
library's native grid (x2_grid_* in include/x2dev.h). Method names and argument meaning follow the reference's: a
worker ``choose``s its next cell and ``claim``s chunks there; the broker (``GpuLane.render_grid``) feeds the GPU lane
and a CPU lane from it.

Every call here runs on the host; none needs a GPU. The library must be built (there is no Python fallback)."""
import ctypes

import numpy as np

from . import abi

CPU = abi.X2_DEVICE_CPU
GPU = abi.X2_DEVICE_GPU
NO_CELL = abi.X2_NO_CELL


def options(cell_size=64, samples_per_chunk=4, mode=2, progressive=True, sample_first=0, sample_count=0):
    """GridOptions (async_grid.h:103-113): mode 0 CPU only, 1 GPU only, 2 both."""
    return abi.GridOptions(cell_size, samples_per_chunk, mode, 1 if progressive else 0, sample_first, sample_count)


class WorkGrid:
    """Grid (async_grid.h:115-400). owners: one X2_DEVICE_* per cell (None: the checkerboard); empty: the cells start
    with nothing, for chunks appended later; viewport: (x0, y0, x1, y1) half open."""

    def __init__(self, width, height, spp, opts=None, owners=None, empty=False, viewport=None):
        self.lib = abi.load()
        self.opts = opts or options()
        self.width, self.height, self.spp = width, height, spp
        handle = ctypes.c_void_p()
        own = None
        n_own = 0
        if owners is not None:
            self._owners = np.ascontiguousarray(owners, np.uint8)
            own = self._owners.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
            n_own = len(self._owners)
        vp = (abi.i32 * 4)(*viewport) if viewport is not None else None
        abi.check(self.lib.x2_grid_create(width, height, spp, ctypes.byref(self.opts), own, n_own, 1 if empty else 0,
                                          vp, ctypes.byref(handle)), "x2_grid_create")
        self.handle = handle

    def close_handle(self):
        if self.handle:
            self.lib.x2_grid_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close_handle()
        except Exception:
            pass

    def cell_count(self):
        return self.lib.x2_grid_cell_count(self.handle)

    def total_chunks(self):
        return self.lib.x2_grid_total_chunks(self.handle)

    def owner(self, cell):
        return self.lib.x2_grid_owner(self.handle, cell)

    def remaining(self, cell):
        return self.lib.x2_grid_remaining(self.handle, cell)

    def claimable(self, cell, device=CPU):
        return self.lib.x2_grid_claimable(self.handle, cell, device)

    def empty(self):
        return bool(self.lib.x2_grid_empty(self.handle))

    def claim(self, cell, n, device=CPU, out=None):
        """Up to n chunks of the cell, as a ctypes array of x2_chunk (the first ``got`` valid) and got."""
        buf = out if out is not None else (abi.Chunk * max(1, n))()
        got = self.lib.x2_grid_claim(self.handle, cell, n, buf, device)
        return buf, got

    def claim_list(self, cell, n, device=CPU):
        buf, got = self.claim(cell, n, device)
        return [(buf[i].tile, buf[i].first, buf[i].count, buf[i].cell) for i in range(got)]

    def append(self, cell, chunks):
        arr = (abi.Chunk * len(chunks))(*[abi.Chunk(*c) for c in chunks])
        abi.check(self.lib.x2_grid_append(self.handle, cell, arr, len(chunks)), "x2_grid_append")

    def choose(self, device, current, salt):
        return self.lib.x2_grid_choose(self.handle, device, current, salt)

    def seed(self, device, worker, workers):
        return self.lib.x2_grid_seed(self.handle, device, worker, workers)

    def frontier(self, device):
        return self.lib.x2_grid_frontier(self.handle, device)

    def ranges(self):
        n = self.lib.x2_grid_ranges(self.handle, None, 0)
        out = (abi.u32 * max(1, 2 * n))()
        self.lib.x2_grid_ranges(self.handle, out, n)
        return [(out[2 * i], out[2 * i + 1]) for i in range(n)]

    def set_scores(self, scores):
        s = np.ascontiguousarray(scores, np.float32)
        abi.check(self.lib.x2_grid_set_scores(self.handle, s.ctypes.data_as(ctypes.POINTER(abi.f32)), len(s)),
                  "x2_grid_set_scores")

    def important(self, cell):
        return bool(self.lib.x2_grid_important(self.handle, cell))

    def reserved(self, cell):
        return bool(self.lib.x2_grid_reserved(self.handle, cell))

    def set_suitability(self, p):
        a = np.ascontiguousarray(p, np.float32)
        abi.check(self.lib.x2_grid_set_suitability(self.handle, a.ctypes.data_as(ctypes.POINTER(abi.f32)), len(a)),
                  "x2_grid_set_suitability")

    def set_rates(self, gpu_samples, cpu_worker_samples, cpu_workers, window_seconds):
        self.lib.x2_grid_set_rates(self.handle, gpu_samples, cpu_worker_samples, cpu_workers, window_seconds)

    def gpu_claimed(self):
        self.lib.x2_grid_gpu_claimed(self.handle)

    def passage(self, device, cell):
        return bool(self.lib.x2_grid_passage(self.handle, device, cell))

    def close(self):
        self.lib.x2_grid_close(self.handle)

    def rendered(self, cell, device, samples):
        self.lib.x2_grid_rendered(self.handle, cell, device, samples)

    def rendered_by(self, cell, device):
        return self.lib.x2_grid_rendered_by(self.handle, cell, device)

    def deferred_to_cpu(self, tile):
        self.lib.x2_grid_deferred_to_cpu(self.handle, tile)

    def cell_of(self, tile):
        return self.lib.x2_grid_cell_of(self.handle, tile)

    def tile_pixels(self, tile):
        return self.lib.x2_grid_tile_pixels(self.handle, tile)


class Ownership:
    """Ownership (async_grid.h:402-437): owners persisting across frames."""

    def __init__(self):
        self.lib = abi.load()
        handle = ctypes.c_void_p()
        abi.check(self.lib.x2_ownership_create(ctypes.byref(handle)), "x2_ownership_create")
        self.handle = handle

    def __del__(self):
        try:
            if self.handle:
                self.lib.x2_ownership_destroy(self.handle)
                self.handle = None
        except Exception:
            pass

    def owners(self, cells):
        """The owners for a grid of that many cells, or None for the checkerboard."""
        out = np.zeros(cells, np.uint8)
        known = self.lib.x2_ownership_owners(self.handle, cells, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
        return out if known else None

    def finished(self, grid):
        abi.check(self.lib.x2_ownership_finished(self.handle, grid.handle), "x2_ownership_finished")
