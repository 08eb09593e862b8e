"""GpuLane: the reference's GPU-lane interface (xpu2/interactive/async_gpu.h:82-1731) over the C-ABI device.

Method names and meaning follow the reference: begin(samples_per_chunk) zeroes the film; ready() allows at most two
launches in flight; launch(chunks, samples) copies the chunks and queues a render, False when there was nothing to
run; poll() never waits and returns the oldest finished launch's Completion or None; download_film()/film_ready()
bring the film to the host; sums()/counts() view it. Failures raise RuntimeError, as the reference's cpu::need
throws std::runtime_error. There is no CPU fallback: without the built library every call raises.
"""
import ctypes

import numpy as np

from . import abi


class GpuLane:
    def __init__(self, device=0, max_chunks_per_launch=16384, ring_chunks=1 << 17, counters=True, wave_slots=0,
                 lane=0):
        self.lib = abi.load()
        opts = abi.Options(device, max_chunks_per_launch, ring_chunks, 1 if counters else 0, wave_slots, lane)
        h = ctypes.c_void_p()
        abi.check(self.lib.x2dev_create(ctypes.byref(opts), ctypes.byref(h)), "x2dev_create")
        self.h = h
        self.config = None
        self._scene = None

    def close(self):
        if getattr(self, "h", None):
            self.lib.x2dev_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def prepare(self, scene):
        """Uploads a scenes.SceneData (BVH8 built on the host)."""
        abi.check(self.lib.x2dev_prepare(self.h, ctypes.byref(scene.c)), "x2dev_prepare")
        self._scene = scene

    def begin(self, config, samples_per_chunk=4):
        abi.check(self.lib.x2dev_begin(self.h, ctypes.byref(config), samples_per_chunk), "x2dev_begin")
        self.config = config

    def begin_adaptive(self, config, samples_per_chunk=4):
        """GpuLane::begin(spc, adaptive=true): the odd-index film kept, chunks read through the pass tables."""
        abi.check(self.lib.x2dev_begin_adaptive(self.h, ctypes.byref(config), samples_per_chunk),
                  "x2dev_begin_adaptive")
        self.config = config

    def upload_pass(self, slot, masks, uniform_base=abi.X2_NO_SLOT):
        """GpuLane::uploadPass: a pass slot's tile masks (uint64 a tile) and bases (uniform_base, or the counts)."""
        m = np.ascontiguousarray(masks, np.uint64)
        abi.check(self.lib.x2dev_upload_pass(self.h, slot, m.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                             uniform_base), "x2dev_upload_pass")

    def errors(self):
        """GpuLane::requestErrors, then errors(): every pixel's adaptive error (float32), waited for."""
        abi.check(self.lib.x2dev_request_errors(self.h), "x2dev_request_errors")
        while not abi.check(self.lib.x2dev_errors_ready(self.h), "x2dev_errors_ready"):
            pass
        e = ctypes.POINTER(ctypes.c_float)()
        abi.check(self.lib.x2dev_errors(self.h, ctypes.byref(e)), "x2dev_errors")
        return np.ctypeslib.as_array(e, shape=(self.config.width * self.config.height,)).copy()

    def seed_film(self, sums, counts, odd=None):
        """GpuLane::seedFilm: the frame starts from a saved film (after begin, ahead of any launch)."""
        s = np.ascontiguousarray(sums, np.float64)
        c = np.ascontiguousarray(counts, np.uint32)
        o = None if odd is None else np.ascontiguousarray(odd, np.float64)
        abi.check(self.lib.x2dev_seed_film(self.h, s.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                           None if o is None else o.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                           c.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))), "x2dev_seed_film")

    def odd(self):
        """The odd-index sums [pixels, 4] of the downloaded film (adaptive frames)."""
        o = ctypes.POINTER(ctypes.c_double)()
        abi.check(self.lib.x2dev_odd(self.h, ctypes.byref(o)), "x2dev_odd")
        n = self.config.width * self.config.height
        return np.ctypeslib.as_array(o, shape=(n * 4,)).reshape(n, 4).copy()

    def render_adaptive(self, config, min_samples, max_samples, target_error, samples_per_chunk=4, device_chain=False,
                        cellular=False, sample_budget=0, noise_floor=0.0):
        """One adaptive frame (host trees, the device pass chain, or cellular generations with an optional sample
        budget); returns (sums, odd, counts, FrameStats, x2_adaptive with passes and tree updates)."""
        st = abi.FrameStats()
        a = abi.Adaptive(min_samples, max_samples, target_error, 0, 0, 1 if device_chain else 0, 1 if cellular else 0,
                         sample_budget, noise_floor, 0)
        r = self.lib.x2dev_render_adaptive(self.h, ctypes.byref(config), samples_per_chunk, ctypes.byref(a),
                                           ctypes.byref(st))
        self.config = config
        if r < 0:
            err = self.lib.x2dev_last_error()
            raise RuntimeError(f"x2dev_render_adaptive failed ({r}): {err.decode() if err else ''}")
        sums, counts = self.film()
        return sums, self.odd(), counts, st, a

    def ready(self):
        return bool(self.lib.x2dev_ready(self.h))

    def in_flight(self):
        return int(self.lib.x2dev_in_flight(self.h))

    def launch_threads(self):
        return int(self.lib.x2dev_launch_threads(self.h))

    def max_chunks_per_launch(self):
        return int(self.lib.x2dev_max_chunks_per_launch(self.h))

    def launch(self, chunks, samples=0, refines=None):
        """chunks: a ctypes array of abi.Chunk (or a slice copy); refines: a ctypes array of abi.Refine to render
        again first (host refines). Returns True when queued."""
        n = len(chunks)
        ptr = ctypes.cast(chunks, ctypes.POINTER(abi.Chunk)) if n else None
        if refines is not None and len(refines):
            rptr = ctypes.cast(refines, ctypes.POINTER(abi.Refine))
            return abi.check(self.lib.x2dev_launch_refines(self.h, ptr, n, rptr, len(refines), samples),
                             "x2dev_launch_refines") == 1
        return abi.check(self.lib.x2dev_launch(self.h, ptr, n, samples), "x2dev_launch") == 1

    def take_refines(self):
        """The provisional samples the collected launches made (GpuLane::takeRefines), a ctypes array of abi.Refine."""
        n = int(self.lib.x2dev_refines_waiting(self.h))
        buf = (abi.Refine * max(1, n))()
        got = ctypes.c_uint32()
        abi.check(self.lib.x2dev_take_refines(self.h, buf, n, ctypes.byref(got)), "x2dev_take_refines")
        return (abi.Refine * got.value)(*buf[:got.value])

    def poll(self):
        c = abi.Completion()
        r = abi.check(self.lib.x2dev_poll(self.h, ctypes.byref(c)), "x2dev_poll")
        return c if r == 1 else None

    def take_deferred(self):
        buf = (abi.Parked * 1)()
        n = ctypes.c_uint32()
        abi.check(self.lib.x2dev_take_deferred(self.h, buf, 1, ctypes.byref(n)), "x2dev_take_deferred")
        return [buf[i] for i in range(n.value)]

    def download_film(self):
        abi.check(self.lib.x2dev_download_film(self.h), "x2dev_download_film")

    def film_ready(self):
        return abi.check(self.lib.x2dev_film_ready(self.h), "x2dev_film_ready") == 1

    def film(self, copy=True):
        """(sums [pixels, 4] float64, counts [pixels] uint32) of the downloaded film: copies, or with copy=False views
        of the lane's pinned host film, valid until its next download (GpuLane::sums/counts)."""
        s = ctypes.POINTER(ctypes.c_double)()
        c = ctypes.POINTER(ctypes.c_uint32)()
        abi.check(self.lib.x2dev_film(self.h, ctypes.byref(s), ctypes.byref(c)), "x2dev_film")
        n = self.config.width * self.config.height
        sums = np.ctypeslib.as_array(s, shape=(n * 4,)).reshape(n, 4)
        counts = np.ctypeslib.as_array(c, shape=(n,))
        return (sums.copy(), counts.copy()) if copy else (sums, counts)

    def set_texture(self, index, texels, on_device=False):
        """texels: a float32 array (host) or a device pointer (int) when on_device."""
        if on_device:
            ptr = ctypes.c_void_p(int(texels))
        else:
            arr = np.ascontiguousarray(texels, dtype=np.float32)
            ptr = arr.ctypes.data_as(ctypes.c_void_p)
        abi.check(self.lib.x2dev_set_texture(self.h, index, ptr, 1 if on_device else 0), "x2dev_set_texture")

    def set_material(self, index, material):
        abi.check(self.lib.x2dev_set_material(self.h, index, ctypes.byref(material)), "x2dev_set_material")

    def resolve_image(self, device_rgb_ptr):
        abi.check(self.lib.x2dev_resolve_image(self.h, ctypes.c_void_p(int(device_rgb_ptr))), "x2dev_resolve_image")

    def adjoint_frame(self, config, d_image_ptr, grad_textures_ptr, grad_roughness_ptr, samples_per_chunk=4):
        """The path-replay adjoint (device pointers: dLoss/dI per pixel rgb, texel and roughness gradients)."""
        st = abi.FrameStats()
        r = self.lib.x2dev_adjoint_frame(self.h, ctypes.byref(config), samples_per_chunk,
                                         ctypes.c_void_p(int(d_image_ptr)), ctypes.c_void_p(int(grad_textures_ptr or 0)),
                                         ctypes.c_void_p(int(grad_roughness_ptr)), ctypes.byref(st))
        if r < 0:
            err = self.lib.x2dev_last_error()
            raise RuntimeError(f"x2dev_adjoint_frame failed ({r}): {err.decode() if err else ''}")
        return st

    def render_frame(self, config, samples_per_chunk=4, copy=True):
        """One uniform frame through the broker counterpart; returns (sums, counts, FrameStats) (copy=False: views of
        the pinned host film, valid until the lane's next download)."""
        st = abi.FrameStats()
        r = self.lib.x2dev_render_frame(self.h, ctypes.byref(config), samples_per_chunk, ctypes.byref(st))
        self.config = config
        if r < 0:
            err = self.lib.x2dev_last_error()
            raise RuntimeError(f"x2dev_render_frame failed ({r}): {err.decode() if err else ''}")
        sums, counts = self.film(copy)
        return sums, counts, st

    def chunk_room(self):
        return self.lib.x2dev_chunk_room(self.h)

    def render_grid(self, config, grid, cpu_render=None, cpu_threads=0, window_ms=0.0, first_chunks=0, gpu_workers=0):
        """One frame from a work grid (x2amd.grid.WorkGrid) through the broker (x2dev_render_grid): the GPU fed by
        launches of about window_ms and, with cpu_render (called as cpu_render(worker, chunk) on cpu_threads worker
        threads, chunk an x2_chunk), a CPU lane beside it. Returns (sums, counts, GridFrame): the GPU's film."""
        lane = None
        keep = None
        if cpu_render is not None and cpu_threads:
            def call(_context, worker, chunk):
                try:
                    return int(cpu_render(worker, chunk.contents) or 0)
                except Exception:
                    return 1
            keep = abi.CPU_RENDER(call)
            lane = abi.CpuLane(cpu_threads, 0, None, keep)
        opts = abi.BrokerOptions(window_ms, first_chunks, gpu_workers)
        out = abi.GridFrame()
        r = self.lib.x2dev_render_grid(self.h, ctypes.byref(config), grid.handle,
                                       ctypes.byref(lane) if lane is not None else None, ctypes.byref(opts),
                                       ctypes.byref(out))
        self.config = config
        del keep
        if r < 0:
            err = self.lib.x2dev_last_error()
            raise RuntimeError(f"x2dev_render_grid failed ({r}): {err.decode() if err else ''}")
        sums, counts = self.film()
        return sums, counts, out


    def page_textures(self, pool_tiles, coarse_side=0, partial=False, host_refines=False):
        """On-demand texture tiles (x2dev_page_textures): every mip level wider or taller than coarse_side (0: 64) in
        32x32 half-float tiles, pool_tiles of them on the device, the coarser levels resident; with partial, a sample
        meeting a missing tile reads the texel's parent on the nearest resident level up, provisionally; with
        host_refines the provisional samples go to the host (take_refines) and come back through launch(refines=)."""
        abi.check(self.lib.x2dev_page_textures(self.h, ctypes.byref(abi.Paging(pool_tiles, coarse_side,
                                                                                 1 if partial else 0,
                                                                                 1 if host_refines else 0))),
                  "x2dev_page_textures")

    def paging_stats(self):
        st = abi.PagingStats()
        abi.check(self.lib.x2dev_paging_stats(self.h, ctypes.byref(st)), "x2dev_paging_stats")
        return st

    def parked_outstanding(self):
        return int(self.lib.x2dev_parked_outstanding(self.h))

    def provisional_film(self):
        """(sums [pixels, 4], counts [pixels]) of the provisional samples, as downloaded with the film."""
        s = ctypes.POINTER(ctypes.c_double)()
        c = ctypes.POINTER(ctypes.c_uint32)()
        abi.check(self.lib.x2dev_provisional_film(self.h, ctypes.byref(s), ctypes.byref(c)), "x2dev_provisional_film")
        n = self.config.width * self.config.height
        return (np.ctypeslib.as_array(s, shape=(n * 4,)).reshape(n, 4).copy(),
                np.ctypeslib.as_array(c, shape=(n,)).copy())

    def render_staged(self, config, path=None, stage_samples=0, resume=False, max_stages=0, samples_per_chunk=4):
        """A checkpoint render (x2dev_render_staged): the frame in stages, its film saved to path after each, taken
        up from a save at path with resume. Returns (sums, counts, StagedFrame): the film so far."""
        st = abi.Staged(path.encode() if path else None, stage_samples, 1 if resume else 0, max_stages, 0)
        out = abi.StagedFrame()
        r = self.lib.x2dev_render_staged(self.h, ctypes.byref(config), samples_per_chunk, ctypes.byref(st),
                                         ctypes.byref(out))
        self.config = config
        if r < 0:
            err = self.lib.x2dev_last_error()
            raise RuntimeError(f"x2dev_render_staged failed ({r}): {err.decode() if err else ''}")
        sums, counts = self.film()
        return sums, counts, out


def render_frame(scene, config, samples_per_chunk=4, device=0):
    lane = GpuLane(device)
    try:
        lane.prepare(scene)
        return lane.render_frame(config, samples_per_chunk)
    finally:
        lane.close()


def adaptive_masks(width, height, target_error, errors):
    """The host's tree step (x2_adaptive_masks): each tile's active pixels from every pixel's error, and whether every
    region's tree converged. Needs no device."""
    lib = abi.load()
    e = np.ascontiguousarray(errors, np.float32)
    masks = np.zeros(((width + 7) // 8) * ((height + 7) // 8), np.uint64)
    done = ctypes.c_uint32(0)
    abi.check(lib.x2_adaptive_masks(width, height, target_error, e.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                    masks.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), ctypes.byref(done)),
              "x2_adaptive_masks")
    return masks, bool(done.value)


def cellular_step(width, height, target_error, alive, in_image, errors, planned=None, samples_left=0.0, max_samples=0,
                  noise_floor=0.0):
    """One generation of the cellular rules on the host (x2_cellular_step): (next alive uint8 per tile, ratios, live)."""
    lib = abi.load()
    a = np.ascontiguousarray(alive, np.uint8)
    m = np.ascontiguousarray(in_image, np.uint64)
    e = np.ascontiguousarray(errors, np.float32)
    out = np.zeros_like(a)
    ratios = np.zeros(len(a), np.float32)
    p = None if planned is None else np.ascontiguousarray(planned, np.uint32)
    live = abi.check(lib.x2_cellular_step(width, height, target_error, a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                          m.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                          e.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                          None if p is None else p.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                          samples_left, max_samples, noise_floor,
                                          out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                          ratios.ctypes.data_as(ctypes.POINTER(ctypes.c_float))), "x2_cellular_step")
    return out, ratios, live


def pixel_errors(sums, odd, counts):
    """pixelError over a film on the host (x2_pixel_errors). Needs no device."""
    lib = abi.load()
    s = np.ascontiguousarray(sums, np.float64)
    o = np.ascontiguousarray(odd, np.float64)
    c = np.ascontiguousarray(counts, np.uint32)
    out = np.zeros(len(c), np.float32)
    abi.check(lib.x2_pixel_errors(s.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                  o.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                  c.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), len(c),
                                  out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))), "x2_pixel_errors")
    return out
