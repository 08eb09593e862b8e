"""Frames from the work grid through the broker (x2dev_render_grid): the GPU alone, and the GPU beside a CPU lane.

The GPU alone renders the oracle's film of the frame (counts exact, sums to the FP64 order) and the frame loop's,
with a viewport and a share of the frame's samples too. Beside a CPU lane that renders its chunks with the oracle
into a host film of its own (the CPU device's integrator: async_cpu.h:1584-1673 renderChunk), the two films together
are the oracle's whole frame: a frame rendered by both devices. Beside a CPU lane stand-in that records the chunks it
claims and takes time over each, so the rights of passage and the end-of-frame split come into play, the two lanes
claim every chunk exactly once. Ownership across frames follows the shares the lanes rendered."""
import threading
import time

import numpy as np
import pytest

import oracle
from helpers import drain
from x2amd import abi, scenes
from x2amd import grid as G
from x2amd.lane import GpuLane

pytestmark = pytest.mark.gpu


def film(lane):
    lane.download_film()
    while not lane.film_ready():
        pass
    return lane.film()


def render_chunks(lane, config, chunks, spc):
    """The chunks (tile, first, count) rendered on a fresh begin of the lane; returns its film."""
    lane.begin(config, spc)
    cap = lane.max_chunks_per_launch()
    for i in range(0, len(chunks), cap):
        part = chunks[i:i + cap]
        arr = (abi.Chunk * len(part))(*[abi.Chunk(t, f, c, 0) for t, f, c in part])
        while not lane.ready():
            lane.poll()
        assert lane.launch(arr, samples=0)
    drain(lane)
    return film(lane)


def close(a_sums, b_sums):
    return np.abs(a_sums - b_sums).max() <= 1e-12 * max(1.0, float(np.abs(b_sums).max()))


@pytest.mark.parametrize("window_ms", [0.0, 100.0])
def test_gpu_alone_equals_frame_loop(window_ms):
    scene, config = scenes.cornell(width=96, height=80, spp=12)
    lane = GpuLane(0)
    try:
        lane.prepare(scene)
        ws, wc, _ = lane.render_frame(config)
        g = G.WorkGrid(config.width, config.height, config.spp, G.options(32, 4, 1))
        gs, gc, fr = lane.render_grid(config, g, window_ms=window_ms)
    finally:
        lane.close()
    print(f"window {window_ms} ms: launches {fr.launches}, chunks {fr.gpu_chunks}, {fr.seconds * 1e3:.1f} ms")
    assert fr.errors == 0 and fr.cpu_chunks == 0 and fr.gpu_chunks == g.total_chunks()
    assert fr.gpu_samples == config.width * config.height * config.spp
    assert (gc == wc).all() and (gc == config.spp).all()
    assert close(gs, ws)
    os_, oc, ok = oracle.OracleScene(scene).render(config)
    assert (gc == oc).all() and close(gs, os_) and ok[2] == 0
    assert fr.rays_trace == ok[0] and fr.rays_shadow == ok[1]


def test_viewport_and_sample_share():
    """A viewport (the render region) and a share of the samples [5, 12): only those, each pixel the indices a whole
    frame gives it."""
    scene, config = scenes.cornell(width=80, height=64, spp=16)
    vp = [9, 5, 70, 50]
    config.viewport[:] = vp
    lane = GpuLane(0)
    try:
        lane.prepare(scene)
        g = G.WorkGrid(config.width, config.height, config.spp, G.options(32, 4, 1, True, 5, 7), viewport=vp)
        gs, gc, fr = lane.render_grid(config, g)
        ranges = g.ranges()
        chunks = [(t, f, c) for t in range(((80 + 7) // 8) * ((64 + 7) // 8)) for f, c in ranges
                  if g.tile_pixels(t)]
        rs, rc = render_chunks(lane, config, chunks, 4)
    finally:
        lane.close()
    y, x = np.divmod(np.arange(80 * 64), 80)
    inside = (x >= vp[0]) & (x < vp[2]) & (y >= vp[1]) & (y < vp[3])
    assert ranges[0] == (5, 1) and sum(c for _, c in ranges) == 7
    assert (gc[inside] == 7).all() and (gc[~inside] == 0).all()
    assert (gc == rc).all() and close(gs, rs)
    assert fr.gpu_samples == int(inside.sum()) * 7


class OracleCpuLane:
    """A CPU lane that renders: each worker adds its chunks' samples, computed by the oracle (the CPU integrator this
    device is pinned to), into its own host film."""

    def __init__(self, scene, config, threads):
        self.o = oracle.OracleScene(scene)
        self.config = config
        n = config.width * config.height
        self.sums = [np.zeros((n, 4), np.float64) for _ in range(threads)]
        self.counts = [np.zeros(n, np.uint32) for _ in range(threads)]
        self.lock = threading.Lock()
        self.chunks = 0
        self.errors = 0

    def __call__(self, worker, chunk):
        one = (abi.Chunk * 1)(abi.Chunk(chunk.tile, chunk.first, chunk.count, chunk.cell))
        k = self.o.render_into(self.config, one, self.sums[worker], self.counts[worker], threads=1)
        with self.lock:
            self.chunks += 1
            self.errors += k[2]
        return 0

    def film(self):
        return sum(self.sums), sum(self.counts)


@pytest.mark.parametrize("size,spp,threads", [(256, 16, 4), (200, 24, 6)])
def test_frame_rendered_by_gpu_and_oracle_cpu_lane(size, spp, threads):
    """The GPU and a CPU lane rendering with the oracle share a frame (both devices, each cell's owner first, then
    stealing): the GPU's film plus the CPU lane's is the oracle's whole frame, counts exact, sums to the FP64 order
    (the samples are bit-identical on either device; only their order of addition differs)."""
    scene, config = scenes.cornell(width=size, height=size, spp=spp)
    cpu = OracleCpuLane(scene, config, threads)
    lane = GpuLane(0)
    try:
        lane.prepare(scene)
        g = G.WorkGrid(config.width, config.height, config.spp, G.options(32, 4, 2))
        gs, gc, fr = lane.render_grid(config, g, cpu_render=cpu, cpu_threads=threads, window_ms=2.0, first_chunks=64)
    finally:
        lane.close()
    cs, cc = cpu.film()
    ref_s, ref_c, ref_k = oracle.OracleScene(scene).render(config)
    print(f"GPU {fr.gpu_chunks} chunks, CPU (oracle) {fr.cpu_chunks} chunks, {fr.seconds * 1e3:.1f} ms")
    assert fr.errors == 0 and cpu.errors == 0 and ref_k[2] == 0
    assert fr.gpu_chunks > 0 and fr.cpu_chunks > 0 and cpu.chunks == fr.cpu_chunks
    assert fr.gpu_chunks + fr.cpu_chunks == g.total_chunks()
    assert (gc + cc == ref_c).all() and (ref_c == config.spp).all()
    assert close(gs + cs, ref_s)
    assert 0 < int(cc.sum()) < int(ref_c.sum())


class SlowCpuLane:
    """A CPU lane stand-in: records each chunk it is handed and takes `seconds` over it."""

    def __init__(self, seconds):
        self.seconds = seconds
        self.lock = threading.Lock()
        self.chunks = []

    def __call__(self, worker, chunk):
        time.sleep(self.seconds)
        with self.lock:
            self.chunks.append((chunk.tile, chunk.first, chunk.count))
        return 0


@pytest.mark.parametrize("mode,cell,seconds,size", [(2, 32, 0.0005, 384), (2, 64, 0.002, 512), (0, 32, 0.0, 96)])
def test_co_rendered_frame_equals_whole_frame(mode, cell, seconds, size):
    """Frames of 23,040 and 40,960 chunks: larger than the GPU's first launches (at least two samples a launch thread),
    so the CPU lane's workers claim beside them."""
    scene, config = scenes.cornell(width=size, height=size, spp=32 if mode else 8)
    cpu = SlowCpuLane(seconds)
    lane = GpuLane(0)
    try:
        lane.prepare(scene)
        ws, wc, _ = lane.render_frame(config)
        g = G.WorkGrid(config.width, config.height, config.spp, G.options(cell, 4, mode))
        gs, gc, fr = lane.render_grid(config, g, cpu_render=cpu, cpu_threads=4, window_ms=2.0, first_chunks=64)
        cs, cc = render_chunks(lane, config, cpu.chunks, 4)
    finally:
        lane.close()
    print(f"mode {mode} cell {cell}: GPU {fr.gpu_chunks} chunks in {fr.launches} launches (stolen {fr.gpu_stolen}), "
          f"CPU {fr.cpu_chunks} (stolen {fr.cpu_stolen}), {fr.seconds * 1e3:.1f} ms")
    assert fr.errors == 0
    assert fr.cpu_chunks == len(cpu.chunks) and fr.gpu_chunks + fr.cpu_chunks == g.total_chunks()
    assert len(set((t, f) for t, f, _ in cpu.chunks)) == len(cpu.chunks)
    assert (gc + cc == config.spp).all(), "every pixel's samples exactly once between the lanes"
    assert close(gs + cs, ws)
    if mode == 0:
        assert fr.gpu_chunks == 0 and fr.launches == 0
    else:
        assert fr.gpu_chunks > 0 and fr.cpu_chunks > 0


def test_ownership_follows_the_lanes():
    """Frames from grids whose owners come from the last frames: a slow CPU lane leaves most cells to the GPU, and the
    owners flip towards it after the 8-frame window; every frame stays whole."""
    scene, config = scenes.cornell(width=128, height=128, spp=8)
    own = G.Ownership()
    lane = GpuLane(0)
    flipped = []
    try:
        lane.prepare(scene)
        for frame in range(10):
            owners = own.owners(16)
            g = G.WorkGrid(config.width, config.height, config.spp, G.options(32, 4, 2), owners=owners)
            cpu = SlowCpuLane(0.004)
            gs, gc, fr = lane.render_grid(config, g, cpu_render=cpu, cpu_threads=2, window_ms=2.0, first_chunks=32)
            assert fr.errors == 0 and fr.gpu_chunks + fr.cpu_chunks == g.total_chunks()
            cpu_counts = np.zeros_like(gc)
            for t, f, c in cpu.chunks:
                tx = t % 16
                ty = t // 16
                for yy in range(ty * 8, ty * 8 + 8):
                    cpu_counts[yy * 128 + tx * 8: yy * 128 + tx * 8 + 8] += c
            assert (gc + cpu_counts == config.spp).all(), frame
            own.finished(g)
            flipped.append(sum(1 for i in range(16) if g.owner(i) == G.GPU))
    finally:
        lane.close()
    after = own.owners(16)
    print(f"GPU-owned cells by frame: {flipped}; after: {int((after == G.GPU).sum())}")
    assert flipped[0] == 8 and int((after == G.GPU).sum()) > 8


@pytest.mark.parametrize("window_ms,first,workers,threads", [(1.0, 16, 4, 3), (20.0, 4096, 64, 6), (0.5, 1, 1, 2)])
def test_co_rendered_share_with_a_viewport(window_ms, first, workers, threads):
    """Co-rendering a viewport's share of the samples ([3, 19) of 24) under other broker settings (launch windows, first
    launches, GPU workers, CPU threads): every pixel in the viewport gets its 16 samples exactly once between the lanes,
    none outside, and the two films together equal the GPU's own frame of that share."""
    scene, config = scenes.cornell(width=200, height=144, spp=24)
    vp = [13, 9, 181, 130]
    config.viewport[:] = vp
    cpu = SlowCpuLane(0.0003)
    lane = GpuLane(0)
    try:
        lane.prepare(scene)
        alone = G.WorkGrid(config.width, config.height, config.spp, G.options(32, 4, 1, True, 3, 16), viewport=vp)
        ws, wc, _ = lane.render_grid(config, alone)
        g = G.WorkGrid(config.width, config.height, config.spp, G.options(32, 4, 2, True, 3, 16), viewport=vp)
        gs, gc, fr = lane.render_grid(config, g, cpu_render=cpu, cpu_threads=threads, window_ms=window_ms,
                                      first_chunks=first, gpu_workers=workers)
        cs, cc = render_chunks(lane, config, cpu.chunks, 4)
    finally:
        lane.close()
    y, x = np.divmod(np.arange(config.width * config.height), config.width)
    inside = (x >= vp[0]) & (x < vp[2]) & (y >= vp[1]) & (y < vp[3])
    assert fr.errors == 0 and fr.gpu_chunks + fr.cpu_chunks == g.total_chunks()
    assert ((gc + cc)[inside] == 16).all() and ((gc + cc)[~inside] == 0).all()
    assert (wc == gc + cc).all() and close(gs + cs, ws)
