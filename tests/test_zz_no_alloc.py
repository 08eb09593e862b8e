"""The allocation-free render (x2dev_render_chunks_static) against the lane's own launch (x2dev_launch), through the
C-ABI: the same chunks give the same film bit for bit (two samples a pixel: one FP64 addition a pixel, which is
commutative, so the atomics' order cannot change a bit) and the same ray counts; a second call with fewer chunks
starts from cleared counters (the static array keeps the last call's contents); more chunks than the static array
holds are refused before any device work."""
import ctypes

import numpy as np
import pytest
import torch

from helpers import drain
from x2amd import abi, scenes
from x2amd.lane import GpuLane

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def lane_film(lane, config, chunks):
    """The existing entry point: begin, launch the chunks, poll, download the lane's film."""
    lane.begin(config, samples_per_chunk=1)
    assert lane.launch(chunks)
    done = drain(lane)
    lane.download_film()
    while not lane.film_ready():
        pass
    sums, counts = lane.film()
    return sums, counts, done


def static_film(lane, config, chunks):
    """The allocation-free entry point into a film the caller owns (device memory from torch)."""
    pixels = config.width * config.height
    sums = torch.zeros(pixels * 4, dtype=torch.float64, device=DEV)
    counts = torch.zeros(pixels, dtype=torch.int32, device=DEV)
    torch.cuda.synchronize()
    out = abi.Completion()
    r = lane.lib.x2dev_render_chunks_static(lane.h, ctypes.byref(config), 1, ctypes.cast(chunks, ctypes.POINTER(abi.Chunk)),
                                            len(chunks), ctypes.c_void_p(sums.data_ptr()),
                                            ctypes.c_void_p(counts.data_ptr()), ctypes.byref(out))
    abi.check(r, "x2dev_render_chunks_static")
    torch.cuda.synchronize()
    return (sums.cpu().numpy().reshape(pixels, 4), counts.cpu().numpy().view(np.uint32), out)


def test_static_render_equals_the_lane_launch():
    scene, config = scenes.cornell(width=72, height=52, spp=2)
    all_chunks = scenes.uniform_chunks(config, samples_per_chunk=1)
    first_range = (abi.Chunk * (len(all_chunks) // 2))(*[c for c in all_chunks if c.first == 0])
    assert len(first_range) * 2 == len(all_chunks)
    lane = GpuLane(0)
    try:
        lane.prepare(scene)
        # two calls in a row, the second with a smaller input
        for chunks, spp in ((all_chunks, 2), (first_range, 1)):
            print(f"static render: {len(chunks)} chunks, {spp} samples a pixel", flush=True)
            ref_sums, ref_counts, done = lane_film(lane, config, chunks)
            sums, counts, out = static_film(lane, config, chunks)
            assert (ref_counts == spp).all() and (counts == ref_counts).all()
            assert np.array_equal(sums.view(np.uint64), ref_sums.view(np.uint64)), "bit-exact film"
            assert out.errors == 0 and out.chunks == len(chunks)
            assert out.samples == config.width * config.height * spp
            assert out.rays_trace == sum(c.rays_trace for c in done)
            assert out.rays_shadow == sum(c.rays_shadow for c in done)
            assert out.nodes_visited == sum(c.nodes_visited for c in done) and out.device_seconds > 0
    finally:
        lane.close()


def test_static_render_refuses_more_chunks_than_it_holds():
    scene, config = scenes.cornell(width=64, height=64, spp=1)
    lane = GpuLane(0)
    try:
        lane.prepare(scene)
        n = abi.X2_STATIC_MAX_CHUNKS + 1
        chunks = (abi.Chunk * n)()  # tile 0, no samples: valid chunks, too many of them
        film = torch.zeros(64 * 64 * 4, dtype=torch.float64, device=DEV)
        counts = torch.zeros(64 * 64, dtype=torch.int32, device=DEV)
        r = lane.lib.x2dev_render_chunks_static(lane.h, ctypes.byref(config), 1, chunks, n,
                                                ctypes.c_void_p(film.data_ptr()), ctypes.c_void_p(counts.data_ptr()),
                                                None)
        assert r == -1 and b"X2_STATIC_MAX_CHUNKS" in lane.lib.x2dev_last_error()
        assert int(counts.sum()) == 0
    finally:
        lane.close()
