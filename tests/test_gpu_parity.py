"""Device vs oracle on the same seeded inputs, through the C-ABI (the parity gate).

Per-sample radiance is computed by the same float expressions in the same order on both sides (no contraction,
correctly rounded division and square root, the shared sine/cosine polynomial), so at one sample per pixel the
device's film equals the oracle's bit for bit except where FP64 atomics add in another order (none at 1 spp).
Counts must be exact at every spp; radiance must meet a mean relative difference <= 1e-3 (the stated tolerance)."""
import ctypes

import numpy as np
import pytest

import oracle
from helpers import drain, mean_relative_difference
from x2amd import abi, scenes
from x2amd.lane import GpuLane

pytestmark = pytest.mark.gpu
TOL = 1e-3


@pytest.fixture(scope="module")
def cornell_lane():
    scene, _ = scenes.cornell()
    lane = GpuLane(0)
    lane.prepare(scene)
    yield scene, lane
    lane.close()


def test_cornell_one_spp_bit_exact(cornell_lane):
    scene, lane = cornell_lane
    _, config = scenes.cornell(spp=1)
    chunks = scenes.uniform_chunks(config, samples_per_chunk=1)
    lane.begin(config, samples_per_chunk=1)
    assert lane.launch(chunks, samples=config.width * config.height)
    done = drain(lane)
    assert sum(c.chunks for c in done) == len(chunks) and all(c.errors == 0 for c in done)
    lane.download_film()
    while not lane.film_ready():
        pass
    sums, counts = lane.film()
    ref_sums, ref_counts, ref_k = oracle.OracleScene(scene).render(config)
    assert (counts == 1).all() and (ref_counts == 1).all()
    exact = np.all(sums == ref_sums, axis=1).mean()
    rel = mean_relative_difference(sums, counts, ref_sums, ref_counts)
    print(f"bit-exact pixels {exact:.6f}, mean relative difference {rel:.3e}")
    assert rel <= TOL
    assert exact == 1.0
    # The device counted the same rays as the oracle traced.
    assert sum(c.rays_trace for c in done) == ref_k[0]
    assert sum(c.rays_shadow for c in done) == ref_k[1]


def test_cornell_frame_counts_and_radiance(cornell_lane):
    scene, lane = cornell_lane
    _, config = scenes.cornell()
    sums, counts, stats = lane.render_frame(config)
    ref_sums, ref_counts, _ = oracle.OracleScene(scene).render(config)
    assert (counts == config.spp).all(), "every pixel exactly spp samples"
    assert stats.audit_failures == 0 and stats.errors == 0
    rel = mean_relative_difference(sums, counts, ref_sums, ref_counts)
    print(f"cornell 256x256x16: launches {stats.launches}, rel {rel:.3e}, {stats.device_seconds * 1e3:.2f} ms device")
    assert rel <= TOL


def test_icospheres_reduced(cornell_lane):
    scene, config = scenes.icospheres(width=160, height=90, spp=2)
    lane = GpuLane(0)
    lane.prepare(scene)
    sums, counts, stats = lane.render_frame(config)
    lane.close()
    ref_sums, ref_counts, ref_k = oracle.OracleScene(scene).render(config)
    assert (counts == config.spp).all()
    rel = mean_relative_difference(sums, counts, ref_sums, ref_counts)
    print(f"icospheres 160x90x2: rel {rel:.3e}, traces {stats.rays_trace} vs {ref_k[0]}")
    assert rel <= TOL
    assert stats.rays_trace == ref_k[0] and stats.rays_shadow == ref_k[1]


def test_launch_poll_protocol(cornell_lane):
    scene, lane = cornell_lane
    _, config = scenes.cornell(width=64, height=48, spp=8)
    chunks = scenes.uniform_chunks(config, samples_per_chunk=4)
    lane.begin(config, samples_per_chunk=4)
    third = len(chunks) // 3
    parts = [chunks[:third], chunks[third:2 * third], chunks[2 * third:]]
    copies = [(abi.Chunk * len(p))(*p) for p in parts]
    assert lane.ready()
    assert lane.launch(copies[0], samples=11)
    assert lane.launch(copies[1], samples=22)
    assert not lane.ready()  # at most two in flight (async_gpu.h:1425)
    with pytest.raises(RuntimeError):
        lane.launch(copies[2])
    done = drain(lane)
    assert [c.samples for c in done] == [11, 22]  # collected oldest first
    assert lane.launch(copies[2], samples=33)
    with pytest.raises(RuntimeError):
        lane.download_film()  # a launch is still in flight
    done += drain(lane)
    assert sum(c.chunks for c in done) == len(chunks)
    assert all(c.device_seconds > 0 for c in done)
    assert lane.launch((abi.Chunk * 0)()) is False  # nothing to run
    lane.download_film()
    while not lane.film_ready():
        pass
    sums, counts = lane.film()
    assert (counts == config.spp).all()
    ref_sums, ref_counts, _ = oracle.OracleScene(scene).render(config)
    assert mean_relative_difference(sums, counts, ref_sums, ref_counts) <= TOL


def test_argument_errors(cornell_lane):
    scene, lane = cornell_lane
    _, config = scenes.cornell(width=16, height=16, spp=2)
    lane.begin(config, samples_per_chunk=1)
    bad = (abi.Chunk * 1)(abi.Chunk(0, 0, 2, 0))  # count above samples_per_chunk
    with pytest.raises(RuntimeError):
        lane.launch(bad)
    bad = (abi.Chunk * 1)(abi.Chunk(4, 0, 1, 0))  # tile outside the image (2x2 tiles)
    with pytest.raises(RuntimeError):
        lane.launch(bad)
    fresh = GpuLane(0)
    with pytest.raises(RuntimeError):
        fresh.begin(config)  # before prepare
    fresh.close()
    cfg = scenes.cornell(width=16, height=16, spp=2)[1]
    cfg.bsdf_samples = 0
    with pytest.raises(RuntimeError):
        lane.begin(cfg)
