"""Failure injection from the host, and the failures it forces (async_gpu.h:685-686, async_render.h:736).

XPU2_SIMULATE_CARD_FAIL=k fails lane k (k >= 1) as it takes a scene, as the reference does for its simulated cards
(multi_device.h:42-46); lane 0 is never failed. XPU2_SIMULATE_SAMPLE_FAIL=p+1 makes every sample of pixel p fail as it
starts on the device: each is counted in its launch's errors, the counts reach the completion, and a frame with any
fails with the reference's message through x2dev_render_frame and x2dev_render_grid alike (the adjoint too). Nothing
faults: the kernels run to their end and only count. The rest of the film is the clean frame's, and with the
variable cleared the same lane renders the clean frame again."""
import numpy as np
import pytest
import torch

from helpers import drain
from x2amd import abi
from x2amd import grid as G
from x2amd import scenes
from x2amd.lane import GpuLane

pytestmark = pytest.mark.gpu

FALLBACK = "whole epoch must use fallback"


def chunk_frame(lane, config, spc=4):
    """Every chunk of a uniform frame through begin/launch/poll; returns (sums, counts, summed completion errors)."""
    chunks = scenes.uniform_chunks(config, spc)
    lane.begin(config, spc)
    errors = 0
    cap = lane.max_chunks_per_launch()
    for i in range(0, len(chunks), cap):
        part = chunks[i:i + cap]
        part = (abi.Chunk * len(part))(*part)
        while not lane.ready():
            c = lane.poll()
            errors += c.errors if c is not None else 0
        assert lane.launch(part, samples=0)
    errors += sum(c.errors for c in drain(lane))
    lane.download_film()
    while not lane.film_ready():
        pass
    sums, counts = lane.film()
    return sums.copy(), counts.copy(), errors


def estimator(name):
    """A small frame for each form of the estimator: the one-kernel lane loop (Cornell box, one light), the kernel
    with wave-pooled shadow rays (icospheres, four lights), the wavefront form, and the prototype estimator."""
    if name == "icospheres_pooled":
        scene, config = scenes.icospheres(width=48, height=32, spp=3, count=4, level=3)
        return scene, config, 0
    scene, config = scenes.cornell(width=40, height=24, spp=5)
    if name == "direct":
        config.direct_only, config.surface_limit = 1, 2
    return scene, config, 4096 if name == "wavefront" else 0


@pytest.mark.parametrize("name", ["cornell", "icospheres_pooled", "wavefront", "direct"])
def test_failed_samples_are_counted_and_leave_the_rest(name, monkeypatch):
    scene, config, wave_slots = estimator(name)
    p = (config.height // 2) * config.width + config.width // 2
    print(f"{name}: {config.width}x{config.height}, spp {config.spp}, failing pixel {p}", flush=True)
    lane = GpuLane(0, wave_slots=wave_slots)
    try:
        lane.prepare(scene)
        monkeypatch.delenv("XPU2_SIMULATE_SAMPLE_FAIL", raising=False)
        clean_sums, clean_counts, clean_errors = chunk_frame(lane, config)
        monkeypatch.setenv("XPU2_SIMULATE_SAMPLE_FAIL", str(p + 1))
        sums, counts, errors = chunk_frame(lane, config)
        monkeypatch.delenv("XPU2_SIMULATE_SAMPLE_FAIL")
        again_sums, again_counts, again_errors = chunk_frame(lane, config)
    finally:
        lane.close()
    assert clean_errors == 0 and again_errors == 0
    assert (clean_counts == config.spp).all() and (again_counts == config.spp).all()
    assert errors == config.spp, "every sample of the failing pixel is one counted error"
    assert counts[p] == 0 and not sums[p].any(), "a failed sample adds nothing to the film"
    rest = np.ones(counts.size, bool)
    rest[p] = False
    assert (counts[rest] == config.spp).all()
    scale = max(1.0, float(np.abs(clean_sums).max()))
    assert np.abs(sums[rest] - clean_sums[rest]).max() <= 1e-12 * scale
    assert np.abs(again_sums - clean_sums).max() <= 1e-12 * scale


def test_counted_failure_fails_the_frame(monkeypatch):
    """async_render.h:736, 955-958: a frame whose completions counted an error stops feeding and fails, before the
    sample audit, with the reference's message: from the frame loop, from the work grid's broker, and from adaptive
    frames with host trees or the device pass chain."""
    scene, config = scenes.cornell(width=40, height=24, spp=4)
    lane = GpuLane(0)
    try:
        lane.prepare(scene)
        monkeypatch.setenv("XPU2_SIMULATE_SAMPLE_FAIL", "1")
        with pytest.raises(RuntimeError, match=FALLBACK):
            lane.render_frame(config)
        g = G.WorkGrid(config.width, config.height, config.spp, G.options(32, 4, 1))
        with pytest.raises(RuntimeError, match=FALLBACK):
            lane.render_grid(config, g)
        for chain in (False, True):
            with pytest.raises(RuntimeError, match=FALLBACK):
                lane.render_adaptive(config, 4, 16, 0.01, device_chain=chain)
        monkeypatch.delenv("XPU2_SIMULATE_SAMPLE_FAIL")
        _, counts, st = lane.render_frame(config)
        assert st.errors == 0 and (counts == config.spp).all()
        g = G.WorkGrid(config.width, config.height, config.spp, G.options(32, 4, 1))
        _, counts, _ = lane.render_grid(config, g)
        assert (counts == config.spp).all()
    finally:
        lane.close()


def test_counted_failure_fails_the_adjoint(monkeypatch):
    scene, config, params = scenes.cornell_textured(width=32, height=32, spp=2, texture_size=64)
    lane = GpuLane(0)
    try:
        lane.prepare(scene)
        dev = torch.device("cuda:0")
        g = torch.ones(config.width * config.height * 3, dtype=torch.float32, device=dev)
        gt = torch.zeros(params["texture"].size, dtype=torch.float32, device=dev)
        gr = torch.zeros(scene.material_count, dtype=torch.float32, device=dev)
        torch.cuda.synchronize()
        monkeypatch.setenv("XPU2_SIMULATE_SAMPLE_FAIL", str(config.width * 5 + 7))
        with pytest.raises(RuntimeError, match="GPU sample failure in the adjoint"):
            lane.adjoint_frame(config, g.data_ptr(), gt.data_ptr(), gr.data_ptr())
        monkeypatch.delenv("XPU2_SIMULATE_SAMPLE_FAIL")
        gt.zero_()
        gr.zero_()
        torch.cuda.synchronize()
        st = lane.adjoint_frame(config, g.data_ptr(), gt.data_ptr(), gr.data_ptr())
        assert st.errors == 0
    finally:
        lane.close()


def test_simulated_card_failure_leaves_the_lane_out(monkeypatch):
    """async_gpu.h:685-686: XPU2_SIMULATE_CARD_FAIL=k fails lane k as it takes the scene; other lanes on the card,
    lane 0 always, prepare and render."""
    scene, config = scenes.cornell(width=32, height=24, spp=2)
    monkeypatch.setenv("XPU2_SIMULATE_CARD_FAIL", "1")
    lanes = [GpuLane(0, lane=k) for k in range(3)]
    try:
        with pytest.raises(RuntimeError, match="XPU2_SIMULATE_CARD_FAIL"):
            lanes[1].prepare(scene)
        for k in (0, 2):
            lanes[k].prepare(scene)
            _, counts, st = lanes[k].render_frame(config)
            assert st.errors == 0 and (counts == config.spp).all()
        monkeypatch.setenv("XPU2_SIMULATE_CARD_FAIL", "0")  # lane 0 is never failed, even when named
        lanes[0].prepare(scene)
        monkeypatch.delenv("XPU2_SIMULATE_CARD_FAIL")
        lanes[1].prepare(scene)
        _, counts, _ = lanes[1].render_frame(config)
        assert (counts == config.spp).all()
    finally:
        for lane in lanes:
            lane.close()
