"""Adaptive frames on the device against the oracle (oracle/oracle_adaptive.py): the pass protocol (tile masks, each
pixel's sample base, the odd-index film), the device's per-pixel errors against the oracle's and the host's, and whole
adaptive frames — the same passes and masks (so every pixel's count equal), radiance within the stated tolerance —
in both forms of the surface estimator."""
import numpy as np
import pytest

import oracle
import oracle_adaptive as A
from helpers import drain, mean_relative_difference
from x2amd import abi, scenes
from x2amd import lane as L
from x2amd.lane import GpuLane

pytestmark = pytest.mark.gpu
TOL = 1e-3


def test_pass_protocol_matches_oracle():
    scene, config = scenes.cornell(width=40, height=24, spp=3)
    w, h = config.width, config.height
    tiles_x = (w + 7) // 8
    inside = A.in_image(config)
    half = inside.copy()
    half[::2] = 0  # the second pass samples every other tile
    lane = GpuLane(0)
    try:
        lane.prepare(scene)
        lane.begin_adaptive(config, samples_per_chunk=2)
        lane.upload_pass(0, inside, 0)  # uniform: every pixel from index 0
        chunks = (abi.Chunk * len(inside))(*[abi.Chunk(t, 0 << abi.X2_PASS_SHIFT, 1, 0) for t in range(len(inside))])
        assert lane.launch(chunks, samples=w * h)
        drain(lane)
        lane.upload_pass(1, half)  # from each pixel's count
        active = [t for t in range(len(half)) if half[t]]
        chunks = (abi.Chunk * len(active))(*[abi.Chunk(t, (1 << abi.X2_PASS_SHIFT) | 0, 2, 0) for t in active])
        assert lane.launch(chunks, samples=0)
        drain(lane)
        errors = lane.errors()
        lane.download_film()
        while not lane.film_ready():
            pass
        sums, counts = lane.film()
        odd = lane.odd()
    finally:
        lane.close()
    o = oracle.OracleScene(scene)
    rs, ro = np.zeros((w * h, 4)), np.zeros((w * h, 4))
    rc = np.zeros(w * h, np.uint32)
    o.render_pixels(config, np.zeros(w * h, np.uint32), np.ones(w * h, np.uint32), rs, ro, rc)
    first, count = np.ones(w * h, np.uint32), np.zeros(w * h, np.uint32)
    for t in active:
        for q in range(64):
            x, y = (t % tiles_x) * 8 + q % 8, (t // tiles_x) * 8 + q // 8
            if x < w and y < h:
                count[y * w + x] = 2
    o.render_pixels(config, first, count, rs, ro, rc)
    assert (counts == rc).all() and set(np.unique(counts)) == {1, 3}
    assert np.all(sums == rs, axis=1).mean() >= 0.999 and np.all(odd == ro, axis=1).mean() >= 0.999
    assert (errors.view(np.uint32) == L.pixel_errors(sums, odd, counts).view(np.uint32)).all()
    ref = A.pixel_errors(rs, ro, rc)
    assert (errors.view(np.uint32) == ref.view(np.uint32)).mean() >= 0.999
    assert (errors[counts == 1] == -1).all() and (errors[counts == 3] >= 0).all()


@pytest.mark.parametrize("wave_slots", [0, 4096], ids=["one_kernel", "wavefront"])
@pytest.mark.parametrize("target", [5e-4, 1e-3])
def test_adaptive_frame_matches_oracle(target, wave_slots):
    scene, config = scenes.cornell(width=64, height=64, spp=64)
    lane = GpuLane(0, wave_slots=wave_slots)
    try:
        lane.prepare(scene)
        sums, odd, counts, st, a = lane.render_adaptive(config, 4, 64, target)
        errors = lane.errors()
    finally:
        lane.close()
    rs, ro, rc, history = A.adaptive_frame(oracle.OracleScene(scene), config, 4, 64, target)
    print(f"target {target}: passes {a.passes} (oracle {len(history)}), tree updates {a.tree_updates}, samples a pixel "
          f"{counts.mean():.1f} (min {counts.min()}, max {counts.max()})")
    assert st.errors == 0 and st.audit_failures == 0
    assert a.passes == len(history)
    assert (counts == rc).all(), "the same passes sampled the same pixels"
    assert counts.min() < counts.max(), "the frame was adaptive"
    assert mean_relative_difference(sums, counts, rs, rc) <= TOL
    assert mean_relative_difference(odd, np.maximum(counts // 2, 1), ro, np.maximum(rc // 2, 1)) <= TOL
    ref = A.pixel_errors(rs, ro, rc)
    assert (errors.view(np.uint32) == ref.view(np.uint32)).mean() >= 0.99
    assert np.abs(errors - ref).max() <= 1e-5 * max(1.0, float(np.abs(ref).max()))


@pytest.mark.parametrize("target", [5e-4, 2e-3])
def test_device_chain_matches_host_trees(target):
    """The device pass chain (trees, masks and active tiles made on the device, every pass queued at once) samples
    exactly the pixels the host's trees do; the uniform stage is one step there, so only the film's FP64 order differs."""
    scene, config = scenes.cornell(width=72, height=56, spp=64)
    lane = GpuLane(0)
    try:
        lane.prepare(scene)
        hs, ho, hc, hst, ha = lane.render_adaptive(config, 4, 64, target)
        cs, co, cc, cst, ca = lane.render_adaptive(config, 4, 64, target, device_chain=True)
    finally:
        lane.close()
    print(f"target {target}: host passes {ha.passes} tree updates {ha.tree_updates}; chain passes {ca.passes} tree "
          f"updates {ca.tree_updates}; samples a pixel {hc.mean():.1f}; chain {cst.device_seconds * 1e3:.2f} ms device")
    assert cst.errors == 0 and cst.audit_failures == 0
    assert (cc == hc).all() and hc.min() < hc.max()
    assert ca.passes == ha.passes
    assert mean_relative_difference(cs, cc, hs, hc) <= 1e-12
    assert mean_relative_difference(co, np.maximum(cc // 2, 1), ho, np.maximum(hc // 2, 1)) <= 1e-12
    _, _, rc, history = A.adaptive_frame(oracle.OracleScene(scene), config, 4, 64, target)
    assert (cc == rc).all() and len(history) == ca.passes


@pytest.mark.parametrize("target,budget", [(1e-3, 0), (3e-3, 0), (1e-3, 64 * 64 * 24), (1e-3, 64 * 64 * 12)])
def test_cellular_frame_matches_oracle(target, budget):
    """Cellular generations (and frames by a sample budget): the same tiles live at every generation, so every pixel's
    count equals the oracle's; a budget is spent exactly."""
    scene, config = scenes.cornell(width=64, height=64, spp=64)
    lane = GpuLane(0)
    try:
        lane.prepare(scene)
        sums, odd, counts, st, a = lane.render_adaptive(config, 4, 64, target, cellular=True, sample_budget=budget)
    finally:
        lane.close()
    rs, ro, rc, history = A.cellular_frame(oracle.OracleScene(scene), config, 4, 64, target, budget)
    print(f"cellular target {target} budget {budget}: passes {a.passes}, generations {a.tree_updates}, samples "
          f"{int(counts.sum())} ({counts.mean():.1f} a pixel, {counts.min()}..{counts.max()})")
    assert st.errors == 0 and st.audit_failures == 0
    assert a.passes == len(history)
    assert (counts == rc).all()
    assert mean_relative_difference(sums, counts, rs, rc) <= TOL
    if budget:
        assert int(counts.sum()) == budget
