"""The wavefront form of the surface estimator (x2dev_options.wave_slots) against the oracle and against the
one-kernel form: the same samples, so at 1 spp the film equals the oracle's bit for bit with identical ray counts,
and at higher spp the two forms' films differ only by the order of FP64 atomic adds. A launch's iterations are
estimated on the host and its drain kernel finishes whatever they leave; a small slot pool makes the drain do most
of the work, which must not change the result."""
import numpy as np
import pytest

import oracle
from helpers import drain, mean_relative_difference
from x2amd import abi, scenes
from x2amd.lane import GpuLane

pytestmark = pytest.mark.gpu


def frame(scene, config, wave_slots, spc=4):
    lane = GpuLane(0, wave_slots=wave_slots)
    try:
        lane.prepare(scene)
        sums, counts, st = lane.render_frame(config, samples_per_chunk=spc)
    finally:
        lane.close()
    return sums, counts, st


def test_wavefront_one_spp_bit_exact():
    scene, config = scenes.cornell(spp=1)
    lane = GpuLane(0, wave_slots=4096)
    try:
        lane.prepare(scene)
        chunks = scenes.uniform_chunks(config, samples_per_chunk=1)
        lane.begin(config, samples_per_chunk=1)
        assert lane.launch(chunks, samples=config.width * config.height)
        done = drain(lane)
        lane.download_film()
        while not lane.film_ready():
            pass
        sums, counts = lane.film()
    finally:
        lane.close()
    assert sum(c.chunks for c in done) == len(chunks) and all(c.errors == 0 for c in done)
    ref_sums, ref_counts, ref_k = oracle.OracleScene(scene).render(config)
    assert (counts == 1).all()
    exact = np.all(sums == ref_sums, axis=1).mean()
    print(f"wavefront cornell 1 spp: bit-exact pixels {exact:.6f}")
    assert exact == 1.0
    assert sum(c.rays_trace for c in done) == ref_k[0]
    assert sum(c.rays_shadow for c in done) == ref_k[1]


@pytest.mark.parametrize("slots", [256, 65536])
def test_wavefront_matches_one_kernel(slots):
    scene, config = scenes.cornell()
    a_sums, a_counts, a_st = frame(scene, config, 0)
    b_sums, b_counts, b_st = frame(scene, config, slots)
    assert (a_counts == config.spp).all() and (b_counts == a_counts).all()
    assert b_st.errors == 0 and b_st.audit_failures == 0
    assert (a_st.rays_trace, a_st.rays_shadow) == (b_st.rays_trace, b_st.rays_shadow)
    rel = mean_relative_difference(b_sums, b_counts, a_sums, a_counts)
    print(f"{slots} slots: {b_st.launches} launches, mean relative difference to one kernel {rel:.2e}")
    assert rel <= 1e-12  # the FP64 atomics' order only
    ref_sums, ref_counts, _ = oracle.OracleScene(scene).render(config)
    assert mean_relative_difference(b_sums, b_counts, ref_sums, ref_counts) <= 1e-3


def test_wavefront_headline_scene_reduced():
    scene, config = scenes.icospheres(width=160, height=90, spp=2)
    sums, counts, st = frame(scene, config, 8192, spc=1)
    ref_sums, ref_counts, ref_k = oracle.OracleScene(scene).render(config)
    assert (counts == config.spp).all()
    assert st.rays_trace == ref_k[0] and st.rays_shadow == ref_k[1]
    assert mean_relative_difference(sums, counts, ref_sums, ref_counts) <= 1e-3


def test_wave_slots_must_be_whole_workgroups():
    with pytest.raises(RuntimeError, match="wave_slots"):
        GpuLane(0, wave_slots=300)
