"""Batched jobs on the device (x2batch_*): one persistent kernel over a device queue of several independent scenes,
replayed from a captured hipGraph, against the oracle job by job (counts exact, radiance within 1e-3, the same
rays), against the one-launch-per-job mode and a second graph replay (the same samples: the films agree to the
FP64 atomics' order), with every job's completion stamp inside the run."""
import numpy as np
import pytest

import oracle
from helpers import mean_relative_difference
from x2amd import scenes
from x2amd.batch import Batch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small_batch():
    jobs = scenes.batch_jobs(count=6, seed=11, min_tris=1_000, max_tris=30_000, min_side=16, max_side=96, max_spp=4)
    b = Batch(0)
    for scene, config in jobs:
        b.add(scene, config)
    yield jobs, b
    b.close()


def test_batch_matches_oracle_job_by_job(small_batch):
    jobs, b = small_batch
    st, secs = b.run()
    assert st.graph == 1 and st.errors == 0 and st.jobs == len(jobs)
    traces = shadows = 0
    for j, (scene, config) in enumerate(jobs):
        sums, counts = b.film(j)
        ref_sums, ref_counts, k = oracle.OracleScene(scene).render(config)
        assert (counts == config.spp).all(), j
        rel = mean_relative_difference(sums, counts, ref_sums, ref_counts)
        print(f"job {j}: {scene.triangle_count} tris, {config.width}x{config.height}x{config.spp}, rel {rel:.2e}, "
              f"done at {secs[j] * 1e3:.2f} ms")
        assert rel <= 1e-3
        traces += k[0]
        shadows += k[1]
    assert st.rays_trace == traces and st.rays_shadow == shadows
    assert np.all(secs > 0) and np.all(secs <= st.device_seconds * 1.01 + 1e-4)


def test_per_job_launches_and_replay_agree(small_batch):
    jobs, b = small_batch
    b.run()
    first = [b.film(j) for j in range(len(jobs))]
    st2, _ = b.run()  # a second replay of the same graph
    again = [b.film(j) for j in range(len(jobs))]
    st3, secs = b.run(per_job_launches=True)
    per = [b.film(j) for j in range(len(jobs))]
    assert st3.graph == 0 and st3.launches == len(jobs) + 2
    for (s1, c1), (s2, c2), (s3, c3) in zip(first, again, per):
        assert (c1 == c2).all() and (c1 == c3).all()
        assert np.allclose(s1, s2, rtol=1e-12, atol=0) and np.allclose(s1, s3, rtol=1e-12, atol=0)
    assert np.all(secs > 0)


def test_tuner_picks_a_valid_shape(small_batch):
    jobs, b = small_batch
    t = b.tune()
    print(f"tuned: {t.block_threads} threads, claim {t.claim_items}, {t.waves_per_simd} waves/SIMD, {t.seconds * 1e3:.3f} ms")
    assert t.block_threads in (64, 128, 256) and t.claim_items % 64 == 0 and t.waves_per_simd in (2, 3)
    assert t.seconds > 0
    st, _ = b.run()
    assert st.tuning.block_threads == t.block_threads and st.tuning.claim_items == t.claim_items
    for j, (scene, config) in enumerate(jobs[:2]):
        sums, counts = b.film(j)
        ref_sums, ref_counts, _ = oracle.OracleScene(scene).render(config)
        assert mean_relative_difference(sums, counts, ref_sums, ref_counts) <= 1e-3


def test_batch_argument_errors():
    b = Batch(0)
    try:
        with pytest.raises(RuntimeError):
            b.run()  # no jobs
        scene, config = scenes.cornell(width=16, height=16, spp=1)
        config.bsdf_samples = 0
        with pytest.raises(RuntimeError):
            b.add(scene, config)
        with pytest.raises(RuntimeError):
            b.set_tuning(96, 64, 3)
        textured, tconfig, _ = scenes.cornell_textured(width=16, height=16, spp=1, texture_size=16)
        with pytest.raises(RuntimeError):
            b.add(textured, tconfig)
    finally:
        b.close()


def test_batched_workload_of_64_jobs_through_the_graph():
    """The batched workload itself (64 seeded jobs of 10k-500k triangles, 64^2-1024^2 pixels, 1-64 spp) from the
    captured graph: every pixel of every job takes exactly its job's spp, and each job's film equals the oracle's over
    a 16 x 16 window of the image at the job's full spp (the same samples: sums to the FP64 order)."""
    from x2amd import abi
    jobs = scenes.batch_jobs()
    assert len(jobs) == 64
    b = Batch(0)
    try:
        for scene, config in jobs:
            b.add(scene, config)
        st, secs = b.run()
        assert st.graph == 1 and st.errors == 0 and st.jobs == 64
        assert st.paths == sum(c.width * c.height * c.spp for _, c in jobs)
        rng = np.random.default_rng(5)
        for j, (scene, config) in enumerate(jobs):
            sums, counts = b.film(j)
            assert (counts == config.spp).all(), j
            x0 = int(rng.integers(0, config.width - 16 + 1))
            y0 = int(rng.integers(0, config.height - 16 + 1))
            window = abi.Config.from_buffer_copy(config)
            window.viewport[:] = [x0, y0, x0 + 16, y0 + 16]
            ref_sums, ref_counts, k = oracle.OracleScene(scene).render(window)
            inside = ref_counts > 0
            assert k[2] == 0 and int(inside.sum()) == 256 and (ref_counts[inside] == config.spp).all(), j
            assert (counts[inside] == ref_counts[inside]).all(), j
            err = np.abs(sums[inside] - ref_sums[inside]).max() / max(1.0, float(np.abs(ref_sums[inside]).max()))
            assert err <= 1e-12, (j, err)
        print(f"64 jobs, {st.paths} paths, {st.device_seconds * 1e3:.1f} ms on the device; job p50 "
              f"{np.percentile(secs, 50) * 1e3:.1f} ms, p99 {np.percentile(secs, 99) * 1e3:.1f} ms")
        assert np.all(secs > 0) and np.all(secs <= st.device_seconds * 1.01 + 1e-4)
    finally:
        b.close()
