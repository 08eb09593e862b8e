"""The prototype estimator on the device (directKernel) against the oracle: bit-exact films at 1 spp and identical
ray counts, on the headline icosphere scene at reduced size and on the Cornell box with surface limit 4."""
import numpy as np
import pytest

import oracle
from x2amd import scenes
from x2amd.lane import GpuLane

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["icospheres", "cornell"])
def test_direct_bit_exact(name):
    if name == "icospheres":
        scene, config = scenes.icospheres(width=160, height=90, spp=1)
        config.surface_limit = 1
    else:
        scene, config = scenes.cornell(width=64, height=64, spp=1)
        config.surface_limit = 4
    config.direct_only = 1
    config.sky[:] = [0.2, 0.25, 0.3]
    lane = GpuLane(0)
    try:
        lane.prepare(scene)
        sums, counts, st = lane.render_frame(config, samples_per_chunk=1)
    finally:
        lane.close()
    ref_sums, ref_counts, ref_k = oracle.OracleScene(scene).render(config)
    assert (counts == 1).all() and (ref_counts == 1).all() and st.errors == 0 and ref_k[2] == 0
    exact = np.all(sums == ref_sums, axis=1).mean()
    print(f"{name} direct: bit-exact pixels {exact:.5f}, traces {st.rays_trace} vs {ref_k[0]}, shadows "
          f"{st.rays_shadow} vs {ref_k[1]}")
    assert exact == 1.0
    assert st.rays_trace == ref_k[0] and st.rays_shadow == ref_k[1]
