"""The medium on the device against the oracle, through the C-ABI: the cloud scene (brick-sparse FP16 density and
temperature, ratio-tracked shadows, checkpointed marches replayed for the light sampling, single and multiple
scattering, blackbody emission, the glowing medium lighting the ground and its own steps through its emission
region) at reduced image sizes. Counts must be exact and radiance within the stated mean
relative difference of 1e-3. The device's expf/logf/atan2f/tanf/powf are the GPU library's, the oracle's the host
libm's (both faithfully rounded), so unlike the surface-only path the per-sample radiance is not bit-exact; the ray
counts, which follow the samples' decisions, are compared within 1e-4."""
import numpy as np
import pytest

import oracle
from helpers import mean_relative_difference
from x2amd import scenes
from x2amd.lane import GpuLane

pytestmark = pytest.mark.gpu
TOL = 1e-3


def check(scene, config, label):
    lane = GpuLane(0)
    try:
        lane.prepare(scene)
        sums, counts, st = lane.render_frame(config)
    finally:
        lane.close()
    ref_sums, ref_counts, ref_k = oracle.OracleScene(scene).render(config)
    assert (counts == config.spp).all() and (ref_counts == config.spp).all()
    assert st.errors == 0 and ref_k[2] == 0
    rel = mean_relative_difference(sums, counts, ref_sums, ref_counts)
    exact = np.all(sums == ref_sums, axis=1).mean()
    dt = abs(int(st.rays_trace) - int(ref_k[0])) / max(1, ref_k[0])
    ds = abs(int(st.rays_shadow) - int(ref_k[1])) / max(1, ref_k[1])
    print(f"{label}: rel {rel:.3e}, bit-exact pixels {exact:.4f}, traces {st.rays_trace} vs {ref_k[0]}, "
          f"shadows {st.rays_shadow} vs {ref_k[1]}, {st.device_seconds * 1e3:.1f} ms device")
    assert rel <= TOL
    assert dt <= 1e-4 and ds <= 1e-4
    return sums, counts


def test_cloud_small_grid():
    scene, config = scenes.clouds(width=160, height=90, spp=4, res=64)
    check(scene, config, "clouds 64^3 160x90x4")


def test_cloud_full_grid():
    scene, config = scenes.clouds(width=128, height=72, spp=2, res=256)
    check(scene, config, "clouds 256^3 128x72x2")


def test_cloud_single_scattering_and_depth():
    scene, config = scenes.clouds(width=96, height=54, spp=4, res=64, max_volume_depth=1, max_depth=4)
    check(scene, config, "clouds single scattering 96x54x4")


def test_glowing_cube_lights_the_ground():
    """The region's light samples and BSDF samples at a surface (computeRadianceEmissiveRegionsBundled), each through
    the emissive integral: the ground beside a glowing, non-scattering cube."""
    from test_oracle_region import glow_cube_scene
    scene, config, _ = glow_cube_scene(64)
    sums, counts = check(scene, config, "glowing cube over the ground 8x8x64")
    assert (sums[:, 0] > 0).all()  # every ground pixel is lit by the cube


def test_glowing_cube_lights_its_own_steps():
    """The region's in-scattering at the camera ray's steps (volumeScatterRegions) and at multiple scattering's
    scatter points, in a scattering glowing cube seen through its middle."""
    n = 32
    d = np.zeros((n, n, n), np.float32)
    d[4:28, 4:28, 4:28] = 1.0
    t = np.full((n, n, n), 1500.0, np.float32)
    vol = scenes.VolumeData((-0.5, 0.5, -0.5), (0.5, 1.5, 0.5), d, t, sigma_t_scale=0.6, albedo=(0.8, 0.8, 0.8),
                            g=0.5, emission_scale=1.0, ramp=scenes.blackbody_ramp(1500.0))
    pos = np.array([[-20, 0, -20], [20, 0, -20], [20, 0, 20], [-20, 0, 20]], np.float32)
    scene = scenes.SceneData(pos, np.array([[0, 2, 1], [0, 3, 2]]), [0, 2], [0], [scenes.material((0.5, 0.5, 0.5))],
                             [scenes.env_light((0.1, 0.1, 0.1))], "glow", volume=vol)
    config = scenes.make_config(32, 32, 8, eye=(0, 1.2, 3), target=(0, 0.9, 0), fov=30.0, light_count=1, max_depth=4)
    config.max_volume_depth = 2
    check(scene, config, "scattering glowing cube 32x32x8")


def test_adjoint_refuses_a_medium():
    import ctypes
    scene, config = scenes.clouds(width=32, height=32, spp=1, res=64)
    lane = GpuLane(0)
    try:
        lane.prepare(scene)
        with pytest.raises(RuntimeError):
            lane.adjoint_frame(config, 16, 16, 16)  # refused before any pointer is read
    finally:
        lane.close()
