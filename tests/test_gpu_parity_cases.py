"""Device vs oracle on the estimator's less travelled branches, at 1 spp where the device's film must equal the
oracle's bit for bit (per-sample radiance is computed by the same float expressions in the same order), with
identical ray counts: ragged image sizes and a partial viewport, several light and BSDF samples (stratified
first-hit sets), smart clamping, Russian roulette from the first bounce and none at all, hemispherical and
two-sided/reverse-sided lights, emissive and GGX-only materials, direct lighting only. Every case runs twice: with
the surface estimator as one kernel, and in the wavefront form (shading and traversal in separate kernels)."""
import math

import numpy as np
import pytest

import oracle
from x2amd import abi, scenes
from x2amd.lane import GpuLane

pytestmark = pytest.mark.gpu
FORM = {"wave_slots": 0}


@pytest.fixture(autouse=True, params=[0, 512], ids=["one_kernel", "wavefront"])
def form(request):
    """Runs each case in both forms of the surface estimator (wave_slots 0: one kernel)."""
    FORM["wave_slots"] = request.param
    yield request.param
    FORM["wave_slots"] = 0


def render_both(scene, config):
    lane = GpuLane(0, wave_slots=FORM["wave_slots"])
    try:
        lane.prepare(scene)
        sums, counts, st = lane.render_frame(config, samples_per_chunk=1)
    finally:
        lane.close()
    ref_sums, ref_counts, ref_k = oracle.OracleScene(scene).render(config)
    return sums, counts, st, ref_sums, ref_counts, ref_k


def assert_parity(scene, config, label):
    sums, counts, st, ref_sums, ref_counts, ref_k = render_both(scene, config)
    vp = config.viewport
    y, x = np.divmod(np.arange(config.width * config.height), config.width)
    inside = (x >= vp[0]) & (y >= vp[1]) & (x < vp[2]) & (y < vp[3])
    assert (counts[inside] == config.spp).all() and (counts[~inside] == 0).all(), label
    assert (ref_counts == counts).all(), label
    assert st.errors == 0 and ref_k[2] == 0, label
    exact = np.all(sums[inside] == ref_sums[inside], axis=1).mean()
    print(f"{label}: bit-exact pixels {exact:.5f}, traces {st.rays_trace} vs {ref_k[0]}, shadows {st.rays_shadow} "
          f"vs {ref_k[1]}")
    assert exact == 1.0, label  # one sample a pixel: no FP64 atomic order to differ
    assert st.rays_trace == ref_k[0] and st.rays_shadow == ref_k[1], label


def hemisphere_env(radiance):
    L = scenes.env_light(radiance)
    L.hemispherical = 1
    L.inv_area = float(np.float32(1.0 / (2.0 * math.pi)))
    L.local_to_render[:] = [1.0, 0.0, 0.0, 0.0, 0.0, -1.0, 0.0, 1.0, 0.0]  # local +z = render +y
    L.render_to_local[:] = [1.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0, -1.0, 0.0]
    L.render_to_local_rot[:] = L.render_to_local[:]
    return L


def test_ragged_image_and_viewport():
    scene, config = scenes.cornell(width=37, height=23, spp=1)
    config.viewport[:] = [3, 2, 31, 21]
    assert_parity(scene, config, "cornell 37x23, viewport [3,2)-(31,21)")


def test_icospheres_two_light_samples():
    """Four lights with two samples each make up to ten shadow rays a vertex: more than the wave's shadow pool holds
    for a vertex, so the frame runs with every shadow ray traced where it is made (the pooled kernel renders the cases
    with at most six rays a vertex and two lights or more; the Cornell box's one light is traced where it is made)."""
    scene, config = scenes.icospheres(width=64, height=36, spp=1, count=4, level=3)
    config.light_samples = 2
    assert_parity(scene, config, "icospheres light_samples 2")


def test_several_light_and_bsdf_samples():
    scene, config = scenes.cornell(width=48, height=40, spp=1)
    config.light_samples, config.bsdf_samples = 3, 2
    assert_parity(scene, config, "cornell light_samples 3, bsdf_samples 2")


def test_clamp_and_roulette_settings():
    scene, config = scenes.cornell(width=48, height=40, spp=1, rr_min_depth=1)
    config.clamp_value, config.clamp_depth = 0.5, 1
    assert_parity(scene, config, "cornell clamp 0.5 from depth 1, RR from depth 1")
    scene, config = scenes.cornell(width=48, height=40, spp=1)
    config.rr_threshold = 0.0
    assert_parity(scene, config, "cornell RR off")


def test_direct_lighting_only():
    scene, config = scenes.cornell(width=48, height=40, spp=1)
    config.max_depth = 0
    assert_parity(scene, config, "cornell max depth 0")
    scene, config = scenes.icospheres(width=64, height=36, spp=1, count=4, level=3)
    config.max_depth = 1
    assert_parity(scene, config, "icospheres max depth 1")


def test_light_kinds_and_materials():
    s = 20.0
    pos = np.array([[-s, 0, -s], [s, 0, -s], [s, 0, s], [-s, 0, s],
                    [-1, 0.5, -1], [1, 0.5, -1], [1, 0.5, 1], [-1, 0.5, 1],
                    [-0.5, 2, 0], [0.5, 2, 0], [0, 2.5, 0.3]], np.float32)
    idx = np.array([[0, 2, 1], [0, 3, 2], [4, 6, 5], [4, 7, 6], [8, 9, 10]])
    mats = [scenes.material((0.6, 0.6, 0.6)),
            scenes.material((0.0, 0.0, 0.0), diffuse=0.0, specular=1.0, roughness=0.35),  # GGX only
            scenes.material((0.2, 0.3, 0.4), emission=(2.0, 1.0, 0.5))]  # emissive
    lights = [hemisphere_env((0.3, 0.35, 0.4)),
              scenes.rect_light((2.0, 3.0, 0.0), (0.5, 0, 0), (0, 0, 0.5), (6.0, 6.0, 6.0), sidedness=abi.X2_SIDE_TWO),
              scenes.rect_light((-2.0, 3.0, 0.0), (0.5, 0, 0), (0, 0, -0.5), (4.0, 5.0, 6.0),
                                sidedness=abi.X2_SIDE_REVERSE)]
    scene = scenes.SceneData(pos, idx, [0, 2, 4, 5], [0, 1, 2], mats, lights, "lights")
    config = scenes.make_config(64, 48, 1, eye=(0.0, 2.0, 6.0), target=(0.0, 0.8, 0.0), light_count=3, max_depth=6)
    assert_parity(scene, config, "hemispherical env, two-sided and reverse rect lights, GGX-only, emissive")


def test_single_sided_meshes():
    """Back-face culling: the Cornell box with every mesh single-sided (the walls face into the box, the boxes out)
    and a single-sided screen between the camera and the box facing into it, so that camera rays pass through its
    back while rays leaving the box meet its front; the walls' outsides are culled too."""
    base, config = scenes.cornell(width=40, height=32, spp=1)
    screen = np.array([[-0.6, -0.6, 1.5], [-0.6, 0.6, 1.5], [0.6, 0.6, 1.5], [0.6, -0.6, 1.5]], np.float32)
    # (b - a) x (c - a) = (0, 1.2, 0) x (1.2, 1.2, 0) points to -z: into the box, away from the camera
    n = base.positions.shape[0]
    positions = np.concatenate([base.positions, screen])
    indices = np.concatenate([base.indices, np.array([[n, n + 1, n + 2], [n, n + 2, n + 3]])])
    first = np.concatenate([base.geom_first, [base.geom_first[-1] + 2]])
    mats = np.concatenate([base.geom_material, [0]])
    geoms = len(first) - 1
    materials = [base.materials[i] for i in range(base.material_count)]
    lights = [base.lights[i] for i in range(base.light_count)]
    scene = scenes.SceneData(positions, indices, first, mats, materials, lights, "cornell single-sided",
                             single_sided=range(geoms))
    assert_parity(scene, config, "cornell single-sided with a screen")
    # the screen is seen through from the camera: the camera's first hits are the box's, as without the screen
    two = scenes.SceneData(positions, indices, first, mats, materials, lights, "cornell two-sided")
    a = oracle.OracleScene(scene).render(config)[0]
    b = oracle.OracleScene(two).render(config)[0]
    assert not np.array_equal(a, b)
