"""The oracle's medium pinned by what the mathematics fixes:

- trilinear density: voxel centres return the voxel, a linear field is reproduced, empty bricks read zero;
- ratio tracking: its mean over many sample identities is exp(-optical depth), the depth integrated independently
  along the ray (a homogeneous box, and a varying field integrated by quadrature of the density);
- Henyey-Greenstein: the phase integrates to one, its samples' mean cosine is g and their histogram its pdf;
- the march: transmittance exp(-sigma_t L) in a homogeneous medium, and the emission integral
  e (1 - exp(-sigma_t L)) / sigma_t;
- the whole estimator: an emitting, non-scattering box seen by the camera gives the same integral per pixel; an
  optically thin isotropic box under a constant environment gives single scattering albedo * E * (1 - exp(-tau)).
The medium model itself (box-bounded FP16 bricks, ratio tracking, the blackbody ramp) is the build's own, so these
checks, not the reference, pin it."""
import ctypes
import math

import numpy as np
import pytest

import oracle
from x2amd import scenes

f32 = ctypes.c_float


def fa(v):
    return (f32 * len(v))(*[float(x) for x in v])


def dummy_scene(volume, lights=(), materials=None):
    """A far, tiny triangle (the BVH wants one) and the medium."""
    pos = np.array([[500, -500, 500], [501, -500, 500], [500, -500, 501]], np.float32)
    mats = materials or [scenes.material((0.5, 0.5, 0.5))]
    return scenes.SceneData(pos, np.array([[0, 1, 2]]), [0, 1], [0], mats, list(lights), "medium", volume=volume)


def test_trilinear_density_closed_forms():
    n = 16
    k, j, i = np.mgrid[0:n, 0:n, 0:n]
    lin = ((i + 2 * j + 3 * k) / 64.0).astype(np.float32)  # exact in halves
    vol = scenes.VolumeData((0, 0, 0), (2, 2, 2), lin)
    o = oracle.OracleScene(dummy_scene(vol))
    lib = o.lib
    vox = 2.0 / n
    for (ii, jj, kk) in [(0, 0, 0), (3, 5, 7), (15, 15, 15), (8, 1, 12)]:
        p = [(ii + 0.5) * vox, (jj + 0.5) * vox, (kk + 0.5) * vox]
        assert lib.oracle_volume_density(o.h, fa(p)) == np.float32(lin[kk, jj, ii])
    rng = np.random.default_rng(1)
    for _ in range(300):
        q = rng.uniform(0.5, n - 1.5, 3)  # continuous voxel coordinates away from the edge
        p = (q + 0.5) * vox
        want = (q[0] + 2 * q[1] + 3 * q[2]) / 64.0
        assert abs(lib.oracle_volume_density(o.h, fa(p)) - want) < 2e-5


def test_empty_bricks_read_zero():
    n = 32
    d = np.zeros((n, n, n), np.float32)
    d[8:16, 8:16, 8:16] = 1.0  # one brick
    vol = scenes.VolumeData((0, 0, 0), (1, 1, 1), d)
    assert vol.occupancy() == 1 / 64
    o = oracle.OracleScene(dummy_scene(vol))
    vox = 1.0 / n
    at = lambda q: o.lib.oracle_volume_density(o.h, fa([(x + 0.5) * vox for x in q]))
    assert at((11, 11, 11)) == 1.0
    assert at((7.5, 11, 11)) == 0.5  # half way to the neighbour brick's first voxel (the apron)
    assert at((7, 11, 11)) == 0.0 and at((25, 25, 25)) == 0.0 and at((0, 0, 0)) == 0.0


def tracking_mean(o, config, org, dr, t0, t1, samples=4000):
    return np.mean([o.lib.oracle_volume_transmittance(o.h, ctypes.addressof(config), fa(org), fa(dr), t0, t1, 7, s, 3)
                    for s in range(samples)])


def test_ratio_tracking_homogeneous_box():
    d = np.full((16, 16, 16), 0.5, np.float32)
    vol = scenes.VolumeData((-1, -1, -1), (1, 1, 1), d, sigma_t_scale=1.0)
    o = oracle.OracleScene(dummy_scene(vol))
    _, cfg = scenes.cornell(width=8, height=8, spp=1)
    cfg.volume_transmittance_threshold = 0.0
    dr = np.array([1.0, 0.3, -0.2])
    dr /= np.linalg.norm(dr)
    # along the whole chord from outside: the slab in x is crossed over 2 / dx
    L = 2.0 / dr[0]
    mean = tracking_mean(o, cfg, (-3 * dr[0], -3 * dr[1], -3 * dr[2]), dr, 0.0, 100.0)
    want = math.exp(-0.5 * L)
    # majorant = density here: each estimate is 0 or 1, a binomial with p = want
    assert abs(mean - want) < 4 * math.sqrt(want * (1 - want) / 4000), (mean, want)
    # a stretch ending inside the box
    mean = tracking_mean(o, cfg, (0, 0, 0), dr, 0.2, 0.9)
    want = math.exp(-0.5 * 0.7)
    assert abs(mean - want) < 4 * math.sqrt(want * (1 - want) / 4000), (mean, want)


def test_ratio_tracking_matches_quadrature_of_a_varying_field():
    n = 32
    k, j, i = np.mgrid[0:n, 0:n, 0:n]
    d = (0.2 + 0.8 * np.abs(np.sin(i / 5.0) * np.cos(j / 7.0 + k / 9.0))).astype(np.float32)
    d[:, :, :8] = 0.0  # an empty brick column on the way
    vol = scenes.VolumeData((0, 0, 0), (2, 2, 2), d, sigma_t_scale=1.5)
    o = oracle.OracleScene(dummy_scene(vol))
    _, cfg = scenes.cornell(width=8, height=8, spp=1)
    cfg.volume_transmittance_threshold = 0.0
    org = np.array([-0.5, 0.31, 0.47])
    dr = np.array([1.0, 0.41, 0.29])
    dr /= np.linalg.norm(dr)
    ts = np.linspace(0.0, 6.0, 20001)
    dens = np.array([o.lib.oracle_volume_density(o.h, fa(org + dr * t)) for t in ts])
    tau = 1.5 * np.trapezoid(dens, ts)
    want = math.exp(-tau)
    est = np.array([o.lib.oracle_volume_transmittance(o.h, ctypes.addressof(cfg), fa(org), fa(dr), 0.0, 6.0, 1, s, 5)
                    for s in range(6000)])
    assert est.min() >= 0.0 and est.max() <= 1.0
    assert abs(est.mean() - want) < 4 * est.std() / math.sqrt(len(est)) + 1e-3, (est.mean(), want, tau)


@pytest.mark.parametrize("g", [0.0, 0.6, -0.3])
def test_henyey_greenstein(g):
    lib = oracle.load()
    ray = np.array([0.2, -0.5, 0.84])
    ray /= np.linalg.norm(ray)
    wo = -ray
    # normalisation: a midpoint rule over (cos theta, phi) about the ray
    nt, nphi = 400, 64
    total = 0.0
    frame_u = np.cross(ray, [1.0, 0, 0])
    frame_u /= np.linalg.norm(frame_u)
    frame_v = np.cross(ray, frame_u)
    for a in range(nt):
        ct = -1 + (a + 0.5) * 2 / nt
        st = math.sqrt(1 - ct * ct)
        for b in range(nphi):
            ph = (b + 0.5) * 2 * math.pi / nphi
            wi = ct * ray + st * (math.cos(ph) * frame_u + math.sin(ph) * frame_v)
            total += lib.oracle_phase_eval(g, fa(wo), fa(wi)) * (2 / nt) * (2 * math.pi / nphi)
    assert abs(total - 1.0) < 2e-3, total
    # sampling: mean cosine to the ray is g; the histogram of the cosine follows the pdf
    m = 64
    cos = []
    for a in range(m):
        for b in range(m):
            out = (f32 * 3)()
            lib.oracle_phase_sample(g, fa(wo), (a + 0.5) / m, (b + 0.5) / m, out)
            w = np.array(out[:])
            assert abs(np.linalg.norm(w) - 1) < 1e-5
            cos.append(w @ ray)
    cos = np.array(cos)
    assert abs(cos.mean() - g) < 5e-3, (cos.mean(), g)
    hist, edges = np.histogram(cos, bins=8, range=(-1, 1))
    pdf = lambda c: 0.5 * (1 - g * g) / (1 + g * g - 2 * g * c) ** 1.5
    for h, lo, hi in zip(hist, edges[:-1], edges[1:]):
        cs = np.linspace(lo, hi, 101)
        want = np.trapezoid(pdf(cs), cs) * len(cos)
        assert abs(h - want) <= 0.02 * len(cos), (lo, h, want)


def emissive_box(sigma, temperature=1500.0, n=16, albedo=(0.0, 0.0, 0.0)):
    d = np.ones((n, n, n), np.float32)
    t = np.full((n, n, n), temperature, np.float32)
    ramp = np.tile(np.array([[0.0, 0.0, 0.0]], np.float32), (2, 1))
    ramp[1] = (2.0, 1.0, 0.5)  # emission rgb at 1500 K; linear ramp below
    return scenes.VolumeData((-1, -1, -1), (1, 1, 1), d, t, sigma_t_scale=sigma, albedo=albedo, g=0.0,
                             emission_scale=1.0, ramp=ramp, ramp_max_kelvin=1500.0)


def test_march_transmittance_and_emission_integral():
    sigma = 0.8
    vol = emissive_box(sigma)
    o = oracle.OracleScene(dummy_scene(vol))
    _, cfg = scenes.cornell(width=8, height=8, spp=1)
    cfg.inv_volume_quality = 0.125  # steps of an eighth of a voxel
    out = (f32 * 8)()
    dr = [0.0, 0.0, 1.0]
    n = o.lib.oracle_volume_march(o.h, ctypes.addressof(cfg), fa([0.2, -0.1, -5]), fa(dr), 0.0, 100.0, 0, 9, out)
    # trilinear density is 1/2 on the box's faces and reaches 1 half a voxel in: the chord's optical length is its
    # length less a quarter voxel
    L = 2.0 - 0.125 / 4
    assert n > 100 and out[7] == 0.0
    assert abs(out[0] - math.exp(-sigma * L)) < 5e-4 * math.exp(-sigma * L)  # right-end rule over the face ramps
    assert abs(out[4] - 4.0) < 1e-5 and abs(out[5] - 6.0) < 1e-5  # the steps cover the chord
    want = (1 - math.exp(-sigma * L)) / sigma
    for c, e in zip(out[1:4], (2.0, 1.0, 0.5)):
        assert abs(c - e * want) < 0.02 * e * want, (c, e * want)


def test_estimator_sees_the_emission_integral():
    sigma = 0.8
    vol = emissive_box(sigma)
    scene = dummy_scene(vol)
    cfg = scenes.make_config(16, 16, 4, eye=(0.0, 0.0, 6.0), target=(0.0, 0.0, 0.0), fov=10.0, light_count=0)
    cfg.inv_volume_quality = 0.125
    sums, counts, k = oracle.OracleScene(scene).render(cfg)
    assert (counts == 4).all() and k[2] == 0
    rgb = sums[:, :3] / counts[:, None]
    # near the image centre the camera rays cross the box almost along z (optical length 2 less a quarter voxel)
    L = 2.0 - 0.125 / 4
    want = (1 - math.exp(-sigma * L)) / sigma
    centre = rgb.reshape(16, 16, 3)[7:9, 7:9].reshape(-1, 3)
    assert np.allclose(centre, np.array([2.0, 1.0, 0.5]) * want, rtol=0.03), centre.mean(axis=0)
    alpha = sums[:, 3] / counts
    assert np.allclose(alpha.reshape(16, 16)[7:9, 7:9], 1 - math.exp(-sigma * L), rtol=0.02)


def test_estimator_single_scattering_in_a_thin_medium():
    sigma, E = 0.02, 1.0
    n = 16
    vol = scenes.VolumeData((-1, -1, -1), (1, 1, 1), np.ones((n, n, n), np.float32), sigma_t_scale=sigma,
                            albedo=(1.0, 1.0, 1.0), g=0.0)
    scene = dummy_scene(vol, lights=[scenes.env_light((E, E, E))])
    cfg = scenes.make_config(16, 16, 16, eye=(0.0, 0.0, 6.0), target=(0.0, 0.0, 0.0), fov=10.0, light_count=1,
                             max_depth=4)
    cfg.max_volume_depth = 1
    sums, counts, k = oracle.OracleScene(scene).render(cfg)
    rgb = (sums[:, :3] / counts[:, None]).reshape(16, 16, 3)[6:10, 6:10].reshape(-1, 3)
    # single scattering of a constant environment through an almost transparent box: albedo E (1 - exp(-tau)),
    # less the shadow rays' own extinction (a mean chord of about one box half-length, tau/2 of it)
    tau = sigma * 2.0
    want = E * (1 - math.exp(-tau)) * (1 - 0.5 * tau)
    assert abs(rgb.mean() - want) < 0.05 * want, (rgb.mean(), want)
