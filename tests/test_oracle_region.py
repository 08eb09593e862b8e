"""The oracle's glowing medium as a light source (the emissive region: volume_emission.h, async_integrator.h:478-566,
:999-1057, :1127-1143, :2481-2613), pinned by what the mathematics fixes:

- the region's support: every point where the medium can glow lies in a cell of positive probability (the emission
  at a point interpolates voxels up to one voxel away, so a histogram without the one-voxel margin fails this);
- its sampling: drawn places fall in each cell with the cell's probability, uniformly within the cell;
- its direction pdf: equal to a brute-force line integral of the 3D density (t^2 dt) along the ray, and integrating
  to one over the sphere of directions;
- the emissive integral: e (1 - exp(-sigma L)) / sigma along a ray inside a homogeneous glowing medium;
- the whole estimator: a Lambert ground lit only by a glowing, non-scattering cube gives (albedo / pi) times the
  volume integral of emission x transmittance x cos / r^2, worked out here by quadrature over the cube with a
  trilinear field of its own; light samples alone, BSDF samples alone and both under MIS agree with it.
The histogram (cells of 4^3 voxels, each the bound of its emission) is the build's own; the reference builds its
distribution from a MoonRay VDB histogram (cpu_scene.h:1634-1642) that this image cannot read."""
import ctypes
import math

import numpy as np
import pytest

import oracle
from x2amd import scenes

f32 = ctypes.c_float


def fa(v):
    return (f32 * len(v))(*[float(x) for x in v])


def dummy_scene(volume, lights=(), ground=False):
    if ground:  # a 40 x 40 Lambert ground at y = 0, albedo 0.5
        s = 20.0
        pos = np.array([[-s, 0, -s], [s, 0, -s], [s, 0, s], [-s, 0, s]], np.float32)
        idx = np.array([[0, 2, 1], [0, 3, 2]])
        return scenes.SceneData(pos, idx, [0, 2], [0], [scenes.material((0.5, 0.5, 0.5))], list(lights), "glow",
                                volume=volume)
    pos = np.array([[500, -500, 500], [501, -500, 500], [500, -500, 501]], np.float32)
    return scenes.SceneData(pos, np.array([[0, 1, 2]]), [0, 1], [0], [scenes.material((0.5, 0.5, 0.5))],
                            list(lights), "medium", volume=volume)


def region(o):
    res = (ctypes.c_int * 3)()
    inv = f32()
    m = (f32 * 12)()
    if not o.lib.oracle_region_info(o.h, res, ctypes.byref(inv), m):
        return None
    return np.array(res[:]), inv.value, np.array(m[:], np.float64)


def cell_pdfs(o, res):
    p = np.zeros((res[2], res[1], res[0]))
    for z in range(res[2]):
        for y in range(res[1]):
            for x in range(res[0]):
                p[z, y, x] = o.lib.oracle_region_cell_pdf(o.h, x, y, z)
    return p


def trilinear(grid, bmin, bmax, pts):
    """An independent trilinear interpolation of grid [z, y, x] (voxel centres, zero outside) at pts [n, 3]."""
    grid = np.asarray(grid, np.float64)
    res = np.array(grid.shape[::-1], np.float64)
    q = (pts - np.asarray(bmin)) / (np.asarray(bmax) - np.asarray(bmin)) * res - 0.5
    i0 = np.floor(q).astype(np.int64)
    f = q - i0
    padded = np.pad(grid, 1)
    out = np.zeros(len(pts))
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                w = (f[:, 0] if dx else 1 - f[:, 0]) * (f[:, 1] if dy else 1 - f[:, 1]) * (f[:, 2] if dz else 1 - f[:, 2])
                ix = np.clip(i0[:, 0] + dx + 1, 0, padded.shape[2] - 1)
                iy = np.clip(i0[:, 1] + dy + 1, 0, padded.shape[1] - 1)
                iz = np.clip(i0[:, 2] + dz + 1, 0, padded.shape[0] - 1)
                out += w * padded[iz, iy, ix]
    return out


def blob_volume(res=32, seed=5):
    """A smooth hot blob in a unit box: density and temperature both from one radial profile with noise."""
    rng = np.random.default_rng(seed)
    x = (np.arange(res) + 0.5) / res * 2 - 1
    r = np.sqrt(x[None, None, :] ** 2 + (x[None, :, None] * 1.3) ** 2 + x[:, None, None] ** 2)
    d = np.clip(1.0 - r, 0, 1) * (0.7 + 0.3 * rng.random((res, res, res)))
    t = 1500.0 * np.clip(1.2 - 1.5 * r, 0, 1)
    return scenes.VolumeData((-0.5, 0.0, -0.5), (0.5, 1.0, 0.5), d.astype(np.float32), t.astype(np.float32),
                             sigma_t_scale=2.0, albedo=(0.8, 0.8, 0.8), g=0.3, emission_scale=3.0,
                             ramp=scenes.blackbody_ramp(1500.0))


def test_no_region_without_glow():
    vol = blob_volume()
    vol.c.emission_scale = 0.0
    assert region(oracle.OracleScene(dummy_scene(vol))) is None
    d = np.zeros((16, 16, 16), np.float32)
    d[4:12, 4:12, 4:12] = 1
    assert region(oracle.OracleScene(dummy_scene(scenes.VolumeData((0, 0, 0), (1, 1, 1), d)))) is None


def test_region_support_covers_the_glow():
    vol = blob_volume()
    o = oracle.OracleScene(dummy_scene(vol))
    res, inv, m = region(o)
    assert tuple(res) == (8, 8, 8)
    assert inv == pytest.approx(512.0, rel=1e-6)  # 8^3 cells in a unit box
    pdf = cell_pdfs(o, res)
    assert pdf.sum() == pytest.approx(1.0, rel=1e-5)
    dens = vol.density.view(np.float16).astype(np.float64).reshape(32, 32, 32)
    temp = vol.temperature.view(np.float16).astype(np.float64).reshape(32, 32, 32)
    rng = np.random.default_rng(2)
    pts = rng.uniform((-0.5, 0.0, -0.5), (0.5, 1.0, 0.5), (200000, 3))
    ramp = scenes.blackbody_ramp(1500.0).astype(np.float64)
    lum = ramp @ np.array([0.299, 0.587, 0.114])
    kelvin = trilinear(temp, (-0.5, 0, -0.5), (0.5, 1, 0.5), pts)
    glow = trilinear(dens, (-0.5, 0, -0.5), (0.5, 1, 0.5), pts) * np.interp(kelvin, np.linspace(0, 1500, len(lum)), lum)
    cells = np.clip(((pts - (-0.5, 0.0, -0.5)) * 8).astype(int), 0, 7)
    lit = glow > 0
    assert lit.sum() > 1000
    assert (pdf[cells[lit, 2], cells[lit, 1], cells[lit, 0]] > 0).all()
    assert (pdf == 0).any()  # the box's empty corners stay out


def test_region_points_follow_the_cells():
    vol = blob_volume()
    o = oracle.OracleScene(dummy_scene(vol))
    res, inv, m = region(o)
    pdf = cell_pdfs(o, res)
    n = 120000
    rng = np.random.default_rng(3)
    u = rng.random((n, 3))
    out = (f32 * 3)()
    pts = np.empty((n, 3))
    for i in range(n):
        o.lib.oracle_region_point(o.h, fa(u[i]), out)
        pts[i] = out[:]
    q = (pts - (-0.5, 0.0, -0.5)) * 8
    cells = np.clip(np.floor(q).astype(int), 0, 7)
    counts = np.zeros((8, 8, 8))
    np.add.at(counts, (cells[:, 2], cells[:, 1], cells[:, 0]), 1)
    assert counts[pdf == 0].sum() == 0
    expect = pdf * n
    keep = expect > 20
    chi2 = (((counts - expect) ** 2)[keep] / expect[keep]).sum() / keep.sum()
    assert chi2 < 1.3, chi2  # about 1 for the cells' probabilities
    frac = q - np.floor(q)
    for a in range(3):  # uniform within the cells
        h, _ = np.histogram(frac[:, a], bins=10, range=(0, 1))
        assert np.abs(h / n - 0.1).max() < 0.006


def chord(p, w, bmin=(-0.5, 0.0, -0.5), bmax=(0.5, 1.0, 0.5)):
    """Where the ray p + t w (t >= 0) enters and leaves the box (slabs, in float64)."""
    with np.errstate(divide="ignore"):
        a = (np.asarray(bmin) - p) / w
        b = (np.asarray(bmax) - p) / w
    return max(0.0, float(np.minimum(a, b).max())), float(np.maximum(a, b).min())


def line_integral(pdf, inv, p, w, steps=400000):
    """The 3D pdf (cell probability x cells per unit volume) x t^2, summed at fine midpoints over the ray's chord."""
    ta, tb = chord(p, w)
    if ta >= tb:
        return 0.0
    t = ta + (np.arange(steps) + 0.5) / steps * (tb - ta)
    x = p[None, :] + t[:, None] * w[None, :]
    c = np.clip(np.floor((x - (-0.5, 0.0, -0.5)) * 8).astype(int), 0, 7)
    return float((pdf[c[:, 2], c[:, 1], c[:, 0]] * inv * t * t).sum() * (tb - ta) / steps)


def test_region_pdf_is_the_line_integral():
    vol = blob_volume()
    o = oracle.OracleScene(dummy_scene(vol))
    res, inv, m = region(o)
    pdf = cell_pdfs(o, res)
    rng = np.random.default_rng(4)
    tested = 0
    for i in range(60):
        inside = i % 3 == 0
        p = rng.uniform((-0.4, 0.1, -0.4), (0.4, 0.9, 0.4)) if inside else rng.uniform((-2, -1, -2), (2, 2.5, 2))
        if not inside and np.all((p > (-0.5, 0, -0.5)) & (p < (0.5, 1, 0.5))):
            continue
        target = rng.uniform((-0.3, 0.2, -0.3), (0.3, 0.8, 0.3))
        w = target - p
        w /= np.linalg.norm(w)
        tEnd = f32()
        got = o.lib.oracle_region_pdf(o.h, fa(p), fa(w), ctypes.byref(tEnd))
        want = line_integral(pdf, inv, p, w)
        assert got == pytest.approx(want, rel=2e-3), (p, w)
        assert tEnd.value == pytest.approx(chord(p, w)[1], rel=1e-5)  # where the ray leaves the box
        tested += 1
    assert tested > 30


def test_region_pdf_integrates_to_one_over_the_sphere():
    vol = blob_volume()
    o = oracle.OracleScene(dummy_scene(vol))
    tEnd = f32()
    for p in [(0.3, 1.6, 0.8), (0.05, 0.45, -0.1)]:  # outside and inside the box
        nu, nphi = 160, 320
        total = 0.0
        for i in range(nu):
            cz = 1 - 2 * (i + 0.5) / nu
            sz = math.sqrt(max(0.0, 1 - cz * cz))
            row = 0.0
            for j in range(nphi):
                phi = 2 * math.pi * (j + 0.5) / nphi
                w = (sz * math.cos(phi), cz, sz * math.sin(phi))
                row += o.lib.oracle_region_pdf(o.h, fa(p), fa(w), ctypes.byref(tEnd))
            total += row
        total *= 4 * math.pi / (nu * nphi)
        assert total == pytest.approx(1.0, abs=0.02), (p, total)


def test_emissive_integral_closed_form():
    n = 32
    d = np.full((n, n, n), 0.05, np.float32)
    t = np.full((n, n, n), 1500.0, np.float32)
    ramp = scenes.blackbody_ramp(1500.0)
    vol = scenes.VolumeData((0, 0, 0), (4, 4, 4), d, t, sigma_t_scale=1.0, emission_scale=2.0, ramp=ramp)
    o = oracle.OracleScene(dummy_scene(vol))
    c = scenes.make_config(8, 8, 1, eye=(0, 0, 10), target=(0, 0, 0))
    c.inv_volume_quality = 0.25  # steps of a quarter voxel: the march's Riemann sum within 1e-3 of the integral
    sigma = float(np.float16(0.05)) * 1.0
    e = 2.0 * float(np.float16(0.05)) * ramp[-1].astype(np.float64)
    out = (f32 * 3)()
    for L in (0.5, 1.7, 3.0):
        vals = []
        for b in range(64):  # the lattice offset is random: the march's mean over it
            o.lib.oracle_emissive_integral(o.h, ctypes.addressof(c), fa((0.5, 0.5, 0.5)), fa(np.ones(3) / np.sqrt(3)),
                                           0.0, L, 1, b + 1, out)
            vals.append(out[:])
        want = e * (1 - math.exp(-sigma * L)) / sigma
        np.testing.assert_allclose(np.mean(vals, axis=0), want, rtol=2e-3)


def glow_cube_scene(spp, bsdf_samples=1, light_samples=1):
    """A glowing, non-scattering cube beside a 5-degree camera looking straight down at the ground around the origin
    from y = 1: the medium's box is x in [1, 2], y in [0.5, 1.5], z in [-0.5, 0.5] at 32^3 voxels, its density 1 in
    the inner 24^3 voxels and 0 around them, its temperature 1500 K throughout (so the emission is the ramp's top
    entry times the density wherever the density is not zero)."""
    n = 32
    d = np.zeros((n, n, n), np.float32)
    d[4:28, 4:28, 4:28] = 1.0
    t = np.full((n, n, n), 1500.0, np.float32)
    vol = scenes.VolumeData((1.0, 0.5, -0.5), (2.0, 1.5, 0.5), d, t, sigma_t_scale=0.2, albedo=(0, 0, 0), g=0.0,
                            emission_scale=1.0, ramp=scenes.blackbody_ramp(1500.0))
    scene = dummy_scene(vol, ground=True)
    config = scenes.make_config(8, 8, spp, eye=(0, 1, 0), target=(0, 0, 0), up=(0, 0, 1), fov=5.0, light_count=0,
                                max_depth=1, bsdf_samples=bsdf_samples, light_samples=light_samples)
    config.inv_volume_quality = 0.25  # quarter-voxel march steps: the Riemann sum's bias below 1e-3
    return scene, config, d


def quadrature_radiance(density):
    """(albedo / pi) * integral over the box of emission x transmittance toward p x cos / r^2 (midpoints, two a
    voxel per axis; the optical depth by 16 midpoints along each segment), averaged over the image's 8 x 8 pixel
    centres on the ground (a symmetric grid, so the image's axis convention does not matter)."""
    bmin, bmax = np.array([1.0, 0.5, -0.5]), np.array([2.0, 1.5, 0.5])
    k = 64
    g = (np.arange(k) + 0.5) / k
    X = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3) * (bmax - bmin) + bmin
    dV = np.prod(bmax - bmin) / k ** 3
    rho = trilinear(density, bmin, bmax, X)
    X, rho = X[rho > 0], rho[rho > 0]
    emit = scenes.blackbody_ramp(1500.0).astype(np.float64)[-1]  # emission scale 1 x density x ramp(1500 K)
    half = math.tan(math.radians(2.5))
    offs = ((np.arange(8) + 0.5) / 8 * 2 - 1) * half
    total = np.zeros(3)
    for a in offs:
        for b in offs:
            p = np.array([a, 0.0, b])
            v = X - p
            r = np.linalg.norm(v, axis=1)
            cos = np.clip(v[:, 1] / r, 0, None)
            tau = np.zeros(len(X))
            for si in (np.arange(16) + 0.5) / 16:
                tau += trilinear(density, bmin, bmax, p + (X - p) * si) * 0.2 * r / 16
            total += (0.5 / math.pi) * (rho * np.exp(-tau) * cos / (r * r) * dV).sum() * emit
    return total / 64


@pytest.fixture(scope="module")
def quad():
    _, _, density = glow_cube_scene(1)
    return quadrature_radiance(density)


@pytest.mark.parametrize("bsdf_samples,light_samples", [(1, 1), (1, 4), (4, 1)])
def test_region_lighting_matches_quadrature(quad, bsdf_samples, light_samples):
    """MIS of region light samples and BSDF samples, at three splits between them (the power heuristic's counts
    enter the weights), against the quadrature."""
    scene, config, _ = glow_cube_scene(256, bsdf_samples, light_samples)
    sums, counts, k = oracle.OracleScene(scene).render(config)
    assert (counts == config.spp).all() and k[2] == 0
    img = (sums[:, :3] / counts[:, None]).mean(axis=0)
    print(f"bsdf {bsdf_samples} light {light_samples}: {img} against {quad}")
    np.testing.assert_allclose(img[:2], quad[:2], rtol=0.005)


def test_region_in_scattering_matches_quadrature():
    """The glowing medium lighting its own steps (single scattering of its emission, max volume depth 1): camera
    rays through the glowing cube's middle see its emission plus its in-scattered light; the latter, the difference
    against the same medium without scattering (its marches draw the same), equals the integral along the ray of
    sigma_s x transmittance x the light arriving over the sphere (the emission along each direction through its own
    extinction) x the phase function, worked out by quadrature in spherical coordinates about each point."""
    n = 32
    d = np.zeros((n, n, n), np.float32)
    d[4:28, 4:28, 4:28] = 1.0
    t = np.full((n, n, n), 1500.0, np.float32)
    bmin, bmax = np.array([-0.5, 0.5, -0.5]), np.array([0.5, 1.5, 0.5])
    eye, g, sigma, albedo = np.array([0.0, 1.0, 3.0]), 0.5, 0.6, 0.8

    def render(alb):  # 8 x 8 pixels of a 0.2-degree view: 64 rays a few thousandths apart
        vol = scenes.VolumeData(bmin, bmax, d, t, sigma_t_scale=sigma, albedo=(alb, alb, alb), g=g, emission_scale=1.0,
                                ramp=scenes.blackbody_ramp(1500.0))
        config = scenes.make_config(8, 8, 512, eye=tuple(eye), target=(0, 1, 0), fov=0.2, light_count=0, max_depth=1)
        config.max_volume_depth = 1
        config.inv_volume_quality = 0.25
        sums, counts, k = oracle.OracleScene(dummy_scene(vol)).render(config)
        assert (counts == 512).all() and k[2] == 0
        return (sums[:, :3] / counts[:, None]).mean(axis=0)

    scattered = render(albedo) - render(0.0)
    dens = d.astype(np.float64)
    dirc = np.array([0.0, 0.0, -1.0])
    nu, nphi, nr, m = 32, 64, 64, 32
    cu = 1 - 2 * (np.arange(nu) + 0.5) / nu
    CU, PH = np.meshgrid(cu, 2 * math.pi * (np.arange(nphi) + 0.5) / nphi, indexing="ij")
    SU = np.sqrt(1 - CU ** 2)
    W = np.stack([SU * np.cos(PH), CU, SU * np.sin(PH)], -1).reshape(-1, 3)
    dw = 4 * math.pi / (nu * nphi)
    cos = W @ dirc
    phase = (1 - g * g) / (4 * math.pi * (1 + g * g - 2 * g * cos) ** 1.5)
    dr = 1.8 / nr
    rr = (np.arange(nr) + 0.5) * dr
    ta, tb = 2.5, 3.5
    total = 0.0
    for ti in ta + (np.arange(m) + 0.5) / m * (tb - ta):
        x = eye + dirc * ti
        rx = trilinear(dens, bmin, bmax, x[None])[0]
        if rx == 0:
            continue
        s = ta + (np.arange(64) + 0.5) / 64 * (ti - ta)
        tau_cam = trilinear(dens, bmin, bmax, eye + dirc * s[:, None]).sum() * sigma * (ti - ta) / 64
        Y = x[None, None, :] + W[:, None, :] * rr[None, :, None]
        rho = trilinear(dens, bmin, bmax, Y.reshape(-1, 3)).reshape(len(W), nr)
        tau = (np.cumsum(rho, axis=1) - 0.5 * rho) * sigma * dr  # optical depth from x out to each radial midpoint
        arriving = (rho * np.exp(-tau)).sum(axis=1) * dr  # emission (per unit of the ramp's colour) along each direction
        total += albedo * sigma * rx * math.exp(-tau_cam) * (phase * arriving).sum() * dw * (tb - ta) / m
    want = total * scenes.blackbody_ramp(1500.0).astype(np.float64)[-1]
    print(f"in-scattered {scattered} against {want}")
    np.testing.assert_allclose(scattered[:2], want[:2], rtol=0.03)
