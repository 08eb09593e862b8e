"""The oracle's estimator pinned where its whole-path checks were open: each test compares an oracle render with a
value the test computes by other means, chosen so that a plausible slip (a dropped MIS weight, a lobe's pdf off by
the half-vector Jacobian, a phase function's sign, a wrong sample address) fails it.

- GGX inside the estimator: a GGX-only plane under a constant environment E gives E times the lobe's directional
  albedo at the view angle, the albedo integrated here by quadrature of the reference's own compiled lobe evaluation
  (tests/reflib.py ref_lobe_eval; shade_lobes_eval.h:93-146) over the half-vector hemisphere (Walter et al. 2007's
  GGX distribution as the change of variables, dw_i = 4 |w_o . h| dw_h).
- The 50/50 Lambert + GGX material: under E, E (albedo w_d + w_s A_ggx); under a rectangle light, a 2-D quadrature
  of f(w_o, w_i) Le cos(theta_l) / r^2 over the light (light_core.h:476-500), f again the reference's compiled lobes.
- The medium under a rectangle light: single scattering along the camera ray against a float64 quadrature of
  sigma_s T_camera HG(g) Le cos(theta_l) / r^2 T_shadow over the ray and the light (async_integrator.h:1152-1291).
- childBranch (async_integrator.h:281-286): the k-th output of SplitMix64 seeded with the parent, k = a * 0x10001 +
  b + 1, from a SplitMix64 generator written here from its published definition (Steele, Lea and Flood 2014), for
  parents and a other than zero.
- First-hit sample addresses (async_integrator.h:2642-2647, 2816-2821): a pixel's first-hit light samples, and its
  lobe samples, each fill every 1-D stratum exactly once in every power-of-two prefix of the sample numbers
  index * n + s, and the light and lobe sets are not the same points."""
import ctypes
import math

import numpy as np
import pytest

import oracle
import reflib
from x2amd import scenes

F3 = ctypes.c_float * 3
FLOAT = ctypes.c_float


def plane_scene(material, lights, size=1000.0):
    s = size
    pos = np.array([[-s, 0, -s], [s, 0, -s], [s, 0, s], [-s, 0, s]], np.float32)
    idx = np.array([[0, 2, 1], [0, 3, 2]])
    return scenes.SceneData(pos, idx, [0, 2], [0], [material], lights, "plane")


def view_config(theta_deg, spp, light_count, size=16, max_depth=2, **kw):
    """A camera 5 units from the origin at theta from the plane's normal (+y), looking at the origin through a
    narrow field, so that every pixel sees the plane at (nearly) the origin from (nearly) one direction."""
    th = math.radians(theta_deg)
    eye = (0.0, 5.0 * math.cos(th), 5.0 * math.sin(th))
    cfg = scenes.make_config(size, size, spp, eye=eye, target=(0.0, 0.0, 0.0), fov=0.02, light_count=light_count,
                             max_depth=max_depth, **kw)
    wo = np.array(eye) / np.linalg.norm(eye)
    return cfg, wo


def oracle_mean(scene, cfg):
    sums, counts, k = oracle.OracleScene(scene).render(cfg)
    assert (counts == cfg.spp).all() and k[2] == 0
    return (sums[:, :3] / counts[:, None]).mean(axis=0)


def ref_f(R, kind, wo, wi, roughness, N=(0.0, 1.0, 0.0)):
    f, pdf = F3(), FLOAT()
    R.ref_lobe_eval(kind, F3(*N), F3(*N), F3(*wo), F3(*wi), roughness, f, ctypes.byref(pdf))
    return np.array(f[:], np.float64)


def ggx_albedo(R, wo, roughness, nt=192, nphi=96):
    """The GGX lobe's directional albedo, integral of f(w_o, w_i) dw_i over the upper hemisphere (f carries the cosine),
    by a midpoint rule in the variables that map the unit square onto GGX half vectors (tan^2 theta_h = alpha^2 t /
    (1 - t), the density D(h) cos theta_h): the integrand becomes f 4 |w_o . h| / (D(h) cos theta_h), which is smooth."""
    alpha = max(0.001, roughness * roughness)
    a2 = alpha * alpha
    total = np.zeros(3)
    for i in range(nt):
        t = (i + 0.5) / nt
        tan2 = a2 * t / (1.0 - t)
        cos_h = 1.0 / math.sqrt(1.0 + tan2)
        sin_h = math.sqrt(max(0.0, 1.0 - cos_h * cos_h))
        D = a2 / (math.pi * cos_h ** 4 * (a2 + tan2) ** 2)
        for j in range(nphi):
            ph = 2.0 * math.pi * (j + 0.5) / nphi
            h = np.array([sin_h * math.cos(ph), cos_h, sin_h * math.sin(ph)])
            oh = float(wo @ h)
            if oh <= 0.0:
                continue
            wi = 2.0 * oh * h - wo
            if wi[1] <= 0.0:
                continue
            total += ref_f(R, 1, wo, wi / np.linalg.norm(wi), roughness) * 4.0 * oh / (D * cos_h)
    return total / (nt * nphi)


@pytest.mark.parametrize("roughness,theta", [(0.05, 35.0), (0.3, 60.0), (0.8, 45.0)])
def test_ggx_plane_under_environment_is_the_directional_albedo(roughness, theta):
    R = reflib.load()
    E = 0.7
    cfg, wo = view_config(theta, 256, 1, light_samples=2, bsdf_samples=2)
    mat = scenes.material((0.0, 0.0, 0.0), diffuse=0.0, specular=1.0, roughness=roughness)
    rgb = oracle_mean(plane_scene(mat, [scenes.env_light((E, E, E))]), cfg)
    A = ggx_albedo(R, wo, roughness)
    # the quadrature has converged: half the points in each direction change it by far less than the tolerance
    coarse = ggx_albedo(R, wo, roughness, nt=96, nphi=48)
    assert np.abs(coarse - A).max() < 1e-3 * A.max(), (coarse, A)
    want = E * A
    print(f"r {roughness} at {theta} deg: oracle {rgb} quadrature {want}")
    assert np.allclose(rgb, want, rtol=5e-3), (rgb, want)
    # a rough lobe loses energy to masking and shadowing (no energy compensation): the albedo is measurably below 1,
    # so a slip that dropped G would not pass; a smooth lobe at this angle keeps nearly all of it
    if roughness >= 0.3:
        assert want.max() < 0.98 * E


def test_lambert_ggx_under_environment():
    R = reflib.load()
    E, rho, wd, ws, r = 0.9, 0.6, 0.5, 0.5, 0.3
    cfg, wo = view_config(50.0, 256, 1)
    mat = scenes.material((rho, rho, rho), diffuse=wd, specular=ws, roughness=r)
    rgb = oracle_mean(plane_scene(mat, [scenes.env_light((E, E, E))]), cfg)
    want = E * (rho * wd + ws * ggx_albedo(R, wo, r))
    print(f"lambert + ggx under E: oracle {rgb} want {want}")
    assert np.allclose(rgb, want, rtol=5e-3), (rgb, want)


def rect_quadrature(R, wo, mat, center, u_axis, v_axis, Le, n=160):
    """Direct light from a one-sided rectangle (lighting the side u x v points to) at the origin of the y = 0 plane:
    the midpoint rule over the light of f(w_o, w_i) Le cos(theta_l) / r^2 dA, f = Lambert (albedo w_d, its cosine)
    + the GGX lobe scaled by w_s, both the reference's compiled evaluation."""
    c, u, v = (np.asarray(x, np.float64) for x in (center, u_axis, v_axis))
    nl = np.cross(u, v)
    nl /= np.linalg.norm(nl)
    area = 4.0 * np.linalg.norm(u) * np.linalg.norm(v)
    albedo, wd, ws, r = mat
    total = np.zeros(3)
    for i in range(n):
        for j in range(n):
            y = c + u * (2 * (i + 0.5) / n - 1) + v * (2 * (j + 0.5) / n - 1)
            d = y
            r2 = float(d @ d)
            wi = d / math.sqrt(r2)
            cos_l = float(-(wi @ nl))
            if cos_l <= 0.0 or wi[1] <= 0.0:
                continue
            f = np.asarray(albedo) * wd * ref_f(R, 0, wo, wi, 1.0) + ws * ref_f(R, 1, wo, wi, r)
            total += f * Le * cos_l / r2
    return total * area / (n * n)


def test_lambert_ggx_under_rect_light():
    R = reflib.load()
    rho, wd, ws, r, Le = (0.7, 0.5, 0.3), 0.5, 0.5, 0.3, 4.0
    center, u_axis, v_axis = (0.0, 1.6, -1.4), (0.6, 0.0, 0.0), (0.0, 0.3, 0.45)  # faces the origin from above
    light = scenes.rect_light(center, u_axis, v_axis, (Le, Le, Le))
    cfg, wo = view_config(55.0, 1024, 1, light_samples=2, bsdf_samples=2)
    mat = scenes.material(rho, diffuse=wd, specular=ws, roughness=r)
    rgb = oracle_mean(plane_scene(mat, [light]), cfg)
    want = rect_quadrature(R, wo, (rho, wd, ws, r), center, u_axis, v_axis, Le)
    coarse = rect_quadrature(R, wo, (rho, wd, ws, r), center, u_axis, v_axis, Le, n=80)
    assert np.abs(coarse - want).max() < 1e-3 * want.max()
    print(f"lambert + ggx under a rect light: oracle {rgb} quadrature {want}")
    assert np.allclose(rgb, want, rtol=5e-3), (rgb, want)
    # the GGX lobe's share is large here (the light sits near the mirror direction): the test sees that lobe
    lambert_only = rect_quadrature(R, wo, (rho, wd, 0.0, r), center, u_axis, v_axis, Le, n=40)
    assert (want - lambert_only).min() > 0.2 * want.min()


def test_medium_single_scattering_under_rect_light():
    """A homogeneous slab (y in [0, 0.25]) seen from above; a rectangle light below it facing up, so that the camera
    sees light scattered forward (g = 0.6). Single scattering only (max volume depth 1)."""
    sig, g, albedo, Le = 0.5, 0.6, 0.9, 20.0
    nx, ny = 64, 8
    dens = np.ones((nx, ny, nx), np.float32)  # [z, y, x]
    vol = scenes.VolumeData((-4.0, 0.0, -4.0), (4.0, 0.25, 4.0), dens, sigma_t_scale=sig, albedo=(albedo,) * 3, g=g)
    lc, lu, lv = np.array([0.0, -1.0, -1.3]), np.array([0.5, 0.0, 0.0]), np.array([0.0, 0.0, -0.4])
    assert np.cross(lu, lv)[1] > 0  # lights the +y side
    light = scenes.rect_light(tuple(lc), tuple(lu), tuple(lv), (Le, Le, Le))
    far = np.array([[500, -500, 500], [501, -500, 500], [500, -500, 501]], np.float32)  # the BVH's one triangle
    scene = scenes.SceneData(far, np.array([[0, 1, 2]]), [0, 1], [0], [scenes.material((0.5, 0.5, 0.5))], [light],
                             "slab", volume=vol)
    eye = np.array([0.0, 2.0, 2.2])
    target = np.array([0.0, 0.125, 0.0])
    cfg = scenes.make_config(32, 32, 64, eye=tuple(eye), target=tuple(target), fov=0.02, light_count=1, max_depth=2)
    cfg.max_volume_depth = 1
    cfg.inv_volume_quality = 0.25
    cfg.volume_transmittance_threshold = 0.0
    sums, counts, k = oracle.OracleScene(scene).render(cfg)
    assert (counts == cfg.spp).all() and k[2] == 0
    rgb = (sums[:, :3] / counts[:, None]).mean(axis=0)

    # float64 quadrature along the camera ray through the slab
    d = (target - eye) / np.linalg.norm(target - eye)
    t_in, t_out = (0.25 - eye[1]) / d[1], (0.0 - eye[1]) / d[1]
    nt = 400
    ts = t_in + (np.arange(nt) + 0.5) * (t_out - t_in) / nt
    dt = (t_out - t_in) / nt
    o = oracle.OracleScene(scene)
    fine = np.linspace(t_in - 0.05, t_out + 0.05, 4001)
    rho_fine = np.array([o.lib.oracle_volume_density(o.h, (FLOAT * 3)(*(eye + d * t))) for t in fine])
    tau_fine = sig * np.concatenate([[0.0], np.cumsum(0.5 * (rho_fine[1:] + rho_fine[:-1]) * np.diff(fine))])
    rho_t = np.array([o.lib.oracle_volume_density(o.h, (FLOAT * 3)(*(eye + d * t))) for t in ts])
    T_cam = np.exp(-np.interp(ts, fine, tau_fine))
    n = 48
    a = (np.arange(n) + 0.5) / n * 2 - 1
    ys = lc[None, None, :] + a[:, None, None] * lu + a[None, :, None] * lv
    ys = ys.reshape(-1, 3)
    dA = 4 * np.linalg.norm(lu) * np.linalg.norm(lv) / (n * n)
    nl = np.array([0.0, 1.0, 0.0])
    L = 0.0
    for t, rho, Tc in zip(ts, rho_t, T_cam):
        x = eye + d * t
        w = ys - x
        r2 = (w * w).sum(axis=1)
        wi = w / np.sqrt(r2)[:, None]
        cos_l = np.maximum(-(wi @ nl), 0.0)
        cos_s = wi @ d  # the scattering angle between the light's travel (-w_i) and the camera's (-d)
        phase = (1 - g * g) / (4 * math.pi * (1 + g * g - 2 * g * cos_s) ** 1.5)
        T_sh = np.exp(-sig * x[1] / np.abs(wi[:, 1]))  # the shadow ray leaves the slab through its floor
        L += albedo * sig * rho * Tc * (phase * Le * cos_l / r2 * T_sh).sum() * dA * dt
    print(f"slab single scattering: oracle {rgb} quadrature {L:.6f}")
    assert np.allclose(rgb, L, rtol=0.02), (rgb, L)
    # forward scattering matters here: the same light with g = -0.6 would give far less
    phase_back = lambda c: (1 - g * g) / (4 * math.pi * (1 + g * g + 2 * g * c) ** 1.5)
    assert phase_back(0.8) < 0.1 * ((1 - g * g) / (4 * math.pi * (1 + g * g - 2 * g * 0.8) ** 1.5))


class SplitMix64:
    """SplitMix64 (Steele, Lea and Flood, OOPSLA 2014; Vigna's splitmix64.c): state += 0x9e3779b97f4a7c15, then the
    output mix of the new state."""

    def __init__(self, seed):
        self.state = seed & 0xFFFFFFFFFFFFFFFF

    def next(self):
        m = 0xFFFFFFFFFFFFFFFF
        self.state = (self.state + 0x9E3779B97F4A7C15) & m
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
        return z ^ (z >> 31)


def test_splitmix64_generator_known_answers():
    """The generator above against SplitMix64's published first outputs from seeds 0 and 1234567."""
    g = SplitMix64(0)
    assert [g.next() for _ in range(3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    g = SplitMix64(1234567)
    assert [g.next() for _ in range(5)] == [6457827717110365317, 3203168211198807973, 9817491932198370423,
                                          4593380528125082431, 16408922859458223821]


@pytest.mark.parametrize("parent", [0x0123456789ABCDEF, 1234567, 0xFFFFFFFFFFFFFFFF, 0x8000000000000001])
def test_child_branch_is_the_kth_splitmix64_output(parent):
    lib = oracle.load()
    cases = [(0, 0), (0, 5), (1, 0), (1, 3), (2, 0xFFFF)]  # a = 1, 2 need k = 65538 ... 196608 steps
    steps = {a * 0x10001 + b + 1: (a, b) for a, b in cases}
    g = SplitMix64(parent)
    out = {}
    for k in range(1, max(steps) + 1):
        v = g.next()
        if k in steps:
            out[steps[k]] = v
    for (a, b), want in out.items():
        assert lib.oracle_child_branch(parent, a, b) == want, (hex(parent), a, b)
    # the device's tags: the light tree 0xfffe, a lobe's samples 0xffff, thinning 0xfff7, ... are a = tag, b = 0 or
    # the reverse; one of each from the generator (k = tag * 0x10001 + 1 steps would be too many to walk, so the
    # generator is jumped: its state after k steps is seed + k * gamma, which the definition above fixes)
    m = 0xFFFFFFFFFFFFFFFF
    for a, b in [(0xFFFE, 0), (3, 0xFFFF), (0xFFF7, 0), (0, 0xFFFB)]:
        k = a * 0x10001 + b + 1
        j = SplitMix64((parent + (k - 1) * 0x9E3779B97F4A7C15) & m)
        assert lib.oracle_child_branch(parent, a, b) == j.next()


def strata_once(values, m):
    cells = np.floor(np.asarray(values) * (1 << m)).astype(np.int64)
    return len(np.unique(cells)) == (1 << m) == len(cells)


def test_first_hit_sample_sets_are_stratified():
    """Two lights (a rectangle and the environment) with two samples each, a Lambert + GGX material with two samples a
    lobe: over sample indices 0 .. 127 a pixel draws 256 light samples per light and 256 per lobe, numbered
    index * 2 + s; every power-of-two prefix of each set fills each of its 2^m 1-D strata once in both coordinates."""
    rect = scenes.rect_light((0.0, 2.0, 0.0), (0.5, 0, 0), (0, 0, 0.5), (3.0, 3.0, 3.0))
    rect_up = scenes.rect_light((0.0, 2.0, 0.0), (0.5, 0, 0), (0, 0, -0.5), (3.0, 3.0, 3.0))
    light = rect if np.cross([0.5, 0, 0], [0, 0, 0.5])[1] < 0 else rect_up  # faces down onto the plane
    mat = scenes.material((0.5, 0.5, 0.5), diffuse=0.5, specular=0.5, roughness=0.4)
    scene = plane_scene(mat, [light, scenes.env_light((0.2, 0.2, 0.2))])
    cfg, _ = view_config(40.0, 1, 2, size=8, light_samples=2, bsdf_samples=2)
    o = oracle.OracleScene(scene)
    for pixel in (0, 27, 63):
        sets = {}
        for index in range(128):
            st, rows = o.first_hit_draws(cfg, pixel, index)
            assert st == 0
            for kind, which, number, r0, r1 in rows:
                sets.setdefault((int(kind), int(which)), {})[index * 2 + int(number)] = (r0, r1)
        assert set(sets) == {(0, 0), (0, 1), (1, 0), (1, 1)}, set(sets)  # both lights, both lobes
        for key, pts in sets.items():
            assert sorted(pts) == list(range(256)), key
            xy = np.array([pts[j] for j in range(256)])
            for m in range(1, 9):
                for c in (0, 1):
                    assert strata_once(xy[: 1 << m, c], m), (pixel, key, m, c)
        # the light samples are stratified apart from the lobe samples: different points
        assert not np.allclose(np.array([sets[(0, 0)][j] for j in range(256)]),
                               np.array([sets[(1, 0)][j] for j in range(256)]))
