"""The prototype estimator (integrateDirect, async_integrator.h:61-126) in the oracle: its Lambert sample bit-exact
against the reference's surface::lambert compiled in place, and the estimator against closed forms (a Lambert plane
under a constant sky returns albedo x sky for any surface limit; an emitter shows its emission)."""
import ctypes

import numpy as np

import oracle
import reflib
from x2amd import scenes

f32 = ctypes.c_float


def fa(v):
    return (f32 * len(v))(*[float(x) for x in v])


def test_lambert_sample_matches_reference():
    ref = reflib.load()
    lib = oracle.load()
    rng = np.random.default_rng(4)
    for _ in range(3000):
        n = rng.normal(size=3)
        n = (n / np.linalg.norm(n)).astype(np.float32)
        albedo = rng.uniform(0, 1, 3)
        u, v = float(rng.random()), float(rng.random())
        outs = []
        for f in (ref.ref_surface_lambert, lib.oracle_surface_lambert):
            wi, fv, pdf = (f32 * 3)(), (f32 * 3)(), f32()
            ok = f(fa(albedo), fa(n), u, v, wi, fv, ctypes.byref(pdf))
            outs.append((ok, np.array(wi[:], np.float32), np.array(fv[:], np.float32), np.float32(pdf.value)))
        (a_ok, a_wi, a_f, a_p), (b_ok, b_wi, b_f, b_p) = outs
        assert a_ok == b_ok
        if a_ok:
            assert np.array_equal(a_wi.view(np.uint32), b_wi.view(np.uint32))
            assert np.array_equal(a_f.view(np.uint32), b_f.view(np.uint32)) and a_p.view(np.uint32) == b_p.view(np.uint32)
    # invalid inputs are refused alike
    for albedo, n, u, v in [((-0.1, 0.5, 0.5), (0, 0, 1), 0.5, 0.5), ((0.5, 0.5, 0.5), (0, 0, 2), 0.5, 0.5),
                            ((0.5, 0.5, 0.5), (0, 0, 1), 1.5, 0.5)]:
        wi, fv, pdf = (f32 * 3)(), (f32 * 3)(), f32()
        assert ref.ref_surface_lambert(fa(albedo), fa(n), u, v, wi, fv, ctypes.byref(pdf)) == 0
        assert lib.oracle_surface_lambert(fa(albedo), fa(n), u, v, wi, fv, ctypes.byref(pdf)) == 0


def plane(albedo, emission=(0.0, 0.0, 0.0)):
    s = 1000.0
    pos = np.array([[-s, 0, -s], [s, 0, -s], [s, 0, s], [-s, 0, s]], np.float32)
    mats = [scenes.material(albedo, emission=emission)]
    return scenes.SceneData(pos, np.array([[0, 2, 1], [0, 3, 2]]), [0, 2], [0], mats, [], "plane")


def direct_config(limit, sky):
    cfg = scenes.make_config(32, 32, 4, eye=(0.0, 5.0, 0.0), target=(0.0, 0.0, 0.0), up=(0, 0, -1), fov=60,
                             light_count=0)
    cfg.direct_only, cfg.surface_limit = 1, limit
    cfg.sky[:] = list(sky)
    return cfg


def test_plane_under_sky_is_albedo_times_sky():
    rho, sky = (0.3, 0.5, 0.7), (0.8, 1.0, 1.2)
    for limit in (1, 2, 4):
        sums, counts, k = oracle.OracleScene(plane(rho)).render(direct_config(limit, sky))
        assert k[2] == 0 and (counts == 4).all()
        rgb = sums[:, :3] / counts[:, None]
        assert np.allclose(rgb, np.array(rho) * np.array(sky), rtol=1e-6), (limit, rgb.mean(axis=0))
        assert (sums[:, 3] / counts == 1.0).all()  # every camera ray hits the plane


def test_emission_is_seen():
    sums, counts, _ = oracle.OracleScene(plane((0.0, 0.0, 0.0), emission=(2.0, 1.0, 0.5))).render(
        direct_config(1, (0.0, 0.0, 0.0)))
    rgb = sums[:, :3] / counts[:, None]
    assert np.allclose(rgb, [2.0, 1.0, 0.5])
