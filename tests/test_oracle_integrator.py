"""The oracle's whole estimator against what the mathematics fixes: a white furnace (a Lambert plane under a
constant environment gives albedo * radiance), the irradiance of a rectangle light (the closed-form point-to-
rectangle form factor), pure emission, and the film's exact counts and determinism."""
import math

import numpy as np

import oracle
from x2amd import scenes


def plane_scene(albedo, lights, size=1000.0, emission=(0.0, 0.0, 0.0), diffuse=1.0):
    s = size
    pos = np.array([[-s, 0, -s], [s, 0, -s], [s, 0, s], [-s, 0, s]], np.float32)
    idx = np.array([[0, 2, 1], [0, 3, 2]])
    mats = [scenes.material(albedo, diffuse=diffuse, emission=emission)]
    return scenes.SceneData(pos, idx, [0, 2], [0], mats, lights, "plane")


def mean_radiance(scene, cfg):
    sums, counts, k = oracle.OracleScene(scene).render(cfg)
    assert (counts == cfg.spp).all() and k[2] == 0
    return (sums[:, :3] / counts[:, None]).mean(axis=0), sums[:, 3] / counts


def test_white_furnace_lambert():
    rho, E = 0.5, 1.0
    scene = plane_scene((rho, rho, rho), [scenes.env_light((E, E, E))])
    cfg = scenes.make_config(64, 64, 16, eye=(0.0, 5.0, 0.0), target=(0.0, 0.0, 0.0), up=(0, 0, -1), fov=60,
                             light_count=1, max_depth=4)
    rgb, alpha = mean_radiance(scene, cfg)
    assert np.allclose(rgb, rho * E, rtol=5e-3), rgb
    assert (alpha == 1.0).all()


def form_factor_corner(x, y):
    """Differential area to a parallel rectangle, the normal through one corner (a standard closed form)."""
    a, b = x / math.sqrt(1 + x * x), y / math.sqrt(1 + y * y)
    return (a * math.atan(y / math.sqrt(1 + x * x)) + b * math.atan(x / math.sqrt(1 + y * y))) / (2 * math.pi)


def test_rect_light_irradiance():
    rho, Le, h, hw, hh = 0.8, 3.0, 1.5, 0.6, 0.4
    light = scenes.rect_light((0.0, h, 0.0), (hw, 0.0, 0.0), (0.0, 0.0, hh), (Le, Le, Le))  # faces down
    scene = plane_scene((rho, rho, rho), [light])
    cfg = scenes.make_config(16, 16, 256, eye=(0.0, 4.0, 0.0), target=(0.0, 0.0, 0.0), up=(0, 0, -1), fov=0.02,
                             light_count=1, max_depth=3)
    rgb, _ = mean_radiance(scene, cfg)
    expected = rho * Le * 4 * form_factor_corner(hw / h, hh / h)
    assert np.allclose(rgb, expected, rtol=5e-3), (rgb, expected)


def test_emission_exact():
    scene = plane_scene((0.0, 0.0, 0.0), [], emission=(3.0, 2.0, 1.0), diffuse=0.0)
    cfg = scenes.make_config(8, 8, 4, eye=(0.0, 5.0, 0.0), target=(0.0, 0.0, 0.0), up=(0, 0, -1), fov=30,
                             light_count=0)
    sums, counts, _ = oracle.OracleScene(scene).render(cfg)
    assert (sums / counts[:, None] == np.array([3.0, 2.0, 1.0, 1.0])).all()


def test_film_deterministic_and_chunked():
    scene, cfg = scenes.cornell(width=40, height=24, spp=6)
    o = oracle.OracleScene(scene)
    a = o.render(cfg, threads=1)
    b = o.render(cfg, threads=8)
    assert (a[0] == b[0]).all() and (a[1] == b[1]).all() and a[2] == b[2]
    chunks = scenes.uniform_chunks(cfg, samples_per_chunk=4)
    c = o.render(cfg, chunks=chunks)
    assert (c[1] == cfg.spp).all()
    assert np.allclose(c[0], a[0], rtol=1e-12, atol=1e-12)  # the same samples in another grouping
    part = chunks[: len(chunks) // 2]
    d = o.render(cfg, chunks=part)
    expect = np.zeros(cfg.width * cfg.height, np.uint32)
    tiles_x = (cfg.width + 7) // 8
    for ch in part:
        for p in range(64):
            x, y = (ch.tile % tiles_x) * 8 + p % 8, (ch.tile // tiles_x) * 8 + p // 8
            if x < cfg.width and y < cfg.height:
                expect[y * cfg.width + x] += ch.count
    assert (d[1] == expect).all()


def test_sample_identity():
    """A sample is a function of (pixel, index, seed): the same on any call, different for another index or seed."""
    scene, cfg = scenes.cornell(width=32, height=32, spp=1)
    o = oracle.OracleScene(scene)
    a = o.sample(cfg, 500, 3)
    assert a == o.sample(cfg, 500, 3) and a[0] == 0
    assert a[1] != o.sample(cfg, 500, 4)[1]
    cfg.seed = scenes.frame_seed(1)
    assert a[1] != o.sample(cfg, 500, 3)[1]
