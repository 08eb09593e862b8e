"""The oracle's texture filter and path-replay adjoint, pinned without the adjoint's own formulas:

- the bilinear tap against the reference's tex::bilinear (texture_core.h:144-153) on the same texels, bit for bit,
  and against closed forms (texel centres, the periodic seam);
- d ln f / d roughness of the GGX lobe against a central difference of the lobe's value (itself pinned bit-exactly
  to the reference's lobeEval);
- the whole adjoint (texel and roughness gradients of a linear image loss) against central differences of the
  oracle's own renders at the same sample identities. Russian roulette and clamping are off, so the fixed-seed
  estimator is a polynomial in the texels and the texel differences are exact up to rounding; the roughness
  differences move only the lobe's value (the oracle's evalRoughness hook) with sampling, pdfs and MIS weights held,
  which is the estimator the adjoint differentiates (sampling detached)."""
import ctypes

import numpy as np
import pytest

import oracle
import reflib
from x2amd import abi, scenes

P = ctypes.POINTER
f32 = ctypes.c_float


def fa(v):
    return (f32 * len(v))(*[float(x) for x in v])


def lookup(tex, u, v):
    t = abi.Texture()
    arr = np.ascontiguousarray(tex, np.float32)
    t.height, t.width = arr.shape[0], arr.shape[1]
    t.texels = arr.ctypes.data_as(P(f32))
    out = (f32 * 3)()
    ok = oracle.load().oracle_texture_lookup(ctypes.addressof(t), u, v, out)
    return (np.array(out[:], np.float32) if ok else None)


def test_texture_bilinear_matches_reference():
    ref = reflib.load()
    rng = np.random.default_rng(11)
    for (h, w) in [(37, 29), (64, 64), (5, 70)]:  # ragged 32 x 32 tiles on both axes
        tex = rng.uniform(-0.2, 1.5, (h, w, 3)).astype(np.float32)
        flat = np.ascontiguousarray(tex.reshape(-1))
        uv = np.concatenate([rng.uniform(-3, 4, (600, 2)), rng.uniform(0, 1, (600, 2)),
                             (rng.integers(-w, 2 * w, (200, 2)) + 0.5) / np.array([w, h])]).astype(np.float32)
        for u, v in uv:
            out = (f32 * 3)()
            assert ref.ref_bilinear(w, h, flat.ctypes.data_as(P(f32)), float(u), float(v), out) == 1
            mine = lookup(tex, float(u), float(v))
            assert mine is not None
            assert np.array_equal(mine.view(np.uint32), np.array(out[:], np.float32).view(np.uint32)), (h, w, u, v)


def test_texture_closed_forms():
    rng = np.random.default_rng(3)
    h, w = 16, 32
    tex = rng.uniform(0, 1, (h, w, 3)).astype(np.float32)
    for y in range(h):  # texel centres return the texel; t = 1 - v runs down the rows
        for x in range(w):
            got = lookup(tex, (x + 0.5) / w, 1 - (y + 0.5) / h)
            assert np.array_equal(got, tex[y, x])
    # u = 0 sits halfway between the last column and the first (periodic addressing)
    got = lookup(tex, 0.0, 1 - 0.5 / h)
    assert np.allclose(got, 0.5 * (tex[0, w - 1] + tex[0, 0]), rtol=1e-6)
    # a bilinear function of the texel coordinates is reproduced exactly away from the seam
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float32)
    lin = np.stack([0.25 * xx + 0.5 * yy, xx - yy, 0.125 * xx * yy], axis=-1).astype(np.float32)
    for _ in range(200):
        s, t = rng.uniform(0.5, w - 1.5), rng.uniform(0.5, h - 1.5)
        got = lookup(lin, float((s + 0.5) / w), float(1 - (t + 0.5) / h))
        fs, ft = np.float32(np.float32((s + 0.5) / w) * w - 0.5), np.float32((1 - np.float32(1 - (t + 0.5) / h)) * h - 0.5)
        want = np.array([0.25 * fs + 0.5 * ft, fs - ft, 0.125 * fs * ft])
        assert np.allclose(got, want, rtol=1e-4, atol=1e-4), (s, t, got, want)


def unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def test_ggx_roughness_log_derivative_is_the_derivative_of_the_lobe():
    lib = oracle.load()
    rng = np.random.default_rng(5)
    checked = 0
    for _ in range(400):
        N = unit(rng.normal(size=3))
        wo, wi = unit(rng.normal(size=3)), unit(rng.normal(size=3))
        if wo @ N < 0.05:
            wo = unit(wo - 2 * (wo @ N) * N + 0.05 * N)
        if wi @ N < 0.05:
            wi = unit(wi - 2 * (wi @ N) * N + 0.05 * N)
        r = float(rng.uniform(0.1, 0.95))
        h = 1e-3

        def lnf(rr):
            f, pdf = (f32 * 3)(), f32()
            lib.oracle_lobe_eval(1, fa(N), fa(N), fa(wo), fa(wi), rr, f, ctypes.byref(pdf))
            return np.log(np.float64(f[0]))

        num = (lnf(r + h) - lnf(r - h)) / (2 * h)
        ana = lib.oracle_ggx_roughness_log_derivative(r, fa(N), fa(wo), fa(wi))
        assert abs(ana - num) <= 2e-3 * max(1.0, abs(num)), (r, ana, num)
        checked += 1
    assert checked == 400
    # where alpha = r^2 is clamped to 0.001 the lobe does not depend on r
    assert lib.oracle_ggx_roughness_log_derivative(0.03, fa(N), fa(wo), fa(wi)) == 0.0


@pytest.fixture(scope="module")
def small_textured():
    scene, config, params = scenes.cornell_textured(width=32, height=32, spp=4, max_depth=8, texture_size=16)
    o = oracle.OracleScene(scene)
    rng = np.random.default_rng(21)
    d_image = rng.normal(size=(config.height * config.width, 3)).astype(np.float32)
    return scene, config, params, o, d_image


def linear_loss(o, config, d_image):
    sums, counts, k = o.render(config)
    assert k[2] == 0
    return float(np.sum((sums[:, :3] / counts[:, None]) * d_image.astype(np.float64)))


def test_oracle_texture_adjoint_matches_differences(small_textured):
    scene, config, params, o, d_image = small_textured
    tex = params["texture"]
    texture_floats = tex.size
    gt, gr, k = o.adjoint(config, d_image, texture_floats, scene.material_count)
    assert k[2] == 0 and np.abs(gt).sum() > 0
    rng = np.random.default_rng(2)
    picks = rng.choice(texture_floats, 48, replace=False)
    h = 1e-3
    fd = []
    for i in picks:
        t = tex.reshape(-1).copy()
        t[i] += h
        o.set_texture(0, t.reshape(tex.shape))
        lp = linear_loss(o, config, d_image)
        t[i] -= 2 * h
        o.set_texture(0, t.reshape(tex.shape))
        lm = linear_loss(o, config, d_image)
        fd.append((lp - lm) / (2 * h))
    o.set_texture(0, tex)
    fd, adj = np.array(fd), gt[picks]
    scale = np.abs(fd).max()
    print(f"texel gradients: max |fd| {scale:.3e}, max |adj - fd| {np.abs(adj - fd).max():.3e}")
    assert scale > 0
    assert np.linalg.norm(adj - fd) <= 1e-3 * np.linalg.norm(fd)
    assert np.all(np.abs(adj - fd) <= 1e-2 * np.abs(fd) + 1e-3 * scale)


def test_oracle_roughness_adjoint_matches_detached_differences(small_textured):
    scene, config, params, o, d_image = small_textured
    gt, gr, _ = o.adjoint(config, d_image, params["texture"].size, scene.material_count)
    h = 1e-3
    for m, r in zip(params["roughness_materials"], params["roughness"]):
        o.set_eval_roughness(m, float(r) + h)
        lp = linear_loss(o, config, d_image)
        o.set_eval_roughness(m, float(r) - h)
        lm = linear_loss(o, config, d_image)
        o.set_eval_roughness(m, -1.0)
        fd = (lp - lm) / (2 * h)
        print(f"material {m} roughness {r:.2f}: adjoint {gr[m]:.6e}, detached difference {fd:.6e}")
        assert abs(gr[m] - fd) <= 1e-2 * abs(fd) + 1e-9, (m, gr[m], fd)
    assert gr[0] == 0.0  # the textured wall is Lambert alone
