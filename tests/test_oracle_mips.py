"""The oracle's mip pyramids, map lookups and ray differentials, pinned:

- against the reference's own functions compiled in place (oracle/_ref/libref.so): tex::lookup over the same levels
  in all three path modes (texture_core.h:161-206), tex::footprint and shade::hitTexDerivatives, the transfer of
  differentials and the texture derivatives they give, trianglePartials, the cosine lobe's differentials and
  scatterDifferentials for Lambert and GGX lobes (shade_material.h, shade_math.h). Bit for bit where the two use the
  same arithmetic; where the oracle takes the shared sine/cosine polynomial and an FP64 square root in place of libm's
  sincosf and hypot, within the rounding those leave (and bit for bit on the taps that do not read theta);
- against what the mathematics fixes: a constant texture reads as that constant at every footprint, a footprint
  wider than the texture reads its mean (the 1 x 1 level), a power-of-two pyramid's level k is the mean of level 0
  over 2^k x 2^k blocks, the half-float pyramid is the float one rounded level by level, and a camera looking
  straight at a textured plane sees a footprint of one pixel's width in texture units, scaled by 1/min(sqrt(spp), 2).
"""
import ctypes

import numpy as np
import pytest

import oracle
import reflib

P = ctypes.POINTER
f32 = ctypes.c_float
u32 = ctypes.c_uint32


def fa(v):
    return (f32 * len(v))(*[float(x) for x in v])


def out(n):
    return (f32 * n)()


def arr(o):
    return np.array(o[:], np.float32)


def ref_lookup(levels, mode, fp, swap=1):
    sizes = (u32 * (2 * len(levels)))(*[v for l in levels for v in (l.shape[1], l.shape[0])])
    flat = np.ascontiguousarray(np.concatenate([l.reshape(-1) for l in levels]), np.float32)
    o = out(3)
    ok = reflib.load().ref_map_lookup(len(levels), sizes, flat.ctypes.data_as(P(f32)), mode, swap, fa(fp), o)
    return arr(o) if ok else None


def test_pyramid_sizes_and_box_means():
    rng = np.random.default_rng(5)
    tex = rng.uniform(0, 1, (64, 128, 3)).astype(np.float32)
    levels = oracle.pyramid(tex)
    assert [l.shape[:2] for l in levels] == [(64 >> k if 64 >> k else 1, 128 >> k) for k in range(8)]
    for k, l in enumerate(levels):
        b = 1 << k
        h, w = l.shape[:2]
        if 64 % b:  # past the height's last halving the rows clamp
            continue
        want = tex.astype(np.float64).reshape(h, b, w, b, 3).mean(axis=(1, 3))
        assert np.allclose(l, want, rtol=0, atol=2e-6 * k + 1e-7), k
    # ragged sides: max(1, n >> 1) a level, down to 1 x 1
    for h, w in [(37, 29), (1, 70), (5, 3)]:
        sizes = [l.shape[:2] for l in oracle.pyramid(rng.uniform(0, 1, (h, w, 3)).astype(np.float32))]
        assert sizes[-1] == (1, 1)
        for (h0, w0), (h1, w1) in zip(sizes, sizes[1:]):
            assert (h1, w1) == (max(1, h0 >> 1), max(1, w0 >> 1))


def test_half_pyramid_is_the_float_pyramid_rounded():
    rng = np.random.default_rng(6)
    tex = rng.uniform(-0.1, 3.0, (33, 40, 3)).astype(np.float32)
    for full, half in zip(oracle.pyramid(tex), oracle.pyramid(tex, half=True)):
        assert np.array_equal(half, full.astype(np.float16).astype(np.float32))


def test_constant_texture_reads_its_value_everywhere():
    tex = np.broadcast_to(np.array([0.25, 0.5, 0.75], np.float32), (48, 80, 3)).copy()
    rng = np.random.default_rng(7)
    for mode in (0, 1, 2):
        for _ in range(300):
            fp = [rng.uniform(-3, 3), rng.uniform(-3, 3)] + list(rng.normal(0, 10.0 ** rng.uniform(-5, 1), 4))
            got = oracle.map_lookup(tex, mode, fp)
            assert got is not None and np.allclose(got, [0.25, 0.5, 0.75], rtol=2e-6), (mode, fp, got)


def test_wide_footprints_read_the_mean():
    rng = np.random.default_rng(8)
    tex = rng.uniform(0, 1, (32, 32, 3)).astype(np.float32)
    mean = tex.astype(np.float64).mean(axis=(0, 1))
    for mode in (0, 1, 2):
        got = oracle.map_lookup(tex, mode, [0.3, 0.6, 2.0, 0.0, 0.0, 2.0])
        assert np.allclose(got, mean, atol=2e-6), mode


@pytest.mark.parametrize("shape", [(64, 64), (37, 29), (5, 70), (128, 32)])
def test_map_lookup_matches_reference(shape):
    rng = np.random.default_rng(sum(shape))
    tex = rng.uniform(0, 1, shape + (3,)).astype(np.float32)
    levels = oracle.pyramid(tex)
    exact = total = 0
    for mode in (0, 1, 2):
        for i in range(700):
            scale = 10.0 ** rng.uniform(-4.5, 0.5)
            d = rng.normal(0, scale, 4)
            if i % 7 == 0:
                d[1] = d[2] = 0.0  # axis-aligned
            if i % 11 == 0:
                d[:] = 0.0  # a zero footprint: level 0
            fp = [rng.uniform(-2, 3), rng.uniform(-2, 3)] + list(d)
            mine = oracle.map_lookup(tex, mode, fp)
            ref = ref_lookup(levels, mode, fp)
            assert (mine is None) == (ref is None), (mode, fp)
            if mine is None:
                continue
            total += 1
            if np.array_equal(mine.view(np.uint32), ref.view(np.uint32)):
                exact += 1
            elif mode != 0:
                raise AssertionError(f"mode {mode} differs: {fp} {mine} {ref}")
            else:  # theta's sine and cosine (polynomial vs libm), hypot (sqrt of the FP64 sum): a tap's position
                assert np.allclose(mine, ref, rtol=2e-4, atol=2e-4), (fp, mine, ref)
    print(f"{shape}: {exact} of {total} lookups bit-exact")
    assert exact >= 0.8 * total


def test_prototype_binding_keeps_cross_terms_in_order():
    """The prototype estimator's default binding (swapCross 0) against the reference's."""
    rng = np.random.default_rng(12)
    tex = rng.uniform(0, 1, (64, 64, 3)).astype(np.float32)
    levels = oracle.pyramid(tex)
    lib = oracle.load()
    t, keep = oracle._texture(tex)
    for _ in range(300):
        fp = [rng.uniform(0, 1), rng.uniform(0, 1)] + list(rng.normal(0, 0.01, 4))
        ref = ref_lookup(levels, 0, fp, swap=0)
        assert ref is not None
        o = out(3)
        assert lib.oracle_map_lookup_swap(ctypes.addressof(t), 0, 0, 0, fa(fp), o) == 1
        assert np.allclose(arr(o), ref, rtol=2e-4, atol=2e-4), (fp, arr(o), ref)
        # swapping the cross terms is reading (s, t, ds/dx, -ds/dy, -dt/dx, dt/dy) in order
        g = [fp[0], fp[1], fp[2], -fp[4], -fp[3], fp[5]]
        a, b = out(3), out(3)
        assert lib.oracle_map_lookup_swap(ctypes.addressof(t), 0, 0, 1, fa(fp), a) == 1
        assert lib.oracle_map_lookup_swap(ctypes.addressof(t), 0, 0, 0, fa(g), b) == 1
        assert np.array_equal(arr(a), arr(b))
    del keep


def random_triangle(rng):
    a = rng.normal(0, 1, 3)
    return a, a + rng.normal(0, 1, 3), a + rng.normal(0, 1, 3)


def test_hit_texture_derivatives_match_reference():
    rng = np.random.default_rng(21)
    lib, ref = oracle.load(), reflib.load()
    for i in range(2000):
        a, b, c = random_triangle(rng)
        uv = rng.uniform(-2, 2, 6) if i % 2 else np.array([0, 0, 1, 0, 1, 1], float)
        u, v = rng.uniform(0, 0.5, 2)
        P_ = (1 - u - v) * a + u * b + v * c
        org = P_ + rng.normal(0, 3, 3)
        d = (P_ - org) / np.linalg.norm(P_ - org)
        dx = d + rng.normal(0, 1e-3, 3)
        dy = d + rng.normal(0, 1e-3, 3)
        dx, dy = dx / np.linalg.norm(dx), dy / np.linalg.norm(dy)
        spp = int(rng.choice([1, 2, 3, 4, 16, 64]))
        mine, theirs = out(4), out(4)
        args = [fa(x) for x in (a, b, c, uv, org, d, dx, dy)]
        lib.oracle_hit_tex_derivatives(*args, spp, mine)
        ref.ref_hit_tex_derivatives(*args, spp, theirs)
        assert np.array_equal(arr(mine).view(np.uint32), arr(theirs).view(np.uint32)), i
        if i % 2 == 0:  # default texture coordinates: tex::footprint's (u + v, v) and derivatives
            fp = out(6)
            if ref.ref_footprint(*[fa(x) for x in (a, b, c, org, d, dx, dy)], float(u), float(v), spp, fp):
                assert np.array_equal(arr(fp)[2:].view(np.uint32), arr(mine).view(np.uint32)), i


def test_camera_footprint_on_a_facing_plane():
    """A camera at distance D looking straight at the plane z = 0, textured (s, t) = (x, y) over a 2 x 2 square: the
    next pixel's ray lands one pixel's width (2 D tan(fov/2) / width) away, scaled by 1 / min(sqrt(spp), 2)."""
    lib = oracle.load()
    D, width, fov = 5.0, 256, np.radians(40.0)
    pixel = 2 * D * np.tan(fov / 2) / width
    a, b, c = np.array([-1.0, -1, 0]), np.array([1.0, -1, 0]), np.array([1.0, 1, 0])
    uv = [-1, -1, 1, -1, 1, 1]  # (s, t) = (x, y)
    org, d = np.array([0.0, 0, D]), np.array([0.0, 0, -1])
    ang = np.arctan(pixel / D)
    dx, dy = np.array([np.sin(ang), 0, -np.cos(ang)]), np.array([0, np.sin(ang), -np.cos(ang)])
    for spp, scale in [(1, 1.0), (2, 1 / np.sqrt(2)), (16, 0.5)]:
        o = out(4)
        lib.oracle_hit_tex_derivatives(*[fa(x) for x in (a, b, c, uv, org, d, dx, dy)], spp, o)
        dsdx, dtdx, dsdy, dtdy = arr(o)
        assert abs(dsdx - pixel * scale) <= 1e-3 * pixel and abs(dtdy - pixel * scale) <= 1e-3 * pixel, (spp, o[:])
        assert abs(dtdx) <= 1e-6 and abs(dsdy) <= 1e-6


def test_transfer_and_derivatives_match_reference():
    rng = np.random.default_rng(31)
    lib, ref = oracle.load(), reflib.load()
    for i in range(2000):
        a, b, c = random_triangle(rng)
        uv = rng.uniform(-2, 2, 6) if i % 3 else np.array([0, 0, 1, 0, 1, 1], float)
        ng = np.cross(b - a, c - a)
        ng /= np.linalg.norm(ng)
        u, v = rng.uniform(0, 0.5, 2)
        P_ = (1 - u - v) * a + u * b + v * c
        org = P_ + rng.normal(0, 3, 3)
        d = (P_ - org) / np.linalg.norm(P_ - org)
        diff = np.concatenate([rng.normal(0, 1e-2, 6), rng.normal(0, 1e-2, 6)])
        dPds, dPdt = out(3), out(3)
        args = [fa(x) for x in (a, b, c, uv)]
        ok = lib.oracle_triangle_partials(*args, dPds, dPdt)
        rs, rt = out(3), out(3)
        assert ok == ref.ref_triangle_partials(*args, rs, rt)
        if ok:
            assert np.array_equal(arr(dPds), arr(rs)) and np.array_equal(arr(dPdt), arr(rt))
        else:
            continue
        valid = int(i % 17 != 0)
        mine, theirs = out(10), out(10)
        targs = [fa(diff), valid] + [fa(x) for x in (org, d, ng, P_)] + [dPds, dPdt]
        mv = lib.oracle_transfer(*targs, mine)
        rv = ref.ref_transfer(*targs, theirs)
        assert mv == rv and np.array_equal(arr(mine).view(np.uint32), arr(theirs).view(np.uint32)), i


def test_lobe_differentials_match_reference():
    rng = np.random.default_rng(41)
    lib, ref = oracle.load(), reflib.load()
    worst = 0.0
    for i in range(3000):
        r1, r2 = rng.uniform(0, 1, 2)
        mine, theirs = out(6), out(6)
        lib.oracle_hemisphere_differentials(r1, r2, mine)
        ref.ref_hemisphere_differentials(r1, r2, theirs)
        m, t = arr(mine), arr(theirs)
        worst = max(worst, float(np.abs(m - t).max() / max(1.0, np.abs(t).max())))
    assert worst <= 2e-6, worst  # the polynomial's sin/cos of phi against libm's
    for i in range(3000):
        kind = i % 2
        N = rng.normal(0, 1, 3)
        N /= np.linalg.norm(N)
        wo = N + rng.normal(0, 0.5, 3)
        wo /= np.linalg.norm(wo)
        wi = N + rng.normal(0, 0.5, 3)
        wi /= np.linalg.norm(wi)
        rough = float(rng.uniform(0.0, 0.9))
        incoming = rng.normal(0, 1e-2, 12)
        dP = rng.normal(0, 1e-2, 6)
        P_ = rng.normal(0, 1, 3)
        child = P_ + rng.normal(0, 1e-4, 3)
        r1, r2 = rng.uniform(0, 1, 2)
        scale = float(rng.choice([1.0, 0.5, 1 / np.sqrt(2)]))
        valid = int(i % 13 != 0)
        args = [kind, fa(N), rough, fa(incoming), fa(dP), valid, fa(P_), fa(wo), fa(wi), float(r1), float(r2), fa(child),
                scale]
        mine, theirs = out(12), out(12)
        mv = lib.oracle_scatter_differentials(*args, mine)
        rv = ref.ref_scatter_differentials(*args, theirs)
        assert mv == rv, i
        if not mv:
            continue
        if kind == 1:  # GGX: the same arithmetic
            assert np.array_equal(arr(mine).view(np.uint32), arr(theirs).view(np.uint32)), i
        else:
            # the polynomial's sin/cos of phi, carried through normalize(wi + dD) - wi: relative to the vector
            m, t = arr(mine), arr(theirs)
            assert np.abs(m - t).max() <= 1e-5 * np.abs(t).max(), i
