"""The oracle's closest-hit and any-hit traversal against brute force (every triangle tested, the least (t, geom,
prim) kept), and the watertight triangle test's guarantees. The reference leaves membership to OptiX/Embree, whose
results no reference test pins (parity there is pinned by brute force, as here)."""
import ctypes

import numpy as np

import oracle
from x2amd import scenes

f3 = ctypes.c_float * 3


def random_soup(n, seed):
    rng = np.random.default_rng(seed)
    centers = rng.uniform(-5, 5, (n, 3))
    tris = centers[:, None, :] + rng.normal(0, 0.6, (n, 3, 3))
    positions = tris.reshape(-1, 3).astype(np.float32)
    indices = np.arange(3 * n).reshape(n, 3)
    first = [0, n // 2, n]
    return scenes.SceneData(positions, indices, first, [0, 0], [scenes.material((0.5, 0.5, 0.5))], [], "soup")


def test_closest_matches_brute_force():
    s = oracle.OracleScene(random_soup(3000, 1))
    rng = np.random.default_rng(2)
    hits = 0
    for _ in range(3000):
        o = rng.uniform(-7, 7, 3)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        near = 0.0 if rng.random() < 0.7 else float(rng.uniform(0, 3))
        far = 3.4e38 if rng.random() < 0.7 else float(rng.uniform(3, 12))
        a = s.closest(o, d, near, far, brute=False)
        b = s.closest(o, d, near, far, brute=True)
        assert a == b
        hits += a[0] != 0xFFFFFFFF
        assert s.occluded(o, d, near, far) == s.occluded(o, d, near, far, brute=True) == (b[0] != 0xFFFFFFFF)
    assert hits > 500


def test_cornell_closest_matches_brute_force():
    scene, cfg = scenes.cornell()
    s = oracle.OracleScene(scene)
    rng = np.random.default_rng(3)
    for _ in range(3000):
        o = rng.uniform(-0.95, 0.95, 3)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        assert s.closest(o, d) == s.closest(o, d, brute=True)


def test_watertight_shared_edges():
    """Rays aimed exactly at shared edges and the shared vertex of a closed fan never pass between its triangles."""
    rng = np.random.default_rng(4)
    k = 7
    ang = np.linspace(0, 2 * np.pi, k, endpoint=False)
    rim = np.stack([np.cos(ang), np.sin(ang), np.zeros(k)], 1)
    center = np.zeros(3)
    L = oracle.load()
    missed = 0
    for _ in range(4000):
        i = int(rng.integers(0, k))
        t = float(rng.random()) if rng.random() < 0.9 else 0.0  # a point on edge (center, rim[i]), or the vertex
        target = (center + t * rim[i]).astype(np.float32)
        o = (target + rng.normal(size=3) * 3 + np.array([0, 0, 2.0])).astype(np.float32)
        d = (target.astype(np.float64) - o)
        d = (d / np.linalg.norm(d)).astype(np.float32)
        got = False
        for j in range(k):
            a, b = rim[j].astype(np.float32), rim[(j + 1) % k].astype(np.float32)
            tt = ctypes.c_float()
            if L.oracle_triangle(f3(*o), f3(*d), 0.0, 3.4e38, f3(*center.astype(np.float32)), f3(*a), f3(*b),
                                 ctypes.byref(tt)):
                got = True
        missed += not got
    assert missed == 0


def test_triangle_membership_matches_fp64_away_from_edges():
    """Away from edges the watertight test agrees with exact barycentric membership (computed in FP64)."""
    L = oracle.load()
    rng = np.random.default_rng(5)
    checked = 0
    for _ in range(5000):
        a, b, c = (rng.uniform(-2, 2, 3) for _ in range(3))
        o = rng.uniform(-6, 6, 3)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        a32, b32, c32, o32, d32 = (np.asarray(v, np.float32) for v in (a, b, c, o, d))
        A, B, C, O, D = (np.asarray(v, np.float64) for v in (a32, b32, c32, o32, d32))
        e1, e2, s = B - A, C - A, O - A
        p = np.cross(D, e2)
        det = np.dot(e1, p)
        if abs(det) < 1e-6:
            continue
        u = np.dot(s, p) / det
        q = np.cross(s, e1)
        v = np.dot(D, q) / det
        t = np.dot(e2, q) / det
        margin = min(u, v, 1 - u - v)
        if abs(margin) < 1e-4 or abs(t) < 1e-4:
            continue
        inside = margin > 0 and t > 0
        tt = ctypes.c_float()
        got = L.oracle_triangle(f3(*o32), f3(*d32), 0.0, 3.4e38, f3(*a32), f3(*b32), f3(*c32), ctypes.byref(tt))
        assert bool(got) == inside
        if got:
            assert abs(tt.value - t) <= 1e-4 * max(1, abs(t)) / min(1.0, abs(det))
        checked += 1
    assert checked > 3000


def quad_scene(single):
    """Two triangles of the square [-1, 1]^2 at z = 0 wound so that (b - a) x (c - a) points to +z."""
    pos = np.array([[-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]], np.float32)
    return scenes.SceneData(pos, np.array([[0, 1, 2], [0, 2, 3]]), [0, 2], [0], [scenes.material((0.5, 0.5, 0.5))],
                            [], "quad", single_sided=[0] if single else ())


def test_single_sided_quad():
    """A single-sided mesh is hit only by rays against its (b - a) x (c - a); a two-sided one from either side."""
    one, two = oracle.OracleScene(quad_scene(True)), oracle.OracleScene(quad_scene(False))
    rng = np.random.default_rng(9)
    for _ in range(400):
        p = np.array([*rng.uniform(-0.9, 0.9, 2), 0.0])
        side = 1.0 if rng.random() < 0.5 else -1.0
        o = p + np.array([*rng.uniform(-0.5, 0.5, 2), side * rng.uniform(0.5, 3)])
        d = (p - o) / np.linalg.norm(p - o)
        for brute in (False, True):
            assert two.closest(o, d, brute=brute)[0] != 0xFFFFFFFF
            assert two.occluded(o, d, 0.0, 3.4e38, brute=brute)
            front = side > 0  # from +z the ray runs against the normal
            assert (one.closest(o, d, brute=brute)[0] != 0xFFFFFFFF) == front
            assert one.occluded(o, d, 0.0, 3.4e38, brute=brute) == front


def test_single_sided_soup_matches_two_sided_where_it_faces():
    """Over a soup whose second mesh is single-sided: the BVH walk equals brute force; where the two-sided closest
    hit is a front face or on the two-sided mesh it is also the single-sided scene's; where it is a culled back face,
    the single-sided scene's hit lies beyond it (or there is none) and is not itself a culled back face."""
    soup2 = random_soup(3000, 5)
    soup1 = scenes.SceneData(soup2.positions, soup2.indices, soup2.geom_first, soup2.geom_material,
                             [scenes.material((0.5, 0.5, 0.5))], [], "soup", single_sided=[1])
    one, two = oracle.OracleScene(soup1), oracle.OracleScene(soup2)
    tri = soup2.positions.reshape(-1, 3, 3).astype(np.float64)
    normal = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    rng = np.random.default_rng(6)
    culled_seen = 0
    for _ in range(3000):
        o = rng.uniform(-7, 7, 3)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        a, b = one.closest(o, d), one.closest(o, d, brute=True)
        assert a == b
        assert one.occluded(o, d, 0.0, 3.4e38) == one.occluded(o, d, 0.0, 3.4e38, brute=True) == (a[0] != 0xFFFFFFFF)
        t2 = two.closest(o, d)
        if t2[0] == 0xFFFFFFFF:
            assert a[0] == 0xFFFFFFFF
            continue
        g2, p2 = t2[0], t2[1]  # (geom, prim, t)
        k2 = int(soup2.geom_first[g2]) + p2
        back = g2 == 1 and float(normal[k2] @ d) > 0  # clear of the dot's rounding
        front = g2 == 0 or float(normal[k2] @ d) < 0
        if front:
            assert a == t2
        elif back:
            culled_seen += 1
            if a[0] != 0xFFFFFFFF:
                assert a[2] >= t2[2]
                k1 = int(soup2.geom_first[a[0]]) + a[1]
                assert a[0] == 0 or float(normal[k1] @ d) < 0
    assert culled_seen > 100
