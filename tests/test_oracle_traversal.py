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
