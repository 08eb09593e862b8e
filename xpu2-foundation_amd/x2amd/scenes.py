"""Seeded synthetic scenes and integrator settings for the device's workloads.

The reference builds its flat plan from MoonRay scene files (cpu_scene.h, out of scope here); these generators make
the same kind of flat plan directly: float32 triangle soups grouped into meshes, one material per mesh, lights as
render-space records (light_core.h:59-117) and a pinhole camera as row-vector matrices (camera_core.h:217-250).
Every array is a pure function of the arguments and seeds, so the device and the CPU oracle read the same bytes.

Workloads:
  cornell():    a Cornell-style box, 34 triangles, one ceiling rectangle light (radiance 17, 12, 4), Lambert walls,
                256x256, 16 spp, depth 8, Russian roulette from non-mirror depth 3.
  icospheres(): 32 fBm-displaced level-6 icospheres (81,920 triangles each) over a ground quad, 2,621,442
                triangles; 50/50 Lambert + GGX materials; 3 rectangle lights and a constant environment of 0.2;
                1920x1080, 64 spp, depth 12.
"""
import ctypes
import math

import numpy as np

from . import abi

ALL_LIGHT_VISIBILITY = 0xFC  # rdl2 VisibilityType's lobe bits (light_core.h:188 VisibleAllLights)
MOONRAY_RR_THRESHOLD = 0.0375  # MoonRay's russian_roulette_threshold default
FULL_VIEWPORT = (0, 0, 0x7FFFFFFF, 0x7FFFFFFF)


def frame_seed(frame=0, eye=0, lock=False):
    """MoonRay's initial seed for a frame (surface_core.h:21-23 frameSeed)."""
    if lock:
        return 0
    return ((frame + 8429) * (2 * (1193047 + eye + 6037) + 1)) & 0xFFFFFFFF


class SceneData:
    """A flat triangle scene and the ctypes view of it (kept alive with the arrays)."""

    def __init__(self, positions, indices, geom_first, geom_material, materials, lights, name, uvs=None,
                 textures=(), volume=None, single_sided=()):
        self.name = name
        self.positions = np.ascontiguousarray(positions, dtype=np.float32)
        self.indices = np.ascontiguousarray(indices, dtype=np.uint32)
        self.geom_first = np.ascontiguousarray(geom_first, dtype=np.uint32)
        self.geom_material = np.ascontiguousarray(geom_material, dtype=np.uint32)
        self.materials = (abi.Material * max(1, len(materials)))(*materials)
        self.material_count = len(materials)
        self.lights = (abi.Light * max(1, len(lights)))(*lights)
        self.light_count = len(lights)
        self.c = abi.Scene()
        self.c.positions = self.positions.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        self.c.indices = self.indices.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
        self.c.geom_first = self.geom_first.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
        self.c.geom_material = self.geom_material.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
        self.c.materials = ctypes.cast(self.materials, ctypes.POINTER(abi.Material))
        self.c.lights = ctypes.cast(self.lights, ctypes.POINTER(abi.Light))
        self.c.vertex_count = self.positions.shape[0]
        self.c.triangle_count = self.indices.shape[0]
        self.c.geom_count = len(self.geom_first) - 1
        self.c.material_count = self.material_count
        self.c.light_count = self.light_count
        self.uvs = None if uvs is None else np.ascontiguousarray(uvs, dtype=np.float32)
        if self.uvs is not None:
            assert self.uvs.shape == (self.positions.shape[0], 2)
            self.c.uvs = self.uvs.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        self.texture_arrays = [np.ascontiguousarray(t, dtype=np.float32) for t in textures]
        self.textures = (abi.Texture * max(1, len(self.texture_arrays)))()
        for i, t in enumerate(self.texture_arrays):
            assert t.ndim == 3 and t.shape[2] == 3
            self.textures[i].height, self.textures[i].width = t.shape[0], t.shape[1]
            self.textures[i].texels = t.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        self.c.textures = ctypes.cast(self.textures, ctypes.POINTER(abi.Texture))
        self.c.texture_count = len(self.texture_arrays)
        self.volume = volume  # a VolumeData or None
        if volume is not None:
            self.c.volumes = ctypes.pointer(volume.c)
            self.c.volume_count = 1
        self.geom_flags = None  # the meshes whose back faces are culled (single_sided: their indices)
        if len(single_sided):
            self.geom_flags = np.zeros(len(self.geom_first) - 1, np.uint32)
            self.geom_flags[list(single_sided)] = abi.X2_GEOM_SINGLE_SIDED
            self.c.geom_flags = self.geom_flags.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))

    @property
    def triangle_count(self):
        return int(self.indices.shape[0])


def material(albedo, diffuse=1.0, specular=0.0, roughness=0.5, emission=(0.0, 0.0, 0.0), albedo_map=0):
    m = abi.Material()
    m.albedo_map = albedo_map
    m.albedo[:] = [float(a) for a in albedo]
    m.roughness = float(roughness)
    m.diffuse_weight = float(diffuse)
    m.specular_weight = float(specular)
    m.emission[:] = [float(e) for e in emission]
    return m


def _f32(v):
    return [float(np.float32(x)) for x in v]


def rect_light(center, u_axis, v_axis, radiance, sidedness=abi.X2_SIDE_REGULAR):
    """A rectangle light centred at `center` with half-extent vectors u_axis, v_axis; it lights the side
    u x v points to (its local +Z, light_core.h:66 direction)."""
    c = np.asarray(center, dtype=np.float64)
    u = np.asarray(u_axis, dtype=np.float64)
    v = np.asarray(v_axis, dtype=np.float64)
    hw, hh = np.linalg.norm(u), np.linalg.norm(v)
    ex, ey = u / hw, v / hh
    n = np.cross(ex, ey)
    n /= np.linalg.norm(n)
    m = np.stack([ex, ey, n])  # local-to-render rows, scale 1
    inv = m.T  # rows of the inverse of an orthonormal frame
    L = abi.Light()
    L.type = abi.X2_LIGHT_RECT
    L.sidedness = sidedness
    L.hemispherical = 0
    L.visibility_mask = ALL_LIGHT_VISIBILITY
    L.radiance[:] = _f32(radiance)
    L.inv_area = float(np.float32(1.0 / (4.0 * hw * hh)))
    L.position[:] = _f32(c)
    L.direction[:] = _f32(n)
    L.plane_offset = float(-np.dot(np.asarray(L.direction[:], np.float64), np.asarray(L.position[:], np.float64)))
    L.scale = 1.0
    L.local_to_render[:] = _f32(m.reshape(-1))
    L.render_to_local[:] = _f32(inv.reshape(-1))
    L.render_to_local_rot[:] = _f32(inv.reshape(-1))
    L.half_width = float(np.float32(hw))
    L.half_height = float(np.float32(hh))
    return L


def env_light(radiance):
    """A constant environment over the whole sphere (EnvLight without a texture, light_core.h:478-500)."""
    L = abi.Light()
    L.type = abi.X2_LIGHT_ENV
    L.sidedness = abi.X2_SIDE_REGULAR
    L.hemispherical = 0
    L.visibility_mask = ALL_LIGHT_VISIBILITY
    L.radiance[:] = _f32(radiance)
    L.inv_area = float(np.float32(1.0 / (4.0 * math.pi)))
    L.direction[:] = [0.0, 1.0, 0.0]
    L.scale = 1.0
    eye = [1.0, 0, 0, 0, 1.0, 0, 0, 0, 1.0]
    L.local_to_render[:] = eye
    L.render_to_local[:] = eye
    L.render_to_local_rot[:] = eye
    return L


def camera_matrices(eye, target, up, fov_deg, width, height):
    """Raster-to-camera and camera-to-render matrices (row-vector, m[row][col]) of a pinhole camera whose vertical
    field of view is fov_deg; raster (x, y, -1) maps to camera (X, Y, -1)."""
    t = math.tan(math.radians(fov_deg) * 0.5)
    aspect = width / height
    r2c = np.zeros((4, 4))
    r2c[0, 0] = 2.0 * t * aspect / width
    r2c[3, 0] = -t * aspect
    r2c[1, 1] = -2.0 * t / height
    r2c[3, 1] = t
    r2c[2, 2] = 1.0
    r2c[3, 3] = 1.0
    e = np.asarray(eye, np.float64)
    f = np.asarray(target, np.float64) - e
    f /= np.linalg.norm(f)
    right = np.cross(f, np.asarray(up, np.float64))
    right /= np.linalg.norm(right)
    upv = np.cross(right, f)
    c2r = np.zeros((4, 4))
    c2r[0, :3] = right
    c2r[1, :3] = upv
    c2r[2, :3] = -f
    c2r[3, :3] = e
    c2r[3, 3] = 1.0
    return r2c.astype(np.float32).reshape(-1), c2r.astype(np.float32).reshape(-1)


def make_config(width, height, spp, eye, target, up=(0, 1, 0), fov=40.0, light_count=1, max_depth=8,
                rr_threshold=MOONRAY_RR_THRESHOLD, rr_min_depth=1, seed=None, sampling=3 | 4 | 8,
                shadow_reuse=ALL_LIGHT_VISIBILITY, bsdf_samples=1, light_samples=1):
    """The integrator's settings. sampling 15: policy 3 with the path tracer's bits 4|8 (cpu_jobs.h:14)."""
    c = abi.Config()
    c.width, c.height, c.spp = width, height, spp
    c.seed = frame_seed() if seed is None else seed
    c.sampling = sampling
    c.filter = 0
    c.filter_width = 1.0
    c.near_clip = 0.001
    c.far_clip = 10000.0
    r2c, c2r = camera_matrices(eye, target, up, fov, width, height)
    c.raster_to_camera[:] = [float(x) for x in r2c]
    c.camera_to_render[:] = [float(x) for x in c2r]
    c.light_count = light_count
    c.bsdf_samples = bsdf_samples
    c.light_samples = light_samples
    c.max_depth = c.max_diffuse_depth = c.max_glossy_depth = c.max_mirror_depth = max_depth
    c.clamp_depth = 1
    c.clamp_value = 0.0
    c.rr_threshold = rr_threshold
    c.rr_min_depth = rr_min_depth
    c.shadow_reuse = shadow_reuse
    c.viewport[:] = list(FULL_VIEWPORT)
    # the volume settings' defaults (gpu_jobs.h:19-28)
    c.max_volume_depth = 1
    c.volume_illumination_samples = 4
    c.volume_transmittance_threshold = 0.005
    c.inv_volume_quality = 2.0
    c.volume_attenuation_factor = 0.65
    c.volume_contribution_factor = 0.65
    c.volume_phase_attenuation_factor = 0.5
    c.direct_only = 0  # the path tracer; 1: the prototype estimator (integrateDirect)
    c.surface_limit = 1
    c.sky[:] = [0.0, 0.0, 0.0]
    return c


# ------------------------------------------------------------------------------------------------------------------
# The Cornell-style box.
def _quad(a, b, c, d):
    return [a, b, c, d], [[0, 1, 2], [0, 2, 3]]


def _box(center, size, angle_deg):
    hx, hy, hz = (s * 0.5 for s in size)
    corners = np.array([[x, y, z] for y in (-hy, hy) for z in (-hz, hz) for x in (-hx, hx)], np.float64)
    a = math.radians(angle_deg)
    rot = np.array([[math.cos(a), 0, math.sin(a)], [0, 1, 0], [-math.sin(a), 0, math.cos(a)]])
    v = corners @ rot.T + np.asarray(center, np.float64)
    # corners: 0..3 bottom (y=-h), 4..7 top; within each: (x-,z-), (x+,z-), (x-,z+), (x+,z+)
    faces = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 2, 6, 4), (1, 5, 7, 3), (0, 4, 5, 1), (2, 3, 7, 6)]
    tris = []
    for p, q, r, s in faces:
        tris += [[p, q, r], [p, r, s]]
    return v, tris


def cornell(width=256, height=256, spp=16, max_depth=8, rr_min_depth=3):
    """A Cornell-style box in [-1, 1]^3, open toward +z: 5 wall quads (10 triangles), a tall box turned 15 degrees
    and a short box turned -18 degrees about y (12 triangles each), and a rectangle light under the ceiling."""
    grey, red, green = (0.73, 0.73, 0.73), (0.65, 0.05, 0.05), (0.12, 0.45, 0.15)
    walls = [
        (_quad((-1, -1, 1), (1, -1, 1), (1, -1, -1), (-1, -1, -1)), 0),  # floor
        (_quad((-1, 1, 1), (-1, 1, -1), (1, 1, -1), (1, 1, 1)), 0),  # ceiling
        (_quad((-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1)), 0),  # back
        (_quad((-1, -1, 1), (-1, -1, -1), (-1, 1, -1), (-1, 1, 1)), 1),  # left (red)
        (_quad((1, -1, 1), (1, 1, 1), (1, 1, -1), (1, -1, -1)), 2),  # right (green)
    ]
    meshes = [((np.asarray(v, np.float64), t), m) for (v, t), m in walls]
    meshes.append((_box((-0.35, -0.4, -0.3), (0.6, 1.2, 0.6), 15.0), 0))
    meshes.append((_box((0.35, -0.7, 0.3), (0.6, 0.6, 0.6), -18.0), 0))
    positions, indices, first, mats = [], [], [0], []
    base = 0
    for (v, t), m in meshes:
        positions.append(np.asarray(v, np.float64))
        indices.append(np.asarray(t, np.int64) + base)
        base += len(v)
        first.append(first[-1] + len(t))
        mats.append(m)
    materials = [material(grey), material(red), material(green)]
    lights = [rect_light((0.0, 0.999, 0.0), (0.25, 0.0, 0.0), (0.0, 0.0, 0.2), (17.0, 12.0, 4.0))]
    # u x v = (0.25,0,0) x (0,0,0.2) points to -y: the light faces down into the box.
    scene = SceneData(np.concatenate(positions), np.concatenate(indices), first, mats, materials, lights, "cornell")
    config = make_config(width, height, spp, eye=(0.0, 0.0, 3.75), target=(0.0, 0.0, 0.0), light_count=1,
                         max_depth=max_depth, rr_min_depth=rr_min_depth)
    return scene, config


# ------------------------------------------------------------------------------------------------------------------
# Displaced icospheres.
def icosphere(level):
    """A unit icosphere subdivided `level` times (20 * 4^level triangles)."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.int64)
    for _ in range(level):
        edges = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        key = np.sort(edges, axis=1)
        uniq, inverse = np.unique(key[:, 0] * len(v) + key[:, 1], return_inverse=True)
        a, b = uniq // len(v), uniq % len(v)
        mid = v[a] + v[b]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        m = (inverse + len(v)).reshape(3, -1).T  # midpoints of edges ab, bc, ca
        v = np.concatenate([v, mid])
        f = np.concatenate([
            np.stack([f[:, 0], m[:, 0], m[:, 2]], 1), np.stack([f[:, 1], m[:, 1], m[:, 0]], 1),
            np.stack([f[:, 2], m[:, 2], m[:, 1]], 1), np.stack([m[:, 0], m[:, 1], m[:, 2]], 1)])
    return v, f


def _hash3(ix, iy, iz, seed):
    mix = np.uint64((seed * 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF)
    h = (ix.astype(np.uint64) * np.uint64(0x8DA6B343) ^ iy.astype(np.uint64) * np.uint64(0xD8163841)
         ^ iz.astype(np.uint64) * np.uint64(0xCB1AB31F) ^ mix)
    h ^= h >> np.uint64(29)
    h *= np.uint64(0xBF58476D1CE4E5B9)
    h ^= h >> np.uint64(32)
    return (h & np.uint64(0xFFFFFF)).astype(np.float64) / float(0xFFFFFF) * 2.0 - 1.0


def value_noise(p, seed):
    """Trilinear value noise in [-1, 1] on the integer lattice, smoothstep-blended."""
    i = np.floor(p).astype(np.int64)
    f = p - i
    w = f * f * (3.0 - 2.0 * f)
    out = 0.0
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                h = _hash3(i[:, 0] + dx, i[:, 1] + dy, i[:, 2] + dz, seed)
                wx = w[:, 0] if dx else 1 - w[:, 0]
                wy = w[:, 1] if dy else 1 - w[:, 1]
                wz = w[:, 2] if dz else 1 - w[:, 2]
                out = out + h * wx * wy * wz
    return out


def fbm(p, seed, octaves=5, lacunarity=2.0, gain=0.5):
    total, amp, norm = 0.0, 1.0, 0.0
    for o in range(octaves):
        total = total + amp * value_noise(p * (lacunarity ** o), seed + 101 * o)
        norm += amp
        amp *= gain
    return total / norm


def icospheres(width=1920, height=1080, spp=64, count=32, level=6, max_depth=12, seed=1, rr_min_depth=1):
    """`count` icospheres of `level`, each radius-displaced by fBm (amplitude 0.1) and placed uniformly in a 20^3
    box over a ground quad; per material 50/50 Lambert + GGX, roughness U[0.05, 0.8], albedo U[0.1, 0.9]; three
    rectangle lights and a constant environment of 0.2."""
    rng = np.random.default_rng(seed)
    v0, f0 = icosphere(level)
    positions, indices, first, mats, materials = [], [], [0], [], []
    base = 0
    centers = np.stack([rng.uniform(-10, 10, count), rng.uniform(0, 20, count), rng.uniform(-10, 10, count)], 1)
    for i in range(count):
        r = 1.0 + 0.1 * fbm(v0 * 2.0, seed * 1000 + i)
        positions.append(v0 * r[:, None] + centers[i])
        indices.append(f0 + base)
        base += len(v0)
        first.append(first[-1] + len(f0))
        mats.append(i)
        materials.append(material(rng.uniform(0.1, 0.9, 3), diffuse=0.5, specular=0.5,
                                  roughness=rng.uniform(0.05, 0.8)))
    ground = np.array([[-60, -1.5, -60], [60, -1.5, -60], [60, -1.5, 60], [-60, -1.5, 60]], np.float64)
    positions.append(ground)
    indices.append(np.array([[0, 2, 1], [0, 3, 2]], np.int64) + base)
    first.append(first[-1] + 2)
    mats.append(count)
    materials.append(material((0.5, 0.5, 0.5), diffuse=0.5, specular=0.5, roughness=0.6))
    lights = [
        rect_light((-8.0, 30.0, 0.0), (3.0, 0, 0), (0, 0, 3.0), (20.0, 18.0, 15.0)),
        rect_light((8.0, 30.0, 5.0), (3.0, 0, 0), (0, 0, 3.0), (15.0, 18.0, 20.0)),
        rect_light((0.0, 25.0, 20.0), (4.0, 0, 0), (0, -1.0, 2.0), (18.0, 18.0, 18.0)),
        env_light((0.2, 0.2, 0.2)),
    ]
    scene = SceneData(np.concatenate(positions), np.concatenate(indices), first, mats, materials, lights,
                      "icospheres")
    config = make_config(width, height, spp, eye=(0.0, 10.0, 30.0), target=(0.0, 9.0, 0.0), fov=40.0,
                         light_count=len(lights), max_depth=max_depth, rr_min_depth=rr_min_depth)
    return scene, config


def tiles(config):
    return ((config.width + 7) // 8) * ((config.height + 7) // 8)


def uniform_chunks(config, samples_per_chunk=4, progressive=True):
    """The chunks of a uniform frame as the work grid makes them (async_grid.h:129-175): sample ranges 1, 1, 2,
    then samples_per_chunk (progressive), each over every tile, cell by cell (64-pixel cells)."""
    ranges = []
    first = 0
    while first < config.spp:
        size = samples_per_chunk if not progressive else (1 if first < 2 else 2 if first < 4 else samples_per_chunk)
        ranges.append((first, min(size, config.spp - first)))
        first += ranges[-1][1]
    tiles_x = (config.width + 7) // 8
    tiles_y = (config.height + 7) // 8
    cells_x = (config.width + 63) // 64
    cells_y = (config.height + 63) // 64
    out = []
    for r0, rn in ranges:
        for cy in range(cells_y):
            for cx in range(cells_x):
                for ty in range(cy * 8, min(tiles_y, (cy + 1) * 8)):
                    for tx in range(cx * 8, min(tiles_x, (cx + 1) * 8)):
                        out.append((ty * tiles_x + tx, r0, rn, cy * cells_x + cx))
    arr = (abi.Chunk * len(out))()
    for i, c in enumerate(out):
        arr[i].tile, arr[i].first, arr[i].count, arr[i].cell = c
    return arr


# ------------------------------------------------------------------------------------------------------------------
# The differentiable Cornell box.
def value_noise_texture(size, seed, cells=16, lo=0.1, hi=0.9):
    """A size x size RGB texture of seeded value noise in [lo, hi]: a random lattice of cells x cells per channel,
    smoothstep-interpolated, periodic (it tiles with periodic addressing)."""
    rng = np.random.default_rng(seed)
    out = np.empty((size, size, 3), np.float64)
    x = (np.arange(size) + 0.5) / size * cells
    i0 = np.floor(x).astype(np.int64)
    f = x - i0
    w = f * f * (3 - 2 * f)
    for ch in range(3):
        lat = rng.uniform(lo, hi, (cells, cells))
        a = lat[np.ix_(i0 % cells, i0 % cells)]
        b = lat[np.ix_(i0 % cells, (i0 + 1) % cells)]
        c = lat[np.ix_((i0 + 1) % cells, i0 % cells)]
        d = lat[np.ix_((i0 + 1) % cells, (i0 + 1) % cells)]
        wy, wx = w[:, None], w[None, :]
        out[:, :, ch] = (a * (1 - wx) + b * wx) * (1 - wy) + (c * (1 - wx) + d * wx) * wy
    return out.astype(np.float32)


def cornell_textured(width=512, height=512, spp=32, max_depth=8, texture_size=512, texture=None, roughness=None,
                     seed=7):
    """The Cornell box with its back wall's albedo a texture_size^2 RGB FP32 map (value noise in [0.1, 0.9]) and four
    50/50 Lambert + GGX materials whose roughness is a parameter (floor and ceiling, left, right, the boxes).
    Russian roulette is off so the fixed-seed estimator is differentiable in the parameters (no roulette scale).
    Returns (scene, config, params) with params = {"texture": HxWx3, "roughness": [4], "roughness_materials": [...]}."""
    grey, red, green = (0.73, 0.73, 0.73), (0.65, 0.05, 0.05), (0.12, 0.45, 0.15)
    tex = value_noise_texture(texture_size, seed) if texture is None else np.asarray(texture, np.float32)
    rough = np.array([0.3, 0.5, 0.2, 0.6] if roughness is None else roughness, np.float64)
    walls = [
        (_quad((-1, -1, 1), (1, -1, 1), (1, -1, -1), (-1, -1, -1)), 1),  # floor
        (_quad((-1, 1, 1), (-1, 1, -1), (1, 1, -1), (1, 1, 1)), 1),  # ceiling
        (_quad((-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1)), 0),  # back: textured
        (_quad((-1, -1, 1), (-1, -1, -1), (-1, 1, -1), (-1, 1, 1)), 2),  # left
        (_quad((1, -1, 1), (1, 1, 1), (1, 1, -1), (1, -1, -1)), 3),  # right
    ]
    meshes = [((np.asarray(v, np.float64), t), m) for (v, t), m in walls]
    meshes.append((_box((-0.35, -0.4, -0.3), (0.6, 1.2, 0.6), 15.0), 4))
    meshes.append((_box((0.35, -0.7, 0.3), (0.6, 0.6, 0.6), -18.0), 4))
    positions, indices, first, mats, uvs = [], [], [0], [], []
    base = 0
    for k, ((v, t), m) in enumerate(meshes):
        positions.append(np.asarray(v, np.float64))
        indices.append(np.asarray(t, np.int64) + base)
        base += len(v)
        first.append(first[-1] + len(t))
        mats.append(m)
        # the back wall's corners (-1,-1), (1,-1), (1,1), (-1,1) map to (0,0), (1,0), (1,1), (0,1)
        uvs.append(np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float64) if k == 2 else np.zeros((len(v), 2)))
    materials = [material((1.0, 1.0, 1.0), albedo_map=1),
                 material(grey, diffuse=0.5, specular=0.5, roughness=rough[0]),
                 material(red, diffuse=0.5, specular=0.5, roughness=rough[1]),
                 material(green, diffuse=0.5, specular=0.5, roughness=rough[2]),
                 material(grey, diffuse=0.5, specular=0.5, roughness=rough[3])]
    lights = [rect_light((0.0, 0.999, 0.0), (0.25, 0.0, 0.0), (0.0, 0.0, 0.2), (17.0, 12.0, 4.0))]
    scene = SceneData(np.concatenate(positions), np.concatenate(indices), first, mats, materials, lights,
                      "cornell_textured", uvs=np.concatenate(uvs), textures=[tex])
    config = make_config(width, height, spp, eye=(0.0, 0.0, 3.75), target=(0.0, 0.0, 0.0), light_count=1,
                         max_depth=max_depth, rr_threshold=0.0)
    params = {"texture": tex, "roughness": rough.astype(np.float32), "roughness_materials": [1, 2, 3, 4]}
    return scene, config, params


# ------------------------------------------------------------------------------------------------------------------
# Participating media.
class VolumeData:
    """A box-bounded medium (x2_volume): density and temperature grids as float arrays [z, y, x] (stored as IEEE
    halves, x fastest), a blackbody ramp for the emission."""

    def __init__(self, bounds_min, bounds_max, density, temperature=None, sigma_t_scale=4.0, albedo=(0.8, 0.8, 0.8),
                 g=0.6, emission_scale=0.0, ramp=None, ramp_max_kelvin=1500.0):
        d = np.asarray(density, np.float32)
        assert d.ndim == 3 and all(n % 8 == 0 for n in d.shape)
        self.density = np.ascontiguousarray(d.astype(np.float16)).view(np.uint16)
        self.temperature = None if temperature is None else \
            np.ascontiguousarray(np.asarray(temperature, np.float32).astype(np.float16)).view(np.uint16)
        self.ramp = None if ramp is None else np.ascontiguousarray(ramp, np.float32)
        c = self.c = abi.Volume()
        c.bounds_min[:] = [float(x) for x in bounds_min]
        c.bounds_max[:] = [float(x) for x in bounds_max]
        c.res[:] = [d.shape[2], d.shape[1], d.shape[0]]
        c.sigma_t_scale = sigma_t_scale
        c.albedo[:] = [float(x) for x in albedo]
        c.g = g
        c.emission_scale = emission_scale
        c.ramp_max_kelvin = ramp_max_kelvin
        u16p = ctypes.POINTER(ctypes.c_uint16)
        c.density = self.density.ctypes.data_as(u16p)
        if self.temperature is not None:
            c.temperature = self.temperature.ctypes.data_as(u16p)
            c.ramp = self.ramp.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
            c.ramp_size = self.ramp.shape[0]

    def occupancy(self):
        """The fraction of 8^3 bricks holding a nonzero density."""
        d = self.density.reshape(self.density.shape[0] // 8, 8, self.density.shape[1] // 8, 8,
                                 self.density.shape[2] // 8, 8)
        return float(((d & 0x7FFF) != 0).any(axis=(1, 3, 5)).mean())


def _cmf_lobe(lam, mu, s1, s2):
    s = np.where(lam < mu, s1, s2)
    return np.exp(-0.5 * ((lam - mu) / s) ** 2)


def blackbody_ramp(max_kelvin=1500.0, size=256):
    """Linear sRGB of Planck's law at size temperatures over [0, max_kelvin]: the spectrum (360-830 nm, 1 nm) against
    the CIE 1931 observer in Wyman, Sloan and Shirley's multi-lobe fit (JCGT 2013), XYZ to linear sRGB (IEC
    61966-2-1), negatives clipped, scaled so that the luminance at max_kelvin is 1."""
    lam = np.arange(360.0, 831.0)
    xb = 1.056 * _cmf_lobe(lam, 599.8, 37.9, 31.0) + 0.362 * _cmf_lobe(lam, 442.0, 16.0, 26.7) \
        - 0.065 * _cmf_lobe(lam, 501.1, 20.4, 26.2)
    yb = 0.821 * _cmf_lobe(lam, 568.8, 46.9, 40.5) + 0.286 * _cmf_lobe(lam, 530.9, 16.3, 31.1)
    zb = 1.217 * _cmf_lobe(lam, 437.0, 11.8, 36.0) + 0.681 * _cmf_lobe(lam, 459.0, 26.0, 13.8)
    h, c, k = 6.62607015e-34, 2.99792458e8, 1.380649e-23
    m = lam * 1e-9
    temps = np.linspace(0.0, max_kelvin, size)
    out = np.zeros((size, 3))
    to_rgb = np.array([[3.2406, -1.5372, -0.4986], [-0.9689, 1.8758, 0.0415], [0.0557, -0.2040, 1.0570]])
    for i, t in enumerate(temps):
        if t <= 0:
            continue
        with np.errstate(over="ignore"):
            b = 2 * h * c * c / m ** 5 / np.expm1(np.minimum(h * c / (m * k * t), 700.0))
        xyz = np.array([(b * xb).sum(), (b * yb).sum(), (b * zb).sum()])
        out[i] = np.maximum(to_rgb @ xyz, 0.0)
    top = out[-1] @ np.array([0.299, 0.587, 0.114])
    return (out / top).astype(np.float32)


def _value_noise(res, cells, rng):
    """Seeded value noise on a res^3 grid of voxel centres over `cells` lattice cells, smoothstep-interpolated."""
    lat = rng.random((cells + 1, cells + 1, cells + 1)).astype(np.float32)
    x = (np.arange(res, dtype=np.float32) + 0.5) / res * cells
    i0 = np.minimum(np.floor(x).astype(np.int64), cells - 1)
    f = x - i0
    w = f * f * (3 - 2 * f)
    out = np.empty((res, res, res), np.float32)
    for z in range(res):  # one z slice at a time keeps the temporaries small
        kz, wz = i0[z], w[z]
        a = lat[kz][np.ix_(i0, i0)] * (1 - wz) + lat[kz + 1][np.ix_(i0, i0)] * wz
        b = lat[kz][np.ix_(i0, i0 + 1)] * (1 - wz) + lat[kz + 1][np.ix_(i0, i0 + 1)] * wz
        c = lat[kz][np.ix_(i0 + 1, i0)] * (1 - wz) + lat[kz + 1][np.ix_(i0 + 1, i0)] * wz
        d = lat[kz][np.ix_(i0 + 1, i0 + 1)] * (1 - wz) + lat[kz + 1][np.ix_(i0 + 1, i0 + 1)] * wz
        wy, wx = w[:, None], w[None, :]
        out[z] = (a * (1 - wx) + b * wx) * (1 - wy) + (c * (1 - wx) + d * wx) * wy
    return out


def fbm_grid(res, seed, octaves=4, base_cells=4):
    """Fractional Brownian motion of value noise in [0, 1] (gain 0.5, lacunarity 2)."""
    rng = np.random.default_rng(seed)
    total = np.zeros((res, res, res), np.float32)
    norm = 0.0
    for o in range(octaves):
        amp = 0.5 ** o
        total += amp * _value_noise(res, base_cells * 2 ** o, rng)
        norm += amp
    return total / norm


def cloud_volume(res=256, seed=17, threshold=0.4, bounds_min=(-2.0, 0.2, -2.0), bounds_max=(2.0, 4.2, 2.0)):
    """The participating medium: fBm density thresholded at 0.4 under a spherical falloff (about 30% of the 8^3 bricks
    occupied at 256^3), extinction scale 4, albedo 0.8, HG g = 0.6, and a temperature field of 0-1500 K (a hot core
    with fBm detail) driving blackbody emission."""
    n = fbm_grid(res, seed)
    x = (np.arange(res, dtype=np.float32) + 0.5) / res * 2 - 1
    r = np.sqrt(x[:, None, None] ** 2 + x[None, :, None] ** 2 + x[None, None, :] ** 2)
    shaped = n * np.clip(1.5 - r, 0.0, 1.0) + 0.3 * np.clip(1.0 - r, 0.0, 1.0)
    density = np.clip((shaped - threshold) / (1 - threshold), 0.0, 1.0).astype(np.float32)
    hot = fbm_grid(res // 2, seed + 1, octaves=3, base_cells=3).repeat(2, 0).repeat(2, 1).repeat(2, 2)
    temperature = (1500.0 * np.clip(1.3 * (1.0 - r / 0.8) + 0.6 * (hot - 0.5), 0.0, 1.0)).astype(np.float32)
    return VolumeData(bounds_min, bounds_max, density, temperature, sigma_t_scale=4.0, albedo=(0.8, 0.8, 0.8),
                      g=0.6, emission_scale=2.0, ramp=blackbody_ramp(1500.0), ramp_max_kelvin=1500.0)


def clouds(width=1280, height=720, spp=32, res=256, max_depth=16, max_volume_depth=2, seed=17, volume=None):
    """Participating media: the cloud over a grey ground quad, one rectangle light above it and a constant
    environment of 0.2; 1280x720, 32 spp, depth 16. Returns (scene, config)."""
    vol = volume if volume is not None else cloud_volume(res, seed)
    s = 50.0
    pos = np.array([[-s, 0, -s], [s, 0, -s], [s, 0, s], [-s, 0, s]], np.float32)
    idx = np.array([[0, 2, 1], [0, 3, 2]])
    mats = [material((0.5, 0.5, 0.5))]
    lights = [rect_light((0.0, 8.0, 1.0), (1.5, 0.0, 0.0), (0.0, 0.0, 1.5), (12.0, 11.0, 10.0)),
              env_light((0.2, 0.2, 0.2))]
    scene = SceneData(pos, idx, [0, 2], [0], mats, lights, "clouds", volume=vol)
    config = make_config(width, height, spp, eye=(0.0, 2.6, 9.5), target=(0.0, 2.1, 0.0), light_count=2,
                         max_depth=max_depth)
    config.max_volume_depth = max_volume_depth
    return scene, config


# ------------------------------------------------------------------------------------------------------------------
# Batched jobs.
def batch_jobs(count=64, seed=3, min_tris=10_000, max_tris=500_000, min_side=64, max_side=1024, max_spp=64):
    """`count` seeded icosphere scenes in the headline's style, each of about log-uniform [min_tris, max_tris]
    triangles (n fBm-displaced icospheres of the largest level that gives at least two), a square image of log-uniform
    side in [min_side, max_side] (a multiple of 8) and spp uniform in 1..max_spp. Returns [(scene, config)]."""
    rng = np.random.default_rng(seed)
    jobs = []
    for i in range(count):
        target = math.exp(rng.uniform(math.log(min_tris), math.log(max_tris)))
        for level in (6, 5, 4, 3):
            per = 20 * 4 ** level
            if target / per >= 2 or level == 3:
                n = int(min(64, max(1, round(target / per))))
                break
        side = int(max(8, round(math.exp(rng.uniform(math.log(min_side), math.log(max_side))) / 8) * 8))
        spp = int(rng.integers(1, max_spp + 1))
        scene, config = icospheres(width=side, height=side, spp=spp, count=n, level=level, seed=seed * 1000 + i)
        jobs.append((scene, config))
    return jobs
