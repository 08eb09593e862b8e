"""TEST INFRASTRUCTURE: ctypes loader for the CPU oracle (liboracle.so, built from x2_oracle.cpp by the Makefile
beside it). Only tests/, __graft_entry__.smoke() and bench.py's CPU baseline import this module."""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
_lib = None

P = ctypes.POINTER
u32, u64, f32, f64, c_int = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_float, ctypes.c_double, ctypes.c_int
vp = ctypes.c_void_p


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"oracle not built: {LIB_PATH} (make -C oracle)")
    lib = ctypes.CDLL(LIB_PATH)
    sig = {
        "oracle_philox": (None, [P(u32), P(u32), P(u32)]),
        "oracle_block": (None, [u32, u64, u64, u64, u32, P(u32)]),
        "oracle_word": (u32, [u32, u64, u64, u64, u32]),
        "oracle_unit": (f32, [u32]),
        "oracle_sobol": (u32, [u32, u32]),
        "oracle_directions": (None, [P(u32)]),
        "oracle_owen_scramble": (u32, [u32, u32]),
        "oracle_sampling_word": (u32, [u32, u64, u64, u64, u32, u32]),
        "oracle_child_branch": (u64, [u64, u32, u32]),
        "oracle_stream": (None, [u32, u64, u64, u64, u32, u32, P(f32)]),
        "oracle_sincos": (None, [f32, P(f32), P(f32)]),
        "oracle_lobe_eval": (None, [u32, P(f32), P(f32), P(f32), P(f32), f32, P(f32), P(f32)]),
        "oracle_lobe_sample": (None, [u32, P(f32), P(f32), P(f32), f32, f32, f32, P(f32), P(f32), P(f32)]),
        "oracle_light_sample": (c_int, [vp, P(f32), P(f32), P(f32), P(f32), P(f32), P(f32), P(f32)]),
        "oracle_light_intersect": (c_int, [vp, P(f32), P(f32), P(f32), P(f32), P(f32)]),
        "oracle_light_can_illuminate": (c_int, [vp, P(f32), P(f32)]),
        "oracle_triangle": (c_int, [P(f32), P(f32), f32, f32, P(f32), P(f32), P(f32), P(f32)]),
        "oracle_camera_ray": (c_int, [vp, f32, f32, P(f32), P(f32), P(f32), P(f32)]),
        "oracle_geometry_hit": (c_int, [P(f32), P(f32), P(f32), P(f32), P(f32), P(f32)]),
        "oracle_scene_create": (vp, [vp]),
        "oracle_scene_destroy": (None, [vp]),
        "oracle_closest": (None, [vp, P(f32), P(f32), f32, f32, c_int, P(u32), P(u32), P(f32)]),
        "oracle_occluded": (c_int, [vp, P(f32), P(f32), f32, f32, c_int]),
        "oracle_sample": (c_int, [vp, vp, u32, u32, P(f64), P(u64)]),
        "oracle_first_hit_draws": (c_int, [vp, vp, u32, u32, P(f32), u32, P(u32)]),
        "oracle_render": (c_int, [vp, vp, u32, P(f64), P(u32), P(u64)]),
        "oracle_render_chunks": (c_int, [vp, vp, vp, u32, u32, P(f64), P(u32), P(u64)]),
        "oracle_render_pixels": (c_int, [vp, vp, P(u32), P(u32), u32, P(f64), P(f64), P(u32), P(u64)]),
        "oracle_adjoint": (c_int, [vp, vp, u32, P(f32), P(f64), P(f64), P(u64)]),
        "oracle_set_texture": (None, [vp, u32, P(f32)]),
        "oracle_set_texture_half": (None, [vp, u32, c_int]),
        "oracle_set_texture_coarse": (None, [vp, u32, u32]),
        "oracle_map_lookup": (c_int, [vp, c_int, u32, P(f32), P(f32)]),
        "oracle_map_lookup_swap": (c_int, [vp, c_int, u32, c_int, P(f32), P(f32)]),
        "oracle_pyramid": (u32, [vp, c_int, P(f32), u64, P(u32)]),
        "oracle_hit_tex_derivatives": (None, [P(f32)] * 8 + [u32, P(f32)]),
        "oracle_transfer": (c_int, [P(f32), c_int] + [P(f32)] * 7),
        "oracle_triangle_partials": (c_int, [P(f32)] * 6),
        "oracle_hemisphere_differentials": (None, [f32, f32, P(f32)]),
        "oracle_scatter_differentials": (c_int, [u32, P(f32), f32, P(f32), P(f32), c_int, P(f32), P(f32), P(f32), f32,
                                                 f32, P(f32), f32, P(f32)]),
        "oracle_set_material": (None, [vp, u32, vp]),
        "oracle_set_eval_roughness": (None, [vp, u32, f32]),
        "oracle_ggx_roughness_log_derivative": (f32, [f32, P(f32), P(f32), P(f32)]),
        "oracle_texture_lookup": (c_int, [vp, f32, f32, P(f32)]),
        "oracle_surface_lambert": (c_int, [P(f32), P(f32), f32, f32, P(f32), P(f32), P(f32)]),
        "oracle_volume_density": (f32, [vp, P(f32)]),
        "oracle_volume_transmittance": (f32, [vp, vp, P(f32), P(f32), f32, f32, u32, u64, u64]),
        "oracle_phase_eval": (f32, [f32, P(f32), P(f32)]),
        "oracle_phase_sample": (None, [f32, P(f32), f32, f32, P(f32)]),
        "oracle_volume_march": (c_int, [vp, vp, P(f32), P(f32), f32, f32, u32, u64, P(f32)]),
        "oracle_region_info": (c_int, [vp, P(c_int), P(f32), P(f32)]),
        "oracle_region_cell_pdf": (f32, [vp, c_int, c_int, c_int]),
        "oracle_region_point": (None, [vp, P(f32), P(f32)]),
        "oracle_region_pdf": (f32, [vp, P(f32), P(f32), P(f32)]),
        "oracle_emissive_integral": (None, [vp, vp, P(f32), P(f32), f32, f32, u32, u64, P(f32)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def fa(values):
    """A float32 ctypes array from a sequence (or None)."""
    if values is None:
        return None
    a = (f32 * len(values))(*[float(v) for v in values])
    return a


class _Texture(ctypes.Structure):  # x2_texture
    _fields_ = [("width", u32), ("height", u32), ("texels", P(f32))]


def _texture(texels):
    arr = np.ascontiguousarray(texels, np.float32)
    t = _Texture(arr.shape[1], arr.shape[0], arr.ctypes.data_as(P(f32)))
    return t, arr


def map_lookup(texels, mode, footprint, half=False):
    """The path tracer's map lookup (mode 0 anisotropic, 1 trilinear closest, 2 closest on one level) over the
    texture's pyramid at footprint (s, t, ds/dx, dt/dx, ds/dy, dt/dy); None for an invalid value."""
    t, keep = _texture(texels)
    out = (f32 * 3)()
    ok = load().oracle_map_lookup(ctypes.addressof(t), 1 if half else 0, mode, fa(footprint), out)
    del keep
    return np.array(out[:], np.float32) if ok else None


def pyramid(texels, half=False):
    """The texture's mip levels, each an H x W x 3 float32 array, finest first."""
    t, keep = _texture(texels)
    h, w = keep.shape[:2]
    cap = h * w * 3 * 2 + 16 * 3
    out = np.zeros(cap, np.float32)
    sizes = (u32 * 32)()
    n = load().oracle_pyramid(ctypes.addressof(t), 1 if half else 0, out.ctypes.data_as(P(f32)), cap, sizes)
    levels, at = [], 0
    for k in range(n):
        lw, lh = sizes[2 * k], sizes[2 * k + 1]
        levels.append(out[at:at + lw * lh * 3].reshape(lh, lw, 3).copy())
        at += lw * lh * 3
    return levels


class OracleScene:
    """The oracle's copy of a scene (with its own BVH)."""

    def __init__(self, scene_data):
        self.lib = load()
        self.h = self.lib.oracle_scene_create(ctypes.addressof(scene_data.c))
        if not self.h:
            raise RuntimeError("oracle_scene_create failed")

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.oracle_scene_destroy(self.h)
            self.h = None

    def render(self, config, threads=None, chunks=None):
        """The film (sums [pixels,4] f64, counts [pixels] u32) and counters (traces, shadows, errors)."""
        n = config.width * config.height
        sums = np.zeros((n, 4), np.float64)
        counts = np.zeros(n, np.uint32)
        counters = (u64 * 3)()
        threads = threads or os.cpu_count() or 1
        if chunks is None:
            self.lib.oracle_render(self.h, ctypes.addressof(config), threads,
                                   sums.ctypes.data_as(P(f64)), counts.ctypes.data_as(P(u32)), counters)
        else:
            if isinstance(chunks, list):  # a slice of a ctypes array is a list of records
                chunks = (type(chunks[0]) * len(chunks))(*chunks) if chunks else None
            if chunks is None:
                return sums, counts, (0, 0, 0)
            self.lib.oracle_render_chunks(self.h, ctypes.addressof(config), ctypes.addressof(chunks), len(chunks),
                                          threads, sums.ctypes.data_as(P(f64)), counts.ctypes.data_as(P(u32)),
                                          counters)
        return sums, counts, tuple(counters)

    def render_into(self, config, chunks, sums, counts, threads=1):
        """Adds the chunks' samples (a ctypes array of abi.Chunk) into the caller's film (sums [pixels, 4] float64,
        counts [pixels] uint32, both C-contiguous); returns the counters (traces, shadows, errors)."""
        assert sums.dtype == np.float64 and counts.dtype == np.uint32 and sums.flags.c_contiguous
        assert sums.shape == (config.width * config.height, 4) and counts.shape == (config.width * config.height,)
        counters = (u64 * 3)()
        if len(chunks):
            self.lib.oracle_render_chunks(self.h, ctypes.addressof(config), ctypes.addressof(chunks), len(chunks),
                                          threads, sums.ctypes.data_as(P(f64)), counts.ctypes.data_as(P(u32)),
                                          counters)
        return tuple(counters)

    def adjoint(self, config, d_image, texture_floats, material_count, threads=None):
        """FP64 gradients (textures concatenated, roughness per material) for the image adjoint d_image [pixels, 3]."""
        d = np.ascontiguousarray(d_image, dtype=np.float32).reshape(-1)
        gt = np.zeros(max(1, texture_floats), np.float64)
        gr = np.zeros(material_count, np.float64)
        counters = (u64 * 3)()
        self.lib.oracle_adjoint(self.h, ctypes.addressof(config), threads or os.cpu_count() or 1,
                                d.ctypes.data_as(P(f32)), gt.ctypes.data_as(P(f64)), gr.ctypes.data_as(P(f64)), counters)
        return gt[:texture_floats], gr, tuple(counters)

    def set_texture(self, index, texels):
        t = np.ascontiguousarray(texels, dtype=np.float32)
        self.lib.oracle_set_texture(self.h, index, t.ctypes.data_as(P(f32)))

    def set_texture_half(self, index, half=True):
        """Every level of the texture rounded to half floats, as paged tiles hold them."""
        self.lib.oracle_set_texture_half(self.h, index, 1 if half else 0)

    def set_texture_coarse(self, index, level):
        """Fetches finer than `level` read their ancestor there (16: off)."""
        self.lib.oracle_set_texture_coarse(self.h, index, level)

    def set_material(self, index, material):
        self.lib.oracle_set_material(self.h, index, ctypes.addressof(material))

    def set_eval_roughness(self, material, r):
        self.lib.oracle_set_eval_roughness(self.h, material, r)

    def render_pixels(self, config, first, count, sums, odd, counts, threads=None):
        """Adds pixel p's samples first[p] .. first[p] + count[p] - 1 into the film (sums, odd-index sums, counts)."""
        counters = (u64 * 3)()
        f = np.ascontiguousarray(first, np.uint32)
        n = np.ascontiguousarray(count, np.uint32)
        self.lib.oracle_render_pixels(self.h, ctypes.addressof(config), f.ctypes.data_as(P(u32)),
                                      n.ctypes.data_as(P(u32)), threads or os.cpu_count() or 1,
                                      sums.ctypes.data_as(P(f64)), odd.ctypes.data_as(P(f64)),
                                      counts.ctypes.data_as(P(u32)), counters)
        return tuple(counters)

    def sample(self, config, pixel, index):
        rgba = (f64 * 4)()
        counters = (u64 * 2)()
        st = self.lib.oracle_sample(self.h, ctypes.addressof(config), pixel, index, rgba, counters)
        return st, tuple(rgba), tuple(counters)

    def first_hit_draws(self, config, pixel, index, capacity=512):
        """The stratified draws of the sample's first non-mirror vertex: rows (kind 0 light / 1 lobe, which, number,
        r0, r1) in the order the estimator drew them."""
        out = (f32 * (5 * capacity))()
        n = u32()
        st = self.lib.oracle_first_hit_draws(self.h, ctypes.addressof(config), pixel, index, out, capacity,
                                             ctypes.byref(n))
        assert n.value <= capacity
        return st, np.array(out[: 5 * n.value], np.float64).reshape(-1, 5)

    def closest(self, o, d, near=0.0, far=3.4e38, brute=False):
        g, p, t = u32(), u32(), f32()
        self.lib.oracle_closest(self.h, fa(o), fa(d), near, far, 1 if brute else 0, ctypes.byref(g),
                                ctypes.byref(p), ctypes.byref(t))
        return g.value, p.value, t.value

    def occluded(self, o, d, near, far, brute=False):
        return bool(self.lib.oracle_occluded(self.h, fa(o), fa(d), near, far, 1 if brute else 0))
