"""Loader for oracle/_ref/libref.so: the reference's own component headers compiled in place (oracle/ref_driver.cpp).
Absent on machines without the reference tree at build time; tests that need it skip then."""
import ctypes
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PATH = os.path.join(ROOT, "oracle", "_ref", "libref.so")
P = ctypes.POINTER
u32, u64, f32, c_int, vp = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_float, ctypes.c_int, ctypes.c_void_p


def load():
    if not os.path.exists(PATH):
        pytest.skip("oracle/_ref/libref.so not built (the reference tree was absent at build time)")
    lib = ctypes.CDLL(PATH)
    sig = {
        "ref_philox": (None, [P(u32), P(u32), P(u32)]),
        "ref_block": (None, [u32, u64, u64, u64, u32, P(u32)]),
        "ref_sampling_word": (u32, [u32, u64, u64, u64, u32, u32]),
        "ref_sobol": (u32, [u32, u32]),
        "ref_owen_scramble": (u32, [u32, u32]),
        "ref_unit": (f32, [u32]),
        "ref_directions": (None, [P(u32)]),
        "ref_frame_seed": (u32, [u32, u32, c_int]),
        "ref_sincos": (None, [f32, P(f32), P(f32)]),
        "ref_camera_ray": (c_int, [vp, f32, f32, P(f32), P(f32), P(f32), P(f32)]),
        "ref_box_filter": (c_int, [f32, f32, P(f32)]),
        "ref_geometry_hit": (c_int, [P(f32), P(f32), P(f32), P(f32), P(f32), P(f32)]),
        "ref_light_sample": (c_int, [vp, P(f32), P(f32), P(f32), P(f32), P(f32), P(f32), P(f32)]),
        "ref_light_intersect": (c_int, [vp, P(f32), P(f32), P(f32), P(f32), P(f32)]),
        "ref_light_can_illuminate": (c_int, [vp, P(f32), P(f32)]),
        "ref_lobe_eval": (None, [u32, P(f32), P(f32), P(f32), P(f32), f32, P(f32), P(f32)]),
        "ref_lobe_sample": (None, [u32, P(f32), P(f32), P(f32), f32, f32, f32, P(f32), P(f32), P(f32)]),
        "ref_bilinear": (c_int, [u32, u32, P(f32), f32, f32, P(f32)]),
        "ref_surface_lambert": (c_int, [P(f32), P(f32), f32, f32, P(f32), P(f32), P(f32)]),
        "ref_map_lookup": (c_int, [u32, P(u32), P(f32), u32, u32, P(f32), P(f32)]),
        "ref_footprint": (c_int, [P(f32)] * 7 + [f32, f32, u32, P(f32)]),
        "ref_hit_tex_derivatives": (None, [P(f32)] * 8 + [u32, P(f32)]),
        "ref_transfer": (c_int, [P(f32), c_int] + [P(f32)] * 7),
        "ref_triangle_partials": (c_int, [P(f32)] * 6),
        "ref_hemisphere_differentials": (None, [f32, f32, P(f32)]),
        "ref_scatter_differentials": (c_int, [u32, P(f32), f32, P(f32), P(f32), c_int, P(f32), P(f32), P(f32), f32, f32,
                                              P(f32), f32, P(f32)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib
