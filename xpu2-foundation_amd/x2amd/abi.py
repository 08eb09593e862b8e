"""ctypes mirror of include/x2dev.h: the boundary records and the device library's entry points.

The structures are field-for-field the C ones; sizes are asserted against the C declarations at import so a layout
drift fails loudly. The library is the in-tree ``libx2dev.so`` built by ``__graft_entry__.build()``.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libx2dev.so")

X2_OK = 0
X2_PASS_SHIFT = 16
X2_PASS_SLOTS = 4
X2_NO_SLOT = 0xFFFFFFFF
X2_LIGHT_ENV = 0
X2_LIGHT_RECT = 1
X2_SIDE_REGULAR = 0
X2_SIDE_REVERSE = 1
X2_SIDE_TWO = 2
X2_ERR_FILE = -6
X2_EXR_NONE = 0
X2_EXR_ZIPS = 2
X2_EXR_ZIP = 3
X2_DEVICE_CPU = 0
X2_DEVICE_GPU = 1
X2_NO_CELL = 0xFFFFFFFF
X2_STATIC_MAX_CHUNKS = 65536
NODE_BYTES = 128  # X2_NODE_BYTES: a BVH8 node in one cache line
TRIANGLE_BYTES = 48  # X2_TRIANGLE_BYTES: a triangle record
LEAF_TRIANGLE_BYTES = 36  # X2_LEAF_TRIANGLE_BYTES: the vertices a leaf test reads
FILM_BYTES_PER_PATH = 72
VOXEL_TAP_BYTES = 16  # X2_VOXEL_TAP_BYTES: a trilinear tap set of one grid (8 FP16 voxels)

u32 = ctypes.c_uint32
u64 = ctypes.c_uint64
i32 = ctypes.c_int32
f32 = ctypes.c_float
f64 = ctypes.c_double


class Chunk(ctypes.Structure):
    _fields_ = [("tile", u32), ("first", u32), ("count", u32), ("cell", u32)]


class Parked(ctypes.Structure):
    _fields_ = [("pixel", u32), ("index", u32), ("tries", u32), ("pass_", u32)]


class Refine(ctypes.Structure):  # x2_refine (Refine, async_integrator.h:3069-3073)
    _fields_ = [("pixel", u32), ("index", u32), ("pass_", u32), ("tile", u32), ("rgba", f32 * 4)]


class Light(ctypes.Structure):
    _fields_ = [
        ("type", u32), ("sidedness", u32), ("hemispherical", u32), ("visibility_mask", u32),
        ("radiance", f32 * 3), ("inv_area", f32),
        ("position", f32 * 3), ("plane_offset", f32),
        ("direction", f32 * 3), ("scale", f32),
        ("local_to_render", f32 * 9), ("half_width", f32),
        ("render_to_local", f32 * 9), ("half_height", f32),
        ("render_to_local_rot", f32 * 9), ("pad", f32),
    ]


class Material(ctypes.Structure):
    _fields_ = [
        ("albedo", f32 * 3), ("roughness", f32), ("diffuse_weight", f32), ("specular_weight", f32),
        ("emission", f32 * 3), ("albedo_map", u32),
    ]


class Texture(ctypes.Structure):
    _fields_ = [("width", u32), ("height", u32), ("texels", ctypes.POINTER(f32))]


class Volume(ctypes.Structure):
    _fields_ = [
        ("bounds_min", f32 * 3), ("bounds_max", f32 * 3), ("res", u32 * 3), ("sigma_t_scale", f32),
        ("albedo", f32 * 3), ("g", f32), ("emission_scale", f32), ("ramp_max_kelvin", f32), ("ramp_size", u32),
        ("pad", u32), ("density", ctypes.POINTER(ctypes.c_uint16)), ("temperature", ctypes.POINTER(ctypes.c_uint16)),
        ("ramp", ctypes.POINTER(f32)),
    ]


X2_GEOM_SINGLE_SIDED = 1  # x2_scene.geom_flags: the mesh's back faces are culled


class Scene(ctypes.Structure):
    _fields_ = [
        ("positions", ctypes.POINTER(f32)), ("indices", ctypes.POINTER(u32)),
        ("geom_first", ctypes.POINTER(u32)), ("geom_material", ctypes.POINTER(u32)),
        ("materials", ctypes.POINTER(Material)), ("lights", ctypes.POINTER(Light)),
        ("vertex_count", u32), ("triangle_count", u32), ("geom_count", u32), ("material_count", u32),
        ("light_count", u32), ("pad", u32),
        ("uvs", ctypes.POINTER(f32)), ("textures", ctypes.POINTER(Texture)), ("texture_count", u32), ("pad2", u32),
        ("volumes", ctypes.POINTER(Volume)), ("volume_count", u32), ("pad3", u32),
        ("geom_flags", ctypes.POINTER(u32)),
    ]


class Config(ctypes.Structure):
    _fields_ = [
        ("width", u32), ("height", u32), ("spp", u32), ("seed", u32),
        ("sampling", u32), ("filter", u32), ("filter_width", f32), ("near_clip", f32), ("far_clip", f32),
        ("raster_to_camera", f32 * 16), ("camera_to_render", f32 * 16),
        ("light_count", u32), ("bsdf_samples", u32), ("light_samples", u32),
        ("max_depth", u32), ("max_diffuse_depth", u32), ("max_glossy_depth", u32), ("max_mirror_depth", u32),
        ("clamp_depth", u32), ("clamp_value", f32), ("rr_threshold", f32), ("rr_min_depth", u32),
        ("shadow_reuse", u32), ("viewport", i32 * 4), ("pad", u32 * 2),
        ("max_volume_depth", u32), ("volume_illumination_samples", u32), ("volume_transmittance_threshold", f32),
        ("inv_volume_quality", f32), ("volume_attenuation_factor", f32), ("volume_contribution_factor", f32),
        ("volume_phase_attenuation_factor", f32), ("pad_volume", u32),
        ("direct_only", u32), ("surface_limit", u32), ("sky", f32 * 3), ("pad_direct", u32),
    ]


class Completion(ctypes.Structure):
    _fields_ = [
        ("chunks", u64), ("items", u64), ("samples", u64), ("parked_run", u64), ("refines_run", u64),
        ("errors", u32), ("pad", u32),
        ("device_seconds", f64), ("rays_trace", u64), ("rays_shadow", u64), ("nodes_visited", u64),
        ("triangles_tested", u64), ("voxel_taps", u64),
    ]


class Options(ctypes.Structure):
    _fields_ = [("device", i32), ("max_chunks_per_launch", u32), ("ring_chunks", u32), ("counters", u32),
                ("wave_slots", u32), ("lane", u32)]


class FrameStats(ctypes.Structure):
    _fields_ = [
        ("launches", u64), ("chunks", u64), ("samples", u64), ("errors", u32), ("audit_failures", u32),
        ("device_seconds", f64), ("wall_seconds", f64), ("rays_trace", u64), ("rays_shadow", u64),
        ("nodes_visited", u64), ("triangles_tested", u64), ("voxel_taps", u64),
    ]


class Adaptive(ctypes.Structure):
    _fields_ = [("min_samples", u32), ("max_samples", u32), ("target_error", f32), ("passes", u32),
                ("tree_updates", u32), ("device_chain", u32),
                ("cellular", u32), ("sample_budget", u32), ("noise_floor", f32), ("pad", u32)]


class RngQuery(ctypes.Structure):
    _fields_ = [("pixel", u32), ("dimension", u32), ("policy", u32), ("bounce", u32), ("index", u64), ("stream", u64),
                ("branch", u64), ("a", u32), ("b", u32)]


class BatchTuning(ctypes.Structure):
    _fields_ = [("block_threads", u32), ("claim_items", u32), ("waves_per_simd", u32), ("pad", u32),
                ("seconds", f64)]


class BatchStats(ctypes.Structure):
    _fields_ = [
        ("wall_seconds", f64), ("device_seconds", f64), ("paths", u64), ("rays_trace", u64), ("rays_shadow", u64),
        ("jobs", u32), ("errors", u32), ("launches", u32), ("graph", u32), ("tuning", BatchTuning),
    ]


class GridOptions(ctypes.Structure):
    _fields_ = [("cell_size", u32), ("samples_per_chunk", u32), ("mode", i32), ("progressive", u32),
                ("sample_first", u32), ("sample_count", u32)]


CPU_RENDER = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, u32, ctypes.POINTER(Chunk))


class CpuLane(ctypes.Structure):
    _fields_ = [("threads", u32), ("pad", u32), ("context", ctypes.c_void_p), ("render", CPU_RENDER)]


class BrokerOptions(ctypes.Structure):
    _fields_ = [("window_ms", f64), ("first_chunks", u32), ("gpu_workers", u32)]


class GridFrame(ctypes.Structure):
    _fields_ = [
        ("gpu_chunks", u64), ("cpu_chunks", u64), ("gpu_samples", u64), ("cpu_samples", u64), ("gpu_stolen", u64),
        ("cpu_stolen", u64), ("launches", u64), ("errors", u64), ("seconds", f64), ("gpu_rate", f64),
        ("cpu_rate", f64), ("rays_trace", u64), ("rays_shadow", u64), ("nodes_visited", u64),
        ("triangles_tested", u64),
    ]


class FilmFile(ctypes.Structure):
    _fields_ = [("width", u32), ("height", u32), ("spp", u32), ("done", u32), ("sums", ctypes.POINTER(f64)),
                ("counts", ctypes.POINTER(u32)), ("odd", ctypes.POINTER(f64))]


class ExrImage(ctypes.Structure):
    _fields_ = [("width", u32), ("height", u32), ("sums", ctypes.POINTER(f64)), ("counts", ctypes.POINTER(u32)),
                ("region", i32 * 4), ("half", u32), ("compression", u32), ("weight", u32), ("pad", u32)]


class Staged(ctypes.Structure):
    _fields_ = [("path", ctypes.c_char_p), ("stage_samples", u32), ("resume", u32), ("max_stages", u32), ("pad", u32)]


class StagedFrame(ctypes.Structure):
    _fields_ = [("stages", u32), ("saves", u32), ("launches", u32), ("errors", u32), ("resumed", u32),
                ("resumed_from", u32), ("done", u32), ("pad", u32), ("seconds", f64), ("device_seconds", f64)]


class Paging(ctypes.Structure):
    _fields_ = [("pool_tiles", u32), ("coarse_side", u32), ("partial", u32), ("host_refines", u32)]


class PagingStats(ctypes.Structure):
    _fields_ = [("parked_outstanding", u64), ("parked_total", u64), ("provisional", u64), ("misses", u64),
                ("uploads", u64), ("evictions", u64), ("overflow", u64), ("tiles", u32), ("pool_tiles", u32),
                ("resident", u32), ("pad", u32)]


# Sizes of the C structures (include/x2dev.h); checked against ctypes here and against the C compiler by a
# CPU test that compiles the header.
EXPECTED_SIZES = {
    "x2_chunk": (Chunk, 16), "x2_parked": (Parked, 16), "x2_refine": (Refine, 32), "x2_light": (Light, 184),
    "x2_material": (Material, 40), "x2_texture": (Texture, 16), "x2_volume": (Volume, 96), "x2_scene": (Scene, 120),
    "x2_config": (Config, 292),
    "x2_completion": (Completion, 96), "x2dev_options": (Options, 24), "x2_frame_stats": (FrameStats, 88),
    "x2_batch_tuning": (BatchTuning, 24), "x2_batch_stats": (BatchStats, 80), "x2_rng_query": (RngQuery, 48), "x2_adaptive": (Adaptive, 40),
    "x2_grid_options": (GridOptions, 24), "x2_cpu_lane": (CpuLane, 24), "x2_broker_options": (BrokerOptions, 16),
    "x2_grid_frame": (GridFrame, 120), "x2_film_file": (FilmFile, 40), "x2_exr_image": (ExrImage, 56),
    "x2_staged": (Staged, 24), "x2_staged_frame": (StagedFrame, 48), "x2_paging": (Paging, 16),
    "x2_paging_stats": (PagingStats, 72),
}
for _name, (_cls, _size) in EXPECTED_SIZES.items():
    assert ctypes.sizeof(_cls) == _size, (_name, ctypes.sizeof(_cls), _size)

# Every entry point include/x2dev.h declares, with its ctypes signature.
SIGNATURES = {
    "x2dev_create": (ctypes.c_int, [ctypes.POINTER(Options), ctypes.POINTER(ctypes.c_void_p)]),
    "x2dev_prepare": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Scene)]),
    "x2dev_begin": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Config), u32]),
    "x2dev_ready": (ctypes.c_int, [ctypes.c_void_p]),
    "x2dev_launch": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Chunk), u32, u64]),
    "x2dev_launch_refines": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Chunk), u32, ctypes.POINTER(Refine), u32, u64]),
    "x2dev_take_refines": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Refine), u32, ctypes.POINTER(u32)]),
    "x2dev_refines_waiting": (u64, [ctypes.c_void_p]),
    "x2dev_poll": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Completion)]),
    "x2dev_take_deferred": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Parked), u32, ctypes.POINTER(u32)]),
    "x2dev_download_film": (ctypes.c_int, [ctypes.c_void_p]),
    "x2dev_film_ready": (ctypes.c_int, [ctypes.c_void_p]),
    "x2dev_film": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.POINTER(f64)),
                                  ctypes.POINTER(ctypes.POINTER(u32))]),
    "x2dev_in_flight": (u32, [ctypes.c_void_p]),
    "x2dev_max_chunks_per_launch": (u32, [ctypes.c_void_p]),
    "x2dev_launch_threads": (u32, [ctypes.c_void_p]),
    "x2dev_device_film": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p),
                                         ctypes.POINTER(ctypes.c_void_p)]),
    "x2dev_destroy": (None, [ctypes.c_void_p]),
    "x2dev_last_error": (ctypes.c_char_p, []),
    "x2dev_render_frame": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Config), u32,
                                          ctypes.POINTER(FrameStats)]),
    "x2dev_set_texture": (ctypes.c_int, [ctypes.c_void_p, u32, ctypes.c_void_p, ctypes.c_int]),
    "x2dev_set_material": (ctypes.c_int, [ctypes.c_void_p, u32, ctypes.POINTER(Material)]),
    "x2dev_resolve_image": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "x2dev_adjoint_frame": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Config), u32, ctypes.c_void_p,
                                           ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(FrameStats)]),
    "x2dev_rng_twin": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(RngQuery), u32, u32, ctypes.POINTER(u32),
                                      ctypes.POINTER(f32), ctypes.POINTER(u64)]),
    "x2dev_begin_adaptive": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Config), u32]),
    "x2dev_upload_pass": (ctypes.c_int, [ctypes.c_void_p, u32, ctypes.POINTER(u64), u32]),
    "x2dev_request_errors": (ctypes.c_int, [ctypes.c_void_p]),
    "x2dev_errors_ready": (ctypes.c_int, [ctypes.c_void_p]),
    "x2dev_errors": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.POINTER(f32))]),
    "x2dev_seed_film": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(f64), ctypes.POINTER(f64), ctypes.POINTER(u32)]),
    "x2dev_odd": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.POINTER(f64))]),
    "x2dev_render_adaptive": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Config), u32, ctypes.POINTER(Adaptive),
                                             ctypes.POINTER(FrameStats)]),
    "x2_adaptive_masks": (ctypes.c_int, [u32, u32, f32, ctypes.POINTER(f32), ctypes.POINTER(u64), ctypes.POINTER(u32)]),
    "x2_cellular_step": (ctypes.c_int, [u32, u32, f32, ctypes.POINTER(ctypes.c_uint8), ctypes.POINTER(u64),
                                        ctypes.POINTER(f32), ctypes.POINTER(u32), f64, u32, f32,
                                        ctypes.POINTER(ctypes.c_uint8), ctypes.POINTER(f32)]),
    "x2_pixel_errors": (ctypes.c_int, [ctypes.POINTER(f64), ctypes.POINTER(f64), ctypes.POINTER(u32), u32,
                                       ctypes.POINTER(f32)]),
    "x2batch_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    "x2batch_add_job": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Scene), ctypes.POINTER(Config),
                                       ctypes.POINTER(u32)]),
    "x2batch_tune": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(BatchTuning)]),
    "x2batch_set_tuning": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(BatchTuning)]),
    "x2batch_run": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(f64), ctypes.POINTER(BatchStats)]),
    "x2batch_job_film": (ctypes.c_int, [ctypes.c_void_p, u32, ctypes.POINTER(ctypes.POINTER(f64)),
                                        ctypes.POINTER(ctypes.POINTER(u32))]),
    "x2batch_job_count": (u32, [ctypes.c_void_p]),
    "x2batch_destroy": (None, [ctypes.c_void_p]),
    "x2_grid_create": (ctypes.c_int, [u32, u32, u32, ctypes.POINTER(GridOptions), ctypes.POINTER(ctypes.c_uint8), u32,
                                      ctypes.c_int, ctypes.POINTER(i32), ctypes.POINTER(ctypes.c_void_p)]),
    "x2_grid_destroy": (None, [ctypes.c_void_p]),
    "x2_grid_cell_count": (u32, [ctypes.c_void_p]),
    "x2_grid_total_chunks": (u64, [ctypes.c_void_p]),
    "x2_grid_owner": (u32, [ctypes.c_void_p, u32]),
    "x2_grid_remaining": (u32, [ctypes.c_void_p, u32]),
    "x2_grid_claimable": (u32, [ctypes.c_void_p, u32, u32]),
    "x2_grid_empty": (ctypes.c_int, [ctypes.c_void_p]),
    "x2_grid_claim": (u32, [ctypes.c_void_p, u32, u32, ctypes.POINTER(Chunk), u32]),
    "x2_grid_append": (ctypes.c_int, [ctypes.c_void_p, u32, ctypes.POINTER(Chunk), u32]),
    "x2_grid_choose": (u32, [ctypes.c_void_p, u32, u32, u32]),
    "x2_grid_seed": (u32, [ctypes.c_void_p, u32, u32, u32]),
    "x2_grid_frontier": (u32, [ctypes.c_void_p, u32]),
    "x2_grid_ranges": (u32, [ctypes.c_void_p, ctypes.POINTER(u32), u32]),
    "x2_grid_set_scores": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(f32), u32]),
    "x2_grid_important": (ctypes.c_int, [ctypes.c_void_p, u32]),
    "x2_grid_reserved": (ctypes.c_int, [ctypes.c_void_p, u32]),
    "x2_grid_set_suitability": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(f32), u32]),
    "x2_grid_set_rates": (None, [ctypes.c_void_p, f64, f64, u32, f64]),
    "x2_grid_gpu_claimed": (None, [ctypes.c_void_p]),
    "x2_grid_passage": (ctypes.c_int, [ctypes.c_void_p, u32, u32]),
    "x2_grid_close": (None, [ctypes.c_void_p]),
    "x2_grid_rendered": (None, [ctypes.c_void_p, u32, u32, u64]),
    "x2_grid_rendered_by": (u64, [ctypes.c_void_p, u32, u32]),
    "x2_grid_deferred_to_cpu": (None, [ctypes.c_void_p, u32]),
    "x2_grid_cell_of": (u32, [ctypes.c_void_p, u32]),
    "x2_grid_tile_pixels": (u32, [ctypes.c_void_p, u32]),
    "x2_ownership_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p)]),
    "x2_ownership_destroy": (None, [ctypes.c_void_p]),
    "x2_ownership_owners": (ctypes.c_int, [ctypes.c_void_p, u32, ctypes.POINTER(ctypes.c_uint8)]),
    "x2_ownership_finished": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "x2dev_render_grid": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Config), ctypes.c_void_p,
                                         ctypes.POINTER(CpuLane), ctypes.POINTER(BrokerOptions),
                                         ctypes.POINTER(GridFrame)]),
    "x2dev_chunk_room": (u32, [ctypes.c_void_p]),
    "x2dev_render_chunks_static": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Config), u32, ctypes.POINTER(Chunk),
                                                  u32, ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(Completion)]),
    "x2dev_page_textures": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Paging)]),
    "x2dev_paging_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(PagingStats)]),
    "x2dev_parked_outstanding": (u64, [ctypes.c_void_p]),
    "x2dev_provisional_film": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.POINTER(f64)),
                                              ctypes.POINTER(ctypes.POINTER(u32))]),
    "x2_film_save": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(FilmFile)]),
    "x2_film_load": (ctypes.c_int, [ctypes.c_char_p, u32, u32, u32, ctypes.POINTER(f64), ctypes.POINTER(u32),
                                    ctypes.POINTER(f64), ctypes.POINTER(u32), ctypes.POINTER(u32)]),
    "x2_write_exr": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(ExrImage)]),
    "x2dev_render_staged": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Config), u32, ctypes.POINTER(Staged),
                                           ctypes.POINTER(StagedFrame)]),
}

_lib = None


def load(path=None):
    """Loads the device library (LIB_PATH unless a path is given); raises if it is not built (there is no fallback
    path)."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise RuntimeError(f"x2dev library not built: {path} (run __graft_entry__.build())")
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(code, what):
    if code < 0:
        err = load().x2dev_last_error()
        raise RuntimeError(f"{what} failed ({code}): {err.decode() if err else ''}")
    return code
