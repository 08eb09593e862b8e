"""Benchmark: the path-tracing device on one MI355X.

Workload (the headline configuration): 2,621,442 triangles (32 fBm-displaced level-6 icospheres over a ground quad),
50/50 Lambert + GGX materials, 3 rectangle lights + a constant environment, 1920x1080 at 64 spp, depth 12. A step is
one uniform frame through the broker counterpart (x2dev_render_frame): 132,710,400 paths, chunks launched at most two
in flight (up to LAUNCH_CHUNKS chunks a launch: fewer, larger launches leave fewer idle lanes at launch ends), the
film downloaded and every pixel's count audited. The scene is resident in HBM before timing starts.

Prints one JSON line: Mpaths/s (value), Mrays/s, the roofline of the render kernel (algorithmic bytes: 128 B per
BVH8 node fetched (one cache line) + 36 B per triangle tested (its vertices) + 72 B of film per path, over the kernel's own
HIP-event time)
and the CPU baseline (the oracle on this host's cores over a bounded sample of the same workload). The adjoint
leg times one forward frame and one path-replay adjoint frame of the differentiable Cornell box (512x512, 32 spp):
adjoint_step_over_forward is a gradient step's time (the forward frame + the adjoint frame) over the forward frame's.
The volume leg renders the
participating-media workload (a 256^3 brick-sparse cloud at 1280x720, 32 spp); the batch leg the 64-job batched
workload, persistent kernel from a hipGraph against one launch per job; the direct leg the prototype estimator on
the headline scene; the adaptive leg an adaptive frame of the headline scene through the device pass chain and
through the host's trees; the grid leg the headline frame from the work grid through the broker; the paging leg the
textured Cornell box with its texture resident and in tiles on demand; the checkpoint leg the headline frame in
stages with its film saved after each, and its EXR. roofline.copy_ceiling_gbs is a measured device copy's bandwidth beside the published peak.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "xpu2-foundation_amd"))

HBM_PEAK_GBS = 8000.0


LAUNCH_CHUNKS = 131072  # max_chunks_per_launch of every leg's lane (the reference's default is 16384)


def make_lane(wave_slots=0):
    from x2amd.lane import GpuLane
    return GpuLane(0, max_chunks_per_launch=LAUNCH_CHUNKS, ring_chunks=2 * LAUNCH_CHUNKS, wave_slots=wave_slots)


def timed_cpu(scene, config, runs=3):
    """The oracle (a CPU restatement of the reference's integrator, -O3, no contraction) rendering `config` on every
    host thread (one worker a thread claiming 8x8 tiles, as the reference's CPU lane): one warm-up, then the median of
    `runs` (the reference's protocol). Returns (seconds, rays, cores)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    cores = len(os.sched_getaffinity(0))
    o = oracle.OracleScene(scene)
    o.render(config, threads=cores)
    times, rays = [], 0
    for _ in range(runs):
        t = time.perf_counter()
        _, _, k = o.render(config, threads=cores)
        times.append(time.perf_counter() - t)
        rays = k[0] + k[1]
    times.sort()
    return times[len(times) // 2], rays, cores


def cpu_baseline(scene, width, height, spp):
    """The headline scene on the host's cores at width x height x spp (a bounded sample: the rate is per path)."""
    from x2amd import scenes

    _, cfg = scenes.icospheres(width=width, height=height, spp=spp)
    dt, rays, cores = timed_cpu(scene, cfg)
    paths = width * height * spp
    return {
        "value": paths / dt / 1e6, "unit": "Mpaths/s", "cores": cores, "kind": "port",
        "sample": f"{width}x{height} at {spp} spp ({paths} paths) of the same scene, median of 3 after a warm-up, "
                  f"{dt:.2f} s each",
        "mrays_per_s": rays / dt / 1e6,
    }


def cpu_other_configs():
    """The CPU oracle on the other configurations' workloads: the Cornell box (256x256, 16 spp: the whole frame) and
    the differentiable Cornell box's forward frame (512x512, 32 spp, RR off) sampled at 2 spp."""
    import numpy as np  # noqa: F401
    from x2amd import scenes

    out = {}
    scene, cfg = scenes.cornell()
    dt, rays, cores = timed_cpu(scene, cfg)
    paths = cfg.width * cfg.height * cfg.spp
    out["cornell_256_16spp"] = {"mpaths_per_s": paths / dt / 1e6, "mrays_per_s": rays / dt / 1e6, "cores": cores,
                                "kind": "port", "sample": f"the whole frame ({paths} paths), median of 3, {dt:.3f} s"}
    scene, cfg, _ = scenes.cornell_textured(spp=2)
    dt, rays, cores = timed_cpu(scene, cfg)
    paths = cfg.width * cfg.height * cfg.spp
    out["cornell_textured_512"] = {"mpaths_per_s": paths / dt / 1e6, "mrays_per_s": rays / dt / 1e6, "cores": cores,
                                   "kind": "port",
                                   "sample": f"512x512 at 2 spp ({paths} paths) of the 32-spp forward frame, median of "
                                             f"3, {dt:.2f} s"}
    return out


def roofline_bytes(stats, paths, per_path):
    """Algorithmic bytes of a frame's launches: BVH8 nodes fetched (one 128-B line each), triangles tested (their
    36-B packed vertices), the medium's tap sets (16 B: 8 FP16 voxels), and per_path bytes a path."""
    from x2amd import abi
    return (abi.NODE_BYTES * stats.nodes_visited + abi.LEAF_TRIANGLE_BYTES * stats.triangles_tested +
            abi.VOXEL_TAP_BYTES * stats.voxel_taps + per_path * paths)


def adjoint_leg(reps):
    """The differentiable Cornell box (512x512, 32 spp, depth 8, 512x512 RGB FP32 albedo map, 4 GGX materials): one
    forward frame against one path-replay adjoint frame of an L2 image loss (dLoss/dimage resident on the device),
    each timed over `reps` frames after a warm-up."""
    import numpy as np
    from x2amd import abi, scenes
    from x2amd.hipmem import DeviceBuffer

    scene, config, params = scenes.cornell_textured()
    lane = make_lane()
    lane.prepare(scene)
    rng = np.random.default_rng(1234)
    lane.set_texture(0, np.clip(params["texture"] + rng.normal(0, 0.1, params["texture"].shape), 0.01, 1.0))
    for m, r in zip(params["roughness_materials"], params["roughness"] + rng.normal(0, 0.1, 4)):
        mat = abi.Material.from_buffer_copy(scene.materials[m])
        mat.roughness = float(np.clip(r, 0.05, 0.95))
        lane.set_material(m, mat)
    sums, counts, _ = lane.render_frame(config)
    target = sums[:, :3] / counts[:, None]
    lane.set_texture(0, params["texture"])
    for m in params["roughness_materials"]:
        lane.set_material(m, scene.materials[m])
    sums, counts, _ = lane.render_frame(config)
    d_image = (2 * (sums[:, :3] / counts[:, None] - target)).astype(np.float32)
    dbuf, gt, gr = DeviceBuffer(d_image.nbytes), DeviceBuffer(params["texture"].nbytes), DeviceBuffer(4 * 64)
    dbuf.upload(d_image)
    paths = config.width * config.height * config.spp
    fw = dict(wall=0.0, kernel=0.0, traces=0, bytes=0)
    t0 = time.perf_counter()
    for _ in range(reps):
        _, _, st = lane.render_frame(config)
        fw["kernel"] += st.device_seconds
        fw["traces"] += st.rays_trace + st.rays_shadow
        fw["bytes"] += roofline_bytes(st, paths, 72 + 48)  # film + one texel tap (4 x 12 B) a path's camera hit
    fw["wall"] = time.perf_counter() - t0
    gt.zero()
    gr.zero()
    lane.adjoint_frame(config, dbuf.ptr, gt.ptr, gr.ptr)  # warm-up
    ad = dict(wall=0.0, kernel=0.0, traces=0, bytes=0)
    for _ in range(reps):
        gt.zero()
        gr.zero()
        t0 = time.perf_counter()
        st = lane.adjoint_frame(config, dbuf.ptr, gt.ptr, gr.ptr)
        ad["wall"] += time.perf_counter() - t0
        ad["kernel"] += st.device_seconds
        ad["traces"] += st.rays_trace + st.rays_shadow
        ad["bytes"] += roofline_bytes(st, paths, 12 + 48)  # the pixel's image adjoint + one texel tap a path
    grads = gt.download(np.float32, params["texture"].size)
    for b in (dbuf, gt, gr):
        b.free()
    lane.close()
    assert np.isfinite(grads).all() and np.abs(grads).sum() > 0
    fw_gbs, ad_gbs = fw["bytes"] / fw["kernel"] / 1e9, ad["bytes"] / ad["kernel"] / 1e9
    return {
        "workload": f"differentiable cornell {config.width}x{config.height}, {config.spp} spp, depth "
                    f"{config.max_depth}, {params['texture'].shape[1]}x{params['texture'].shape[0]} albedo map, "
                    "4 GGX roughness, L2 loss",
        "reps": reps,
        "forward_ms": fw["wall"] / reps * 1e3, "adjoint_ms": ad["wall"] / reps * 1e3,
        "forward_kernel_ms": fw["kernel"] / reps * 1e3, "adjoint_kernel_ms": ad["kernel"] / reps * 1e3,
        "adjoint_over_forward": ad["wall"] / fw["wall"],
        "step_over_forward": (fw["wall"] + ad["wall"]) / fw["wall"],
        "step_definition": "a gradient step = the forward frame (the image, for the loss) + the adjoint frame; "
                           "step_over_forward = (forward + adjoint) / forward, adjoint_over_forward = adjoint / forward",
        "roofline": {"bound": "hbm", "unit": "GB/s", "peak": HBM_PEAK_GBS,
                     "forward_achieved": fw_gbs, "forward_frac": fw_gbs / HBM_PEAK_GBS,
                     "adjoint_achieved": ad_gbs, "adjoint_frac": ad_gbs / HBM_PEAK_GBS,
                     "counted": "128 B a BVH8 node, 36 B a tested triangle, a path's film (72 B) or image adjoint "
                                "(12 B) and one bilinear texel tap (48 B); the adjoint's texel-gradient atomics are "
                                "not counted"},
        "forward_mpaths_s": paths * reps / fw["wall"] / 1e6,
        "adjoint_mpaths_s": paths * reps / ad["wall"] / 1e6,
        "rays_replayed_equal": fw["traces"] == ad["traces"],
    }


def volume_leg(reps, cpu):
    """Participating media: the cloud (256^3 brick-sparse FP16 density and temperature, about 31% of the 8^3 bricks
    occupied, extinction scale 4, HG g = 0.6, blackbody emission to 1500 K) over a ground quad under a rectangle light
    and a constant environment, 1280x720 at 32 spp, depth 16, volume depth 2; `reps` frames after a warm-up."""
    from x2amd import scenes

    scene, config = scenes.clouds()
    lane = make_lane()
    lane.prepare(scene)
    lane.render_frame(config)
    wall, kernel, rays, taps, nbytes = 0.0, 0.0, 0, 0, 0
    frame_paths = config.width * config.height * config.spp
    for _ in range(reps):
        t0 = time.perf_counter()
        _, counts, st = lane.render_frame(config)
        wall += time.perf_counter() - t0
        assert st.audit_failures == 0 and st.errors == 0
        kernel += st.device_seconds
        rays += st.rays_trace + st.rays_shadow
        taps += st.voxel_taps
        nbytes += roofline_bytes(st, frame_paths, 72)
    lane.close()
    paths = frame_paths * reps
    achieved = nbytes / kernel / 1e9
    out = {
        "workload": f"clouds {scene.volume.c.res[0]}^3 grid ({scene.volume.occupancy():.3f} of bricks occupied), "
                    f"{config.width}x{config.height}, {config.spp} spp, depth {config.max_depth}, volume depth "
                    f"{config.max_volume_depth}, rect + env lights",
        "reps": reps, "ms_per_frame": wall / reps * 1e3, "kernel_ms_per_frame": kernel / reps * 1e3,
        "mpaths_per_s": paths / wall / 1e6, "mrays_per_s": rays / wall / 1e6,
        "voxel_taps_per_path": taps / paths,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "kernel": "renderVolumeKernel",
                     "counted": "16 B a trilinear tap set (8 FP16 voxels of one grid), 128 B a BVH8 node, 36 B a "
                                "tested triangle, 72 B of film a path"},
    }
    if cpu:
        _, cpu_cfg = scenes.clouds(spp=16, volume=scene.volume)
        dt, cpu_rays, cores = timed_cpu(scene, cpu_cfg)
        n = cpu_cfg.width * cpu_cfg.height * cpu_cfg.spp
        out["cpu"] = {"mpaths_per_s": n / dt / 1e6, "mrays_per_s": cpu_rays / dt / 1e6, "cores": cores,
                      "kind": "port", "sample": f"{cpu_cfg.width}x{cpu_cfg.height} at {cpu_cfg.spp} spp ({n} paths) of "
                                                f"the same frame, median of 3 after a warm-up, {dt:.2f} s each"}
    return out


def batch_leg(reps):
    """Batched renders: 64 seeded icosphere jobs of log-uniform 10k-500k triangles, square images of log-uniform
    side 64-1024 and spp 1-64, rendered by one persistent kernel from a device queue replayed from a hipGraph
    (its shape chosen by the library's autotuner), against one launch per job in the same order."""
    import numpy as np
    from x2amd import scenes
    from x2amd.batch import Batch

    jobs = scenes.batch_jobs()
    b = Batch(0)
    t = time.perf_counter()
    for scene, config in jobs:
        b.add(scene, config)
    add_s = time.perf_counter() - t
    t = time.perf_counter()
    tuning = b.tune()
    tune_s = time.perf_counter() - t
    b.run()  # captures the graph
    paths = sum(c.width * c.height * c.spp for _, c in jobs)
    out = {"workload": f"{len(jobs)} jobs, {sum(s.triangle_count for s, _ in jobs)} triangles, "
                       f"{sum(c.width * c.height for _, c in jobs)} pixels, {paths} paths",
           "add_s": add_s, "tune_s": tune_s,
           "tuning": {"block_threads": tuning.block_threads, "claim_items": tuning.claim_items,
                      "waves_per_simd": tuning.waves_per_simd, "one_sample_pass_ms": tuning.seconds * 1e3}}
    for label, per_job in (("persistent_graph", False), ("per_job_launches", True)):
        wall = dev = 0.0
        secs = None
        for _ in range(reps):
            st, secs = b.run(per_job_launches=per_job)
            assert st.errors == 0
            wall += st.wall_seconds
            dev += st.device_seconds
        out[label] = {"mpaths_per_s": paths * reps / wall / 1e6, "device_mpaths_per_s": paths * reps / dev / 1e6,
                      "ms_per_batch": wall / reps * 1e3, "launches": int(st.launches),
                      "job_p50_ms": float(np.percentile(secs, 50) * 1e3),
                      "job_p99_ms": float(np.percentile(secs, 99) * 1e3)}
    b.close()
    out["speedup_vs_per_job"] = out["persistent_graph"]["mpaths_per_s"] / out["per_job_launches"]["mpaths_per_s"]
    return out


def direct_leg(reps):
    """The prototype estimator (integrateDirect: camera hit, emission, one cosine-sampled Lambert ray to a constant
    sky of 0.2) on the headline scene at 1920x1080 x 64 spp; the reference quotes its GPU lane rate on this
    estimator (400-530 M samples/s on other hardware and scenes, ASYNC_RESULTS.md:58)."""
    from x2amd import scenes

    scene, config = scenes.icospheres()
    config.direct_only, config.surface_limit = 1, 1
    config.sky[:] = [0.2, 0.2, 0.2]
    lane = make_lane()
    lane.prepare(scene)
    lane.render_frame(config)
    wall, rays = 0.0, 0
    for _ in range(reps):
        t0 = time.perf_counter()
        _, _, st = lane.render_frame(config)
        wall += time.perf_counter() - t0
        assert st.errors == 0 and st.audit_failures == 0
        rays += st.rays_trace + st.rays_shadow
    lane.close()
    paths = config.width * config.height * config.spp * reps
    return {"workload": f"icospheres {scene.triangle_count} tris, {config.width}x{config.height}, {config.spp} spp, "
                        "prototype estimator, surface limit 1, sky 0.2",
            "reps": reps, "ms_per_frame": wall / reps * 1e3, "mpaths_per_s": paths / wall / 1e6,
            "mrays_per_s": rays / wall / 1e6}


def adaptive_leg(reps, target=3e-4, min_samples=16, max_samples=64):
    """An adaptive frame of the headline scene (MoonRay's schedule to 64 spp, the passes below 16 samples uniform,
    region trees judged against target): the device pass chain (trees and masks on the device, every pass queued at
    once) against the host's trees between passes. samples_per_pixel is what the trees asked for; the uniform frame
    at 64 spp is the headline."""
    from x2amd import scenes

    scene, config = scenes.icospheres()
    lane = make_lane()
    lane.prepare(scene)
    lane.render_adaptive(config, min_samples, max_samples, target, device_chain=True)
    chain, device, spp = 0.0, 0.0, 0.0
    for _ in range(reps):
        t0 = time.perf_counter()
        _, _, counts, st, a = lane.render_adaptive(config, min_samples, max_samples, target, device_chain=True)
        chain += time.perf_counter() - t0
        device += st.device_seconds
        assert st.errors == 0 and st.audit_failures == 0
        spp = float(counts.mean())
    t0 = time.perf_counter()
    _, _, host_counts, hst, ha = lane.render_adaptive(config, min_samples, max_samples, target)
    host = time.perf_counter() - t0
    lane.close()
    samples = float(counts.sum())
    return {"workload": f"icospheres {scene.triangle_count} tris, {config.width}x{config.height}, adaptive to "
                        f"{max_samples} spp, uniform below {min_samples}, target error {target}",
            "reps": reps, "passes": a.passes, "samples_per_pixel": spp,
            "chain_ms_per_frame": chain / reps * 1e3, "chain_device_ms": device / reps * 1e3,
            "chain_mpaths_per_s": samples / (chain / reps) / 1e6,
            "host_trees_ms_per_frame": host * 1e3, "same_samples": bool((host_counts == counts).all())}


def grid_leg(reps, windows=(100.0, 10.0)):
    """The headline frame from the work grid through the broker (x2dev_render_grid, the GPU alone): the grid's
    progressive sample ranges (1, 1, 2, then 4) handed out by the frontier, the GPU's 32 cellular workers each taking a
    share of a launch, launches sized to about window_ms of the GPU's measured rate (100 ms, and the reference's
    default of 10 ms). Each frame's time includes building its grid."""
    from x2amd import grid as G
    from x2amd import scenes

    scene, config = scenes.icospheres()
    lane = make_lane()
    lane.prepare(scene)
    out = {"workload": f"icospheres {scene.triangle_count} tris, {config.width}x{config.height}, {config.spp} spp, "
                       "work grid of 64-pixel cells, GPU alone", "reps": reps}
    for window in windows:
        lane.render_grid(config, G.WorkGrid(config.width, config.height, config.spp, G.options(64, 4, 1)),
                         window_ms=window)
        wall, launches, chunks = 0.0, 0, 0
        for _ in range(reps):
            t0 = time.perf_counter()
            g = G.WorkGrid(config.width, config.height, config.spp, G.options(64, 4, 1))
            _, counts, fr = lane.render_grid(config, g, window_ms=window)
            wall += time.perf_counter() - t0
            assert fr.errors == 0 and (counts == config.spp).all()
            launches += fr.launches
            chunks = fr.gpu_chunks
        paths = config.width * config.height * config.spp * reps
        out[f"window_{int(window)}ms"] = {"ms_per_frame": wall / reps * 1e3, "mpaths_per_s": paths / wall / 1e6,
                                          "launches_per_frame": launches / reps, "chunks_per_frame": chunks}
    lane.close()
    return out


def paging_leg(reps):
    """Texture tiles on demand on the differentiable Cornell box's forward frame (512x512, 32 spp, its back wall a
    512x512 map): the texture resident, then its levels wider than 64 paged into a pool as large as them (336 tiles;
    the first frame parks every sample that meets a missing tile and runs it again), and paged with partial shading."""
    from x2amd import scenes

    scene, config, params = scenes.cornell_textured()
    out = {"workload": f"cornell textured {config.width}x{config.height}, {config.spp} spp, depth {config.max_depth}, "
                       "512x512 map, levels 512..128 in 336 tiles of 32x32 half floats", "reps": reps}
    for name, paging in [("resident", None), ("paged", (336, False)), ("paged_partial", (336, True))]:
        lane = make_lane()
        lane.prepare(scene)
        if paging:
            lane.page_textures(paging[0], partial=paging[1])
        t0 = time.perf_counter()
        _, counts, st = lane.render_frame(config)
        first = time.perf_counter() - t0
        assert st.errors == 0 and (counts == config.spp).all()
        g = lane.paging_stats() if paging else None  # the first frame's parked and provisional samples
        wall = 0.0
        for _ in range(reps):
            t0 = time.perf_counter()
            _, counts, st = lane.render_frame(config)
            wall += time.perf_counter() - t0
        entry = {"first_frame_ms": first * 1e3, "ms_per_frame": wall / reps * 1e3,
                 "mpaths_per_s": config.width * config.height * config.spp * reps / wall / 1e6}
        if paging:
            g2 = lane.paging_stats()
            entry.update(first_frame_parked=g.parked_total, first_frame_provisional=g.provisional, uploads=g2.uploads,
                         evictions=g2.evictions, tiles=g2.tiles)
        lane.close()
        out[name] = entry
    return out


def checkpoint_leg(stage_samples=16):
    """The headline frame as a checkpoint render (x2dev_render_staged): stages of 16 samples a pixel, the film saved
    whole after each ("X2RS", 75 MB at 1080p, to a temporary file), against the same frame at once; then the finished
    film's beauty as a half-float ZIP EXR."""
    import tempfile

    from x2amd import abi, scenes
    from x2amd import files as F

    scene, config = scenes.icospheres()
    lane = make_lane()
    lane.prepare(scene)
    lane.render_frame(config)
    t0 = time.perf_counter()
    lane.render_frame(config)
    plain = time.perf_counter() - t0
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "headline.exr.xpu2resume")
        t0 = time.perf_counter()
        sums, counts, st = lane.render_staged(config, path, stage_samples=stage_samples)
        staged = time.perf_counter() - t0
        assert st.errors == 0 and (counts == config.spp).all()
        t0 = time.perf_counter()
        F.write_exr(os.path.join(d, "headline.exr"), config.width, config.height, sums, counts, half=True,
                    compression=abi.X2_EXR_ZIP)
        exr = time.perf_counter() - t0
        exr_bytes = os.path.getsize(os.path.join(d, "headline.exr"))
    lane.close()
    return {"workload": f"icospheres {config.width}x{config.height}, {config.spp} spp in stages of {stage_samples}",
            "plain_ms": plain * 1e3, "staged_ms": staged * 1e3, "stages": st.stages, "saves": st.saves,
            "staged_over_plain": staged / plain, "exr_ms": exr * 1e3, "exr_bytes": exr_bytes}


def copy_ceiling(mib=1024):
    """GB/s of a plain device-to-device copy (read + write bytes over the copy's time), the chip's practical
    bandwidth beside the 8 TB/s peak."""
    import numpy as np
    from x2amd.hipmem import DeviceBuffer

    n = mib << 20
    a, b = DeviceBuffer(n), DeviceBuffer(n)
    a.zero()
    best = 1e30
    for _ in range(3):
        a.lib.hipDeviceSynchronize()
        t0 = time.perf_counter()
        if a.lib.hipMemcpy(b.ptr, a.ptr, n, 3) != 0:
            raise RuntimeError("hipMemcpy device to device failed")
        a.lib.hipDeviceSynchronize()
        best = min(best, time.perf_counter() - t0)
    a.free()
    b.free()
    return 2 * n / best / 1e9


def copy_calibration(mib):
    """A plain device-to-device copy of `mib` MiB (the runtime's blit kernel), run under rocprofv3 --pmc to scale
    FETCH_SIZE and WRITE_SIZE: the counters' reading for a copy of known size."""
    import numpy as np
    from x2amd.hipmem import DeviceBuffer

    n = mib << 20
    a, b = DeviceBuffer(n), DeviceBuffer(n)
    a.upload(np.ones(n // 4, np.float32))
    if a.lib.hipMemcpy(b.ptr, a.ptr, n, 3) != 0:  # device to device
        raise RuntimeError("hipMemcpy device to device failed")
    if a.lib.hipDeviceSynchronize() != 0:
        raise RuntimeError("hipDeviceSynchronize failed")
    a.free()
    b.free()
    print(json.dumps({"copy_mib": mib}))


TRAFFIC_FILE = os.path.join("profiles", "renderPoolKernel_icospheres_traffic.json")


def traffic_per_path():
    """HBM bytes per path of the render kernel at the headline workload, from the committed rocprofv3 --pmc
    measurement (FETCH_SIZE + WRITE_SIZE, calibrated by a copy), or None."""
    path = os.path.join(ROOT, TRAFFIC_FILE)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def progress(msg):
    """One line a finished leg, on stderr (the JSON line stays alone on stdout)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    global LAUNCH_CHUNKS
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--cpu-spp", type=int, default=0, help="CPU baseline spp (0: 16 with >= 12 host threads, else 4)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--adjoint-reps", type=int, default=5)
    ap.add_argument("--no-volume", action="store_true")
    ap.add_argument("--volume-reps", type=int, default=2)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-direct", action="store_true")
    ap.add_argument("--no-adaptive", action="store_true")
    ap.add_argument("--no-grid", action="store_true")
    ap.add_argument("--no-paging", action="store_true")
    ap.add_argument("--samples-per-chunk", type=int, default=4, help="sample indices a chunk holds (the headline)")
    ap.add_argument("--no-checkpoint", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, help="MiB: only run a device copy (for --pmc scaling)")
    ap.add_argument("--batch-reps", type=int, default=3)
    ap.add_argument("--launch-chunks", type=int, default=LAUNCH_CHUNKS, help="max_chunks_per_launch of the lanes")
    ap.add_argument("--wave-slots", type=int, default=0,
                    help="render the headline in the wavefront form with this many path slots (0: one kernel)")
    args = ap.parse_args()
    LAUNCH_CHUNKS = args.launch_chunks
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported (the film reduce "
                                   "across GPUs is a design note, not built)"}))
        sys.exit(2)

    from x2amd import scenes

    scene, config = scenes.icospheres(width=args.width, height=args.height, spp=args.spp)
    lane = make_lane(args.wave_slots)
    t = time.perf_counter()
    lane.prepare(scene)
    prepare_s = time.perf_counter() - t
    for _ in range(args.warmup):
        lane.render_frame(config, args.samples_per_chunk)
    totals = dict(launches=0, kernel_s=0.0, traces=0, shadows=0, nodes=0, tris=0)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        # the film lands in the lane's pinned host buffers (first launch -> film landed); no copy of it is timed
        _, counts, st = lane.render_frame(config, args.samples_per_chunk, copy=False)
        assert st.audit_failures == 0 and st.errors == 0
        totals["launches"] += st.launches
        totals["kernel_s"] += st.device_seconds
        totals["traces"] += st.rays_trace
        totals["shadows"] += st.rays_shadow
        totals["nodes"] += st.nodes_visited
        totals["tris"] += st.triangles_tested
    wall = time.perf_counter() - t0
    lane.close()
    progress(f"headline: {args.steps} frames in {wall:.2f} s")

    paths = config.width * config.height * config.spp * args.steps
    rays = totals["traces"] + totals["shadows"]
    from x2amd import abi as _abi
    bytes_total = _abi.NODE_BYTES * totals["nodes"] + _abi.LEAF_TRIANGLE_BYTES * totals["tris"] + \
        _abi.FILM_BYTES_PER_PATH * paths
    per_launch_s = totals["kernel_s"] / max(1, totals["launches"])
    measured = traffic_per_path()
    traffic = None  # HBM bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)
    if measured:
        traffic = measured["hbm_bytes_per_path"] * paths / max(1, totals["launches"])
    achieved = bytes_total / max(totals["kernel_s"], 1e-12) / 1e9
    # the same frame counted with the 256-B node and 48-B triangle records BASELINE.md fixed before the nodes went to
    # one cache line
    bytes_256_48 = 256 * totals["nodes"] + 48 * totals["tris"] + _abi.FILM_BYTES_PER_PATH * paths
    out = {
        "metric": "Mpaths/s + Mrays/s at 1080p 64spp (fraction of HBM roofline); adjoint step/forward",
        "value": paths / wall / 1e6,
        "unit": "Mpaths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"icospheres {scene.triangle_count} tris, {config.width}x{config.height}, "
                               f"{config.spp} spp, depth {config.max_depth}, 3 rect + env lights"},
        "mrays_per_s": rays / wall / 1e6,
        "rays_per_path": rays / paths,
        "nodes_per_ray": totals["nodes"] / max(1, rays),
        "triangles_per_ray": totals["tris"] / max(1, rays),
        "prepare_s": prepare_s,
        "launch": {"max_chunks_per_launch": LAUNCH_CHUNKS, "launches_per_frame": totals["launches"] / args.steps,
                   "samples_per_chunk": args.samples_per_chunk,
                   "form": f"wavefront, {args.wave_slots} path slots" if args.wave_slots else "one kernel"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": f"stored: {TRAFFIC_FILE} (rocprofv3 FETCH_SIZE and WRITE_SIZE of the kernel, "
                                       "each scaled by its reading for a copy of known size), times this run's paths "
                                       "a launch" if measured else None,
                     "kernel": "renderKernel" if args.wave_slots else "renderPoolKernel",
                     "avg_launch_ms": per_launch_s * 1e3,
                     "launches": totals["launches"], "algorithmic_bytes": bytes_total,
                     "bytes_per_node": _abi.NODE_BYTES, "bytes_per_tested_triangle": _abi.LEAF_TRIANGLE_BYTES,
                     "film_bytes_per_path": _abi.FILM_BYTES_PER_PATH,
                     "frac_at_256B_nodes_48B_triangles": bytes_256_48 / max(totals["kernel_s"], 1e-12) / 1e9 /
                     HBM_PEAK_GBS},
    }
    if not args.no_adjoint:
        out["adjoint"] = adjoint_leg(args.adjoint_reps)
        progress("adjoint leg done")
        out["adjoint_step_over_forward"] = out["adjoint"]["step_over_forward"]  # (forward + adjoint) / forward
    if not args.no_volume:
        out["volume"] = volume_leg(args.volume_reps, not args.no_cpu)
        progress("volume leg done")
    if not args.no_batch:
        out["batch"] = batch_leg(args.batch_reps)
        progress("batch leg done")
    if not args.no_direct:
        out["direct"] = direct_leg(2)
        progress("direct leg done")
    if not args.no_adaptive:
        out["adaptive"] = adaptive_leg(2)
        progress("adaptive leg done")
    if not args.no_grid:
        out["grid"] = grid_leg(2)
        progress("grid leg done")
    if not args.no_paging:
        out["paging"] = paging_leg(2)
        progress("paging leg done")
    if not args.no_checkpoint:
        out["checkpoint"] = checkpoint_leg()
        progress("checkpoint leg done")
    out["roofline"]["copy_ceiling_gbs"] = copy_ceiling()
    progress("copy ceiling done")
    if not args.no_cpu:
        cpu_spp = args.cpu_spp or (16 if len(os.sched_getaffinity(0)) >= 12 else 4)
        out["cpu_baseline"] = cpu_baseline(scene, config.width, config.height, cpu_spp)
        progress("cpu baseline done")
        out["gpu_over_cpu"] = out["value"] / out["cpu_baseline"]["value"]
        out["cpu_other_configs"] = cpu_other_configs()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
