"""On-demand texture tiles on the device (x2dev_page_textures; the reference's fine-detail cache and parked samples,
async_gpu.h:2911-3200, async_gpu.cu:350-357): a frame whose texture lives in a small pool of 32x32 half-float tiles —
samples parked on missing tiles and run again once their tiles are sent up, the pool evicting as it fills — is the
oracle's frame of the same half-float texels (counts exact, sums to the FP64 order), and so with partial shading
(provisional samples from the coarse level, then run again in full), through the frame loop, the work grid's broker
and the staged checkpoint render. A launch's provisional samples are the oracle's samples with the coarse level as
the texture. Calls that read resident textures refuse paged ones."""
import numpy as np
import pytest
from helpers import drain

import oracle

from x2amd import abi
from x2amd import grid as G
from x2amd import scenes
from x2amd.lane import GpuLane

pytestmark = pytest.mark.gpu


def close(a, b):
    return np.abs(a - b).max() <= 1e-12 * max(1.0, float(np.abs(b).max()))


def textured(size=128, spp=8, tex=256):
    scene, config, params = scenes.cornell_textured(width=size, height=size, spp=spp, texture_size=tex)
    half = params["texture"].astype(np.float16).astype(np.float32)  # the texels the tiles hold
    resident, _, _ = scenes.cornell_textured(width=size, height=size, spp=spp, texture=half)
    return scene, resident, config


def coarse_level(texture, side=64):
    """The always-resident coarse level as x2dev_page_textures makes it: level 0's half floats averaged over boxes
    (float64 sums, exact for these few halves), at most `side` texels a side."""
    half = np.asarray(texture, np.float32).astype(np.float16).astype(np.float64)
    h, w = half.shape[:2]
    cw, ch = w, h
    while cw > side:
        cw = (cw + 1) // 2
    while ch > side:
        ch = (ch + 1) // 2
    out = np.zeros((ch, cw, 3), np.float32)
    for j in range(ch):
        y0, y1 = j * h // ch, max(j * h // ch + 1, (j + 1) * h // ch)
        for i in range(cw):
            x0, x1 = i * w // cw, max(i * w // cw + 1, (i + 1) * w // cw)
            out[j, i] = (half[y0:y1, x0:x1].sum(axis=(0, 1)) / ((x1 - x0) * (y1 - y0))).astype(np.float32)
    return out


@pytest.fixture(scope="module")
def frames():
    """The paged scene, its config and the oracle's film of the same half-float texels resident."""
    scene, resident, config = textured()
    ws, wc, k = oracle.OracleScene(resident).render(config)
    assert k[2] == 0 and (wc == config.spp).all()
    return scene, config, ws, wc


@pytest.mark.parametrize("pool,partial", [(64, False), (64, True), (80, True)])
def test_paged_frame_equals_resident(frames, pool, partial):
    scene, config, ws, wc = frames
    lane = GpuLane(0)
    try:
        lane.prepare(scene)
        lane.page_textures(pool, partial=partial)
        ps, pc, st = lane.render_frame(config)
        g = lane.paging_stats()
        prov = lane.provisional_film()
    finally:
        lane.close()
    print(f"pool {pool} partial {partial}: parked {g.parked_total}, provisional {g.provisional}, misses {g.misses}, "
          f"uploads {g.uploads}, evictions {g.evictions}, resident {g.resident}/{g.tiles}, launches {st.launches}")
    assert st.errors == 0 and st.audit_failures == 0 and g.overflow == 0 and g.parked_outstanding == 0
    assert g.tiles == 64 and g.parked_total > 0 and g.uploads >= g.resident
    assert (pc == wc).all() and close(ps, ws)
    assert g.evictions == 0  # the pool holds the frame's working set (the whole texture: level 0 is read everywhere)
    # a provisional sample leaves the provisional film when it runs again in full: a finished frame's is empty
    assert int(prov[1].sum()) == 0 and np.abs(prov[0]).max() <= 1e-9
    if partial:
        assert g.provisional > 0
    else:
        assert g.provisional == 0


def test_provisional_film_holds_the_parked_samples(frames):
    """Partial shading through the step calls (async_gpu.cu:282-295): after each launch every sample issued is in the
    film or, provisional, in the provisional film, and the provisional film counts exactly the samples still parked;
    launches of parked samples alone empty it, and the film is then the oracle's resident frame. The first launch
    finds no tile resident, so each of its samples that reads the texture reads the coarse level throughout: the
    film plus the provisional film is then the oracle's frame with the coarse level as the texture."""
    scene, config, ws, wc = frames
    _, _, params = scenes.cornell_textured(width=config.width, height=config.height, spp=config.spp, texture_size=256)
    coarse_scene, _, _ = scenes.cornell_textured(width=config.width, height=config.height, spp=config.spp,
                                                 texture=coarse_level(params["texture"]))
    cs, cc, ck = oracle.OracleScene(coarse_scene).render(config)
    assert ck[2] == 0
    tiles = ((config.width + 7) // 8) * ((config.height + 7) // 8)
    every = [abi.Chunk(t, f, 4, 0) for f in range(0, config.spp, 4) for t in range(tiles)]
    lane = GpuLane(0)
    try:
        lane.prepare(scene)
        lane.page_textures(64, partial=True)
        lane.begin(config)
        trace = []
        for launch in range(64):
            chunks = (abi.Chunk * len(every))(*every) if launch == 0 else (abi.Chunk * 0)()
            if not lane.launch(chunks):
                break
            drain(lane)
            lane.download_film()
            while not lane.film_ready():
                pass
            fs, fc = lane.film()
            ps, pc = lane.provisional_film()
            parked = lane.parked_outstanding()
            trace.append((int(pc.sum()), parked))
            assert (fc + pc == config.spp).all()
            assert int(pc.sum()) == parked
            if launch == 0:  # every textured sample of the first launch read the coarse level
                assert int(pc.sum()) > 0 and close(fs + ps, cs)
            assert np.isfinite(ps).all() and np.abs(ps[pc == 0]).max(initial=0.0) <= 1e-9
            if parked == 0:
                break
        g = lane.paging_stats()
    finally:
        lane.close()
    print(f"provisional samples outstanding, parked after each launch: {trace}; provisional renders {g.provisional}")
    assert trace[0][0] > 0 and trace[-1] == (0, 0) and g.overflow == 0
    assert (fc == wc).all() and close(fs, ws)


def test_frames_with_other_tiles_evict(tmp_path):
    """Camera hits only (depth 0): the left half of the image reads some of the texture's tiles, the right half others.
    In a pool of 48 of the 64 tiles the second frame evicts the first's; each equals its resident frame."""
    scene, resident, config = textured(128, 4, 256)
    config.max_depth = 0
    halves = [[0, 0, 64, 128], [64, 0, 128, 128]]
    want = []
    o = oracle.OracleScene(resident)
    for vp in halves:
        config.viewport[:] = vp
        want.append(o.render(config)[:2])
    lane = GpuLane(0)
    try:
        lane.prepare(scene)
        lane.page_textures(48)
        got = []
        for vp in halves:
            config.viewport[:] = vp
            got.append(lane.render_frame(config)[:2])
        g = lane.paging_stats()
    finally:
        lane.close()
    print(f"two halves: misses {g.misses}, uploads {g.uploads}, evictions {g.evictions}, resident {g.resident}")
    assert g.evictions > 0 and g.overflow == 0 and g.parked_outstanding == 0
    for (gs, gc), (ws, wc) in zip(got, want):
        assert (gc == wc).all() and close(gs, ws)


def test_paged_frames_through_the_grid_and_stages(frames, tmp_path):
    scene, config, ws, wc = frames
    lane = GpuLane(0)
    try:
        lane.prepare(scene)
        lane.page_textures(64, partial=True)
        g = G.WorkGrid(config.width, config.height, config.spp, G.options(64, 4, 1))
        gs, gc, fr = lane.render_grid(config, g, window_ms=5.0)
        ss, sc, sf = lane.render_staged(config, str(tmp_path / "paged.xpu2resume"), stage_samples=3)
        # a second frame on the same lane: its tiles are resident from the first
        rs, rc, _ = lane.render_frame(config)
        st = lane.paging_stats()
    finally:
        lane.close()
    assert fr.errors == 0 and (gc == wc).all() and close(gs, ws)
    assert sf.errors == 0 and (sc == wc).all() and close(ss, ws)
    assert (rc == wc).all() and close(rs, ws) and st.parked_outstanding == 0


def test_paged_adaptive_frame_equals_resident():
    """An adaptive frame (host trees between passes) with paged textures: each pass's parked samples run again before
    its errors are taken, so every pass samples the resident frame's pixels and the films agree."""
    scene, resident, config = textured(96, 32, 256)
    out = []
    for sc, paged in ((resident, False), (scene, True)):
        lane = GpuLane(0)
        try:
            lane.prepare(sc)
            if paged:
                lane.page_textures(64)
            sums, odd, counts, st, a = lane.render_adaptive(config, 4, 32, 1e-3)
            out.append((sums, odd, counts, a.passes))
        finally:
            lane.close()
    (rs, ro, rc, rp), (ps, po, pc, pp) = out
    assert rp == pp and (pc == rc).all() and rc.min() < rc.max()
    assert close(ps, rs) and close(po, ro)


def test_a_pool_too_small_for_a_path_fails_the_frame(frames):
    """A pool smaller than what a path reads at once (two tiles; a bilinear footprint alone can need four): its samples
    park again and again, and the frame fails with the reason rather than running on (the reference's frame fails so
    too, async_render.h:816-818)."""
    scene, config, _, _ = frames
    lane = GpuLane(0)
    try:
        lane.prepare(scene)
        lane.page_textures(2)
        with pytest.raises(RuntimeError, match="cannot hold"):
            lane.render_frame(config)
    finally:
        lane.close()


def test_calls_that_read_resident_textures_refuse_paged_ones():
    scene, _, config = textured(32, 2, 64)
    lane = GpuLane(0)
    try:
        lane.prepare(scene)
        lane.page_textures(4)
        with pytest.raises(RuntimeError, match="paged"):
            lane.set_texture(0, np.zeros((64, 64, 3), np.float32))
        with pytest.raises(RuntimeError, match="paged"):
            lane.page_textures(4)
        config.direct_only, config.surface_limit = 1, 1
        with pytest.raises(RuntimeError, match="paged textures"):
            lane.render_frame(config)
    finally:
        lane.close()
    wave = GpuLane(0, wave_slots=4096)
    try:
        wave.prepare(scene)
        with pytest.raises(RuntimeError, match="wavefront"):
            wave.page_textures(4)
    finally:
        wave.close()
