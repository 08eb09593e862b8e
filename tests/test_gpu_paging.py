"""On-demand texture tiles on the device (x2dev_page_textures; the reference's fine-detail cache and parked samples,
async_gpu.h:2911-3200, async_gpu.cu:350-357): a frame whose texture's finer mip levels live in a small pool of 32x32
half-float tiles — samples parked on missing tiles and run again once their tiles are sent up, the pool evicting as
it fills — is the oracle's frame of the same half-float levels (counts exact, sums to the FP64 order), and so with
partial shading (provisional samples reading missing texels' resident ancestors, then run again in full), through the
frame loop, the work grid's broker and the staged checkpoint render. The levels a frame reads are those its
footprints pick, so a large texture seen small pages in a sliver of its tiles. A first launch's provisional samples
are the oracle's samples with every finer fetch read from its ancestor on the first resident level. Calls that read
resident textures refuse paged ones."""
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


TILES = 80  # a 256^2 texture with coarse_side 64: level 0's 8 x 8 tiles and level 1's 4 x 4; levels 2 on resident
FIRST_RESIDENT = 2


def textured(size=128, spp=8, tex=256):
    scene, config, params = scenes.cornell_textured(width=size, height=size, spp=spp, texture_size=tex)
    return scene, config


def oracle_paged(scene, coarse_from=None):
    """The oracle's scene with the texture's levels held as the tiles hold them (half floats), and with coarse_from
    every finer fetch read from its ancestor there (a first launch with partial shading and no tile resident)."""
    o = oracle.OracleScene(scene)
    o.set_texture_half(0, True)
    if coarse_from is not None:
        o.set_texture_coarse(0, coarse_from)
    return o


@pytest.fixture(scope="module")
def frames():
    """The paged scene, its config and the oracle's film of the same half-float levels resident."""
    scene, config = textured()
    ws, wc, k = oracle_paged(scene).render(config)
    assert k[2] == 0 and (wc == config.spp).all()
    return scene, config, ws, wc


@pytest.mark.parametrize("pool,partial", [(96, False), (96, True), (40, True)])
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
    assert g.tiles == TILES and g.parked_total > 0 and g.uploads >= g.resident
    assert (pc == wc).all() and close(ps, ws)
    if pool >= TILES:
        assert g.evictions == 0
    # a provisional sample leaves the provisional film when it runs again in full: a finished frame's is empty
    assert int(prov[1].sum()) == 0 and np.abs(prov[0]).max() <= 1e-9
    if partial:
        assert g.provisional > 0
    else:
        assert g.provisional == 0


def test_footprints_read_coarser_levels():
    """Mip levels chosen from the footprint: a 1024^2 texture seen at 64 x 64 pixels reads levels near its 64^2 one,
    so a pool of 128 tiles, a tenth of the 1,365 its paged levels hold, keeps the frame's working set without an
    eviction; the frame is the oracle's."""
    scene, config, params = scenes.cornell_textured(width=64, height=64, spp=4, texture_size=1024, max_depth=2)
    ws, wc, k = oracle_paged(scene).render(config)
    lane = GpuLane(0)
    try:
        lane.prepare(scene)
        lane.page_textures(128, coarse_side=16)
        ps, pc, st = lane.render_frame(config)
        g = lane.paging_stats()
    finally:
        lane.close()
    print(f"1024^2 at 64^2: tiles {g.tiles}, misses {g.misses}, uploads {g.uploads}, evictions {g.evictions}")
    assert g.tiles == 1024 + 256 + 64 + 16 + 4 + 1 and g.evictions == 0 and g.overflow == 0
    assert st.errors == 0 and (pc == wc).all() and close(ps, ws)


def test_provisional_film_holds_the_parked_samples(frames):
    """Partial shading through the step calls (async_gpu.cu:282-295): after each launch every sample issued is in the
    film or, provisional, in the provisional film, and the provisional film counts exactly the samples still parked;
    launches of parked samples alone empty it, and the film is then the oracle's resident frame. The first launch
    finds no tile resident, so each fetch of a paged level reads its ancestor on the first resident level: the film
    plus the provisional film is then the oracle's frame read so."""
    scene, config, ws, wc = frames
    cs, cc, ck = oracle_paged(scene, FIRST_RESIDENT).render(config)
    assert ck[2] == 0
    tiles = ((config.width + 7) // 8) * ((config.height + 7) // 8)
    every = [abi.Chunk(t, f, 4, 0) for f in range(0, config.spp, 4) for t in range(tiles)]
    lane = GpuLane(0)
    try:
        lane.prepare(scene)
        lane.page_textures(96, partial=True)
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
            if launch == 0:  # every fetch of a paged level in the first launch read its resident ancestor
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


def test_host_refines_round_trip(frames):
    """The reference's flow for provisional samples (async_gpu.cu:274-295, async_gpu.h:1434-1525, 1713): with host
    refines a provisional sample leaves the device as a Refine record (its pixel, index and value, the tile it missed)
    instead of parking; the host takes them and sends them back through a launch, which renders them again in full
    (refines_run) and takes their values out of the provisional film. The first launch's film plus provisional film is
    the oracle's frame read from the resident ancestors; after the round trip the film is the oracle's frame."""
    scene, config, ws, wc = frames
    cs, cc, ck = oracle_paged(scene, FIRST_RESIDENT).render(config)
    tiles = ((config.width + 7) // 8) * ((config.height + 7) // 8)
    every = [abi.Chunk(t, f, 4, 0) for f in range(0, config.spp, 4) for t in range(tiles)]
    lane = GpuLane(0)
    try:
        lane.prepare(scene)
        lane.page_textures(96, partial=True, host_refines=True)
        lane.begin(config)
        assert lane.launch((abi.Chunk * len(every))(*every))
        done = drain(lane)
        assert len(done) == 1 and done[0].refines_run == 0 and done[0].parked_run == 0 and done[0].errors == 0
        refines = lane.take_refines()
        lane.download_film()
        while not lane.film_ready():
            pass
        fs, fc = lane.film()
        ps, pc = lane.provisional_film()
        n = len(refines)
        assert n > 0 and int(pc.sum()) == n and lane.parked_outstanding() == 0
        assert (fc + pc == config.spp).all() and close(fs + ps, cs)
        mine = np.zeros_like(ps)
        for r in refines:  # each record's value is its sample's share of the provisional film
            assert 0 < r.tile <= TILES
            mine[r.pixel] += np.array(r.rgba[:], np.float64)
        assert np.abs(mine - ps).max() <= 1e-9 * max(1.0, float(np.abs(ps).max()))
        rounds = 0
        while len(refines):  # the tiles they missed go up with the launch that renders them again
            assert lane.launch((abi.Chunk * 0)(), refines=refines)
            done = drain(lane)
            assert sum(c.refines_run for c in done) == len(refines) and all(c.errors == 0 for c in done)
            refines = lane.take_refines()
            rounds += 1
            assert rounds < 8
        lane.download_film()
        while not lane.film_ready():
            pass
        fs, fc = lane.film()
        ps, pc = lane.provisional_film()
        g = lane.paging_stats()
    finally:
        lane.close()
    print(f"host refines: {n} provisional samples, {rounds} round(s) back, provisional renders {g.provisional}")
    assert int(pc.sum()) == 0 and np.abs(ps).max() <= 1e-9 and g.overflow == 0
    assert (fc == wc).all() and close(fs, ws)


def test_host_refines_through_the_frame_loop_and_the_grid(frames):
    scene, config, ws, wc = frames
    lane = GpuLane(0)
    try:
        lane.prepare(scene)
        lane.page_textures(96, partial=True, host_refines=True)
        fs, fc, st = lane.render_frame(config)
        stats = lane.paging_stats()  # the first frame's: its tiles were missing
        g = G.WorkGrid(config.width, config.height, config.spp, G.options(64, 4, 1))
        gs, gc, fr = lane.render_grid(config, g, window_ms=5.0)
        prov = lane.provisional_film()
    finally:
        lane.close()
    assert st.errors == 0 and (fc == wc).all() and close(fs, ws)
    assert fr.errors == 0 and (gc == wc).all() and close(gs, ws)
    assert int(prov[1].sum()) == 0 and stats.provisional > 0 and stats.overflow == 0


def test_frames_with_other_tiles_evict(tmp_path):
    """Camera hits only (depth 0): the left half of the image reads some of the texture's tiles, the right half others.
    In a pool of 48 of the 80 tiles the second frame evicts the first's; each equals the oracle's frame."""
    scene, config = textured(128, 4, 256)
    config.max_depth = 0
    halves = [[0, 0, 64, 128], [64, 0, 128, 128]]
    want = []
    o = oracle_paged(scene)
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
        lane.page_textures(96, partial=True)
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


def half_exact_texture(size=256, block=16, seed=5):
    """A texture constant over block x block squares, each 0.25, 0.5 or 0.75: every level of its pyramid (means of
    such values, multiples of 2^-10 in [0.25, 0.75]) is exact in half floats, so its resident pyramid is the paged one."""
    rng = np.random.default_rng(seed)
    blocks = rng.choice([0.25, 0.5, 0.75], size=(size // block, size // block, 3))
    return np.repeat(np.repeat(blocks, block, axis=0), block, axis=1).astype(np.float32)


def test_paged_adaptive_frame_equals_resident():
    """An adaptive frame (host trees between passes) with paged textures: each pass's parked samples run again before
    its errors are taken, so every pass samples the resident frame's pixels and the films agree (a texture whose
    levels are exact in half floats, so the two read the same values)."""
    tex = half_exact_texture()
    for a, b in zip(oracle.pyramid(tex), oracle.pyramid(tex, half=True)):
        assert np.array_equal(a, b)
    scene, config, _ = scenes.cornell_textured(width=96, height=96, spp=32, texture=tex)
    out = []
    for paged in (False, True):
        lane = GpuLane(0)
        try:
            lane.prepare(scene)
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
    scene, config = textured(32, 2, 64)
    lane = GpuLane(0)
    try:
        lane.prepare(scene)
        with pytest.raises(RuntimeError, match="coarse_side"):
            lane.page_textures(4)  # every level of a 64^2 texture is within the default 64: nothing to page
        lane.page_textures(4, coarse_side=16)
        with pytest.raises(RuntimeError, match="paged"):
            lane.set_texture(0, np.zeros((64, 64, 3), np.float32))
        with pytest.raises(RuntimeError, match="paged"):
            lane.page_textures(4, coarse_side=16)
        config.direct_only, config.surface_limit = 1, 1
        with pytest.raises(RuntimeError, match="paged textures"):
            lane.render_frame(config)
    finally:
        lane.close()
    wave = GpuLane(0, wave_slots=4096)
    try:
        with pytest.raises(RuntimeError, match="wavefront"):
            wave.prepare(scene)  # the wavefront form renders untextured scenes
    finally:
        wave.close()
