"""A frame taken up from a saved film (x2dev_seed_film, GpuLane::seedFilm): the first half of a frame's samples on
one lane, its film saved, the second half on a fresh lane seeded with it — the oracle's whole frame and the same film
as the whole frame at once (counts exact, sums to the FP64 order); an adaptive frame's errors go on from the seeded
film."""
import numpy as np
import pytest

import oracle

from helpers import drain, mean_relative_difference
from x2amd import abi, scenes
from x2amd import lane as L
from x2amd.lane import GpuLane

pytestmark = pytest.mark.gpu


def chunks_for(config, first, count, spc=4):
    tiles = ((config.width + 7) // 8) * ((config.height + 7) // 8)
    out = []
    for f in range(first, first + count, spc):
        out += [abi.Chunk(t, f, min(spc, first + count - f), 0) for t in range(tiles)]
    return (abi.Chunk * len(out))(*out)


def run(lane, chunks, cap=16384):
    for i in range(0, len(chunks), cap):
        part = chunks[i:i + cap]
        arr = (abi.Chunk * len(part))(*part)
        while not lane.ready():
            lane.poll()
        assert lane.launch(arr, samples=0)
    drain(lane)


def film(lane):
    lane.download_film()
    while not lane.film_ready():
        pass
    return lane.film()


def test_resumed_frame_equals_whole_frame():
    scene, config = scenes.cornell(width=96, height=72, spp=16)
    whole = GpuLane(0)
    try:
        whole.prepare(scene)
        ws, wc, _ = whole.render_frame(config)
    finally:
        whole.close()
    first = GpuLane(0)
    try:
        first.prepare(scene)
        first.begin(config)
        run(first, chunks_for(config, 0, 8))
        hs, hc = film(first)
    finally:
        first.close()
    assert (hc == 8).all()
    second = GpuLane(0)
    try:
        second.prepare(scene)
        second.begin(config)
        second.seed_film(hs, hc)
        run(second, chunks_for(config, 8, 8))
        rs, rc = film(second)
    finally:
        second.close()
    assert (rc == wc).all() and (rc == 16).all()
    assert mean_relative_difference(rs, rc, ws, wc) <= 1e-12
    os_, oc, ok = oracle.OracleScene(scene).render(config)
    assert ok[2] == 0 and (rc == oc).all() and mean_relative_difference(rs, rc, os_, oc) <= 1e-12
    # the first lane's half film is the oracle's samples 0..7
    hs_o, hc_o, _ = oracle.OracleScene(scene).render(config, chunks=chunks_for(config, 0, 8))
    assert (hc == hc_o).all() and mean_relative_difference(hs, hc, hs_o, hc_o) <= 1e-12


def test_seeded_adaptive_film_errors():
    scene, config = scenes.cornell(width=40, height=32, spp=8)
    lane = GpuLane(0)
    try:
        lane.prepare(scene)
        s, o, c, _, _ = lane.render_adaptive(config, 8, 8, 1e-3)  # every pass uniform: 8 samples a pixel
        lane.begin_adaptive(config)
        lane.seed_film(s, c, o)
        errors = lane.errors()
    finally:
        lane.close()
    assert (c == 8).all()
    assert (errors.view(np.uint32) == L.pixel_errors(s, o, c).view(np.uint32)).all()
    assert (errors >= 0).all()
