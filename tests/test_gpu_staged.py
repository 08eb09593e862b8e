"""A checkpoint render on the device (x2dev_render_staged, the reference's stagedFrame): the frame in stages with its
film saved after each equals the whole frame (counts exact, sums to the FP64 atomics' order); a render stopped after
two stages and taken up on a fresh lane from its save equals it too; a save of another render is not taken up; the
saved film (read by the oracle's restatement of the reference's save) is the device's film; and the finished frame's
EXR holds the film's means."""
import numpy as np
import pytest

import oracle
import oracle_files as O
from x2amd import abi, scenes
from x2amd import files as F
from x2amd.lane import GpuLane

pytestmark = pytest.mark.gpu


def close(a, b):
    return np.abs(a - b).max() <= 1e-12 * max(1.0, float(np.abs(b).max()))


def test_staged_frame_equals_whole_frame(tmp_path):
    scene, config = scenes.cornell(width=96, height=64, spp=20)
    path = str(tmp_path / "frame.exr.xpu2resume")
    lane = GpuLane(0)
    try:
        lane.prepare(scene)
        ws, wc, _ = lane.render_frame(config)
        ss, sc, st = lane.render_staged(config, path, stage_samples=6)
    finally:
        lane.close()
    print(f"stages {st.stages}, saves {st.saves}, launches {st.launches}, {st.seconds * 1e3:.1f} ms")
    assert st.errors == 0 and st.stages == 4 and st.saves == 4 and st.done == 20 and not st.resumed
    assert (sc == wc).all() and close(ss, ws)
    os_, oc, ok = oracle.OracleScene(scene).render(config)
    assert ok[2] == 0 and (sc == oc).all() and close(ss, os_)
    r = O.read_film(path)
    assert r["done"] == 20 and r["spp"] == 20 and (r["counts"] == sc).all() and (r["sums"] == ss).all()
    exr = str(tmp_path / "frame.exr")
    F.write_exr(exr, config.width, config.height, ss, sc, half=True, compression=abi.X2_EXR_ZIP)
    img = O.read_exr(exr)
    mean = O.beauty(ss, sc).reshape(config.height, config.width, 4)[::-1]
    assert (img["pixels"]["G"] == mean[:, :, 1].astype(np.float16)).all()


def test_stopped_render_taken_up_from_its_save(tmp_path):
    scene, config = scenes.cornell(width=80, height=72, spp=16)
    path = str(tmp_path / "take.xpu2resume")
    lane = GpuLane(0)
    try:
        lane.prepare(scene)
        ws, wc, _ = lane.render_frame(config)
        ps, pc, p = lane.render_staged(config, path, stage_samples=3, max_stages=2)
    finally:
        lane.close()
    assert p.stages == 2 and p.done == 6 and (pc == 6).all()
    second = GpuLane(0)
    try:
        second.prepare(scene)
        rs, rc, r = second.render_staged(config, path, stage_samples=3, resume=True)
    finally:
        second.close()
    print(f"taken up from {r.resumed_from}: {r.stages} more stages")
    assert r.resumed and r.resumed_from == 6 and r.stages == 4 and r.done == 16
    assert (rc == wc).all() and close(rs, ws)
    os_, oc, _ = oracle.OracleScene(scene).render(config)
    assert (rc == oc).all() and close(rs, os_)


def test_save_of_another_render_begins_anew(tmp_path):
    scene, config = scenes.cornell(width=48, height=40, spp=8)
    path = str(tmp_path / "other.xpu2resume")
    n = config.width * config.height
    F.save_film(path, config.width, config.height, 12, 5, np.ones((n, 4)), np.full(n, 5, np.uint32))
    lane = GpuLane(0)
    try:
        lane.prepare(scene)
        ws, wc, _ = lane.render_frame(config)
        rs, rc, r = lane.render_staged(config, path, resume=True)
    finally:
        lane.close()
    assert not r.resumed and r.done == 8 and (rc == wc).all() and close(rs, ws)
    assert O.read_film(path)["spp"] == 8  # overwritten by this render's saves
