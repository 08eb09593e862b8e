"""The on-disk formats on the CPU: the product's film save (x2_film_save / x2_film_load, the reference's "X2RS"
StagedFilm save) and its OpenEXR beauty writer (x2_write_exr) against the oracle (oracle/oracle_files.py).

Pins: a known-answer EXR written byte for byte from the published file layout; every value's half bits against
numpy's IEEE float16 rounding (normals, subnormals, overflow, ties, infinities, NaN); ZIP blocks through zlib and
OpenEXR's reordering and predictor; the data window against the film's rows from the bottom; the film save against
the reference's field order both ways, and its refusals (another render, not a save, cut short)."""
import os
import struct

import numpy as np
import pytest

import oracle_files as O
from x2amd import abi
from x2amd import files as F


def film(rng, w, h, empty=0.1):
    counts = rng.integers(1, 64, w * h).astype(np.uint32)
    counts[rng.random(w * h) < empty] = 0
    sums = rng.gamma(0.8, 2.0, (w * h, 4)) * counts[:, None]
    return sums, counts


def test_known_answer_exr(tmp_path):
    """A 2x1 float image, no compression: the file byte for byte as the OpenEXR layout fixes it."""
    sums = np.array([[1.0, 2.0, 3.0, 1.0], [0.5, 0.25, 0.125, 0.0]])
    counts = np.array([1, 1], np.uint32)
    path = str(tmp_path / "k.exr")
    F.write_exr(path, 2, 1, sums, counts, half=False, compression=abi.X2_EXR_NONE)
    got = open(path, "rb").read()

    def attr(name, typ, value):
        return name.encode() + b"\0" + typ.encode() + b"\0" + struct.pack("<i", len(value)) + value

    chlist = b"".join(n + b"\0" + struct.pack("<iB3xii", 2, 0, 1, 1) for n in (b"A", b"B", b"G", b"R")) + b"\0"
    header = struct.pack("<II", 20000630, 2)
    header += attr("channels", "chlist", chlist)
    header += attr("compression", "compression", b"\0")
    header += attr("dataWindow", "box2i", struct.pack("<4i", 0, 0, 1, 0))
    header += attr("displayWindow", "box2i", struct.pack("<4i", 0, 0, 1, 0))
    header += attr("lineOrder", "lineOrder", b"\0")
    header += attr("pixelAspectRatio", "float", struct.pack("<f", 1.0))
    header += attr("screenWindowCenter", "v2f", struct.pack("<2f", 0.0, 0.0))
    header += attr("screenWindowWidth", "float", struct.pack("<f", 1.0))
    header += b"\0"
    data = struct.pack("<8f", 1.0, 0.0, 3.0, 0.125, 2.0, 0.25, 1.0, 0.5)  # A A, B B, G G, R R
    chunk = struct.pack("<ii", 0, len(data)) + data
    want = header + struct.pack("<Q", len(header) + 8) + chunk
    assert got == want


@pytest.mark.parametrize("compression", [abi.X2_EXR_NONE, abi.X2_EXR_ZIPS, abi.X2_EXR_ZIP])
@pytest.mark.parametrize("half", [True, False])
def test_exr_round_trip(tmp_path, compression, half):
    rng = np.random.default_rng(compression * 2 + half)
    w, h = 53, 37
    sums, counts = film(rng, w, h)
    path = str(tmp_path / "b.exr")
    F.write_exr(path, w, h, sums, counts, half=half, compression=compression, weight=True)
    r = O.read_exr(path)
    assert r["compression"] == compression and r["data_window"] == (0, 0, w - 1, h - 1)
    assert r["display_window"] == (0, 0, w - 1, h - 1)
    assert [c[0] for c in r["channels"]] == ["A", "B", "G", "R", "weight"]
    mean = O.beauty(sums, counts).reshape(h, w, 4)[::-1]  # the file's rows from the top
    for k, name in enumerate("RGBA"):
        want = mean[:, :, k].astype(np.float16) if half else mean[:, :, k]
        assert r["pixels"][name].dtype == want.dtype
        assert (r["pixels"][name].view(np.uint16 if half else np.uint32) ==
                want.view(np.uint16 if half else np.uint32)).all(), name
    assert (r["pixels"]["weight"] == counts.reshape(h, w)[::-1].astype(np.float32)).all()
    assert not os.path.exists(path + ".part")


def test_exr_region_is_the_data_window(tmp_path):
    rng = np.random.default_rng(4)
    w, h = 40, 30
    sums, counts = film(rng, w, h, 0.0)
    region = (5, 3, 31, 22)  # the film's pixels [5, 31) x [3, 22), rows from the bottom
    path = str(tmp_path / "r.exr")
    F.write_exr(path, w, h, sums, counts, region=region, half=False, compression=abi.X2_EXR_ZIP)
    r = O.read_exr(path)
    assert r["data_window"] == (5, h - 22, 30, h - 1 - 3)
    mean = O.beauty(sums, counts).reshape(h, w, 4)
    want = mean[3:22, 5:31][::-1]
    for k, name in enumerate("RGBA"):
        assert (r["pixels"][name] == want[:, :, k]).all()


def test_half_rounding_matches_ieee(tmp_path):
    """Every value category through the writer (a film of one sample a pixel, so each mean is the value itself):
    the half bits equal numpy's float16 rounding."""
    rng = np.random.default_rng(9)
    special = np.array([0.0, -0.0, 1.0, -1.0, 65504.0, 65519.99, 65520.0, 1e6, -1e6, np.inf, -np.inf, np.nan,
                        6.1035156e-05, 6.0975552e-05, 5.9604645e-08, 2.9802322e-08, 2.9802326e-08, 1e-9,
                        1.0009765625, 1.00048828125, 1.000732421875, 2049.0, 2051.0], np.float32)
    bits = rng.integers(0, 2 ** 32, 20000, dtype=np.uint64).astype(np.uint32).view(np.float32)
    bits = bits[np.isfinite(bits)]
    mags = (rng.standard_normal(20000) * np.exp(rng.uniform(-20, 12, 20000))).astype(np.float32)
    values = np.concatenate([special, bits, mags])
    w = 128
    h = (len(values) + w - 1) // w
    v = np.zeros(w * h, np.float32)
    v[:len(values)] = values
    sums = np.zeros((w * h, 4))
    sums[:, 0] = v.astype(np.float64)
    counts = np.ones(w * h, np.uint32)
    path = str(tmp_path / "h.exr")
    F.write_exr(path, w, h, sums, counts, half=True, compression=abi.X2_EXR_NONE)
    got = O.read_exr(path)["pixels"]["R"][::-1].ravel()
    with np.errstate(over="ignore"):
        want = v.astype(np.float16)
    nan = np.isnan(want)
    assert (got[~nan].view(np.uint16) == want[~nan].view(np.uint16)).all()
    assert np.isnan(got[nan]).all()


def test_zip_blocks_stored_when_they_do_not_shrink(tmp_path):
    """A one-pixel ZIPS image of noise: the block does not shrink, so it is stored as it is (the reader tells by its
    size); a smooth image's blocks are deflated."""
    rng = np.random.default_rng(1)
    sums = rng.random((1, 4)) * 1e3
    path = str(tmp_path / "n.exr")
    F.write_exr(path, 1, 1, sums, np.ones(1, np.uint32), half=False, compression=abi.X2_EXR_ZIPS)
    r = O.read_exr(path)
    b = open(path, "rb").read()
    (n,) = struct.unpack_from("<i", b, r["offsets"][0] + 4)
    assert n == 16 and np.allclose(r["pixels"]["R"], O.beauty(sums, np.ones(1, np.uint32))[0, 0])
    smooth = np.tile(np.array([[1.0, 1.0, 1.0, 1.0]]), (64 * 16, 1))
    F.write_exr(path, 64, 16, smooth, np.ones(64 * 16, np.uint32), half=True, compression=abi.X2_EXR_ZIP)
    r = O.read_exr(path)
    b = open(path, "rb").read()
    (n,) = struct.unpack_from("<i", b, r["offsets"][0] + 4)
    assert n < 64 * 16 * 8 // 4 and (r["pixels"]["G"] == 1.0).all()


def test_film_save_both_ways(tmp_path):
    rng = np.random.default_rng(2)
    w, h, spp = 31, 17, 24
    sums, counts = film(rng, w, h)
    odd = sums * 0.5
    a = str(tmp_path / "a.xpu2resume")
    F.save_film(a, w, h, spp, 12, sums, counts, odd)
    r = O.read_film(a)
    assert (r["magic"], r["version"], r["width"], r["height"], r["spp"], r["done"]) == (O.FILM_MAGIC, 1, w, h, spp, 12)
    assert r["channels"] == 0 and r["closest"] == 0 and r["categories"] == 0 and r["end"] == r["size"]
    assert (r["sums"] == sums).all() and (r["counts"] == counts).all() and (r["odd"] == odd).all()
    b = str(tmp_path / "b.xpu2resume")
    O.write_film(b, w, h, spp, 7, sums, counts)
    s2, c2, o2, done = F.load_film(b, w, h, spp)
    assert done == 7 and o2 is None and (s2 == sums).all() and (c2 == counts).all()
    s3, c3, o3, done = F.load_film(a, w, h, spp)
    assert done == 12 and (o3 == odd).all() and (s3 == sums).all()
    assert not os.path.exists(a + ".part")


def test_film_load_refusals(tmp_path):
    rng = np.random.default_rng(3)
    w, h, spp = 16, 8, 4
    sums, counts = film(rng, w, h)
    path = str(tmp_path / "s.xpu2resume")
    with pytest.raises(RuntimeError, match="no saved render"):
        F.load_film(path, w, h, spp)
    F.save_film(path, w, h, spp, 4, sums, counts)
    with pytest.raises(RuntimeError, match="another render"):
        F.load_film(path, w, h, spp + 1)
    with pytest.raises(RuntimeError, match="another render"):
        F.load_film(path, w + 8, h, spp)
    data = open(path, "rb").read()
    open(path, "wb").write(data[:len(data) - 9])
    with pytest.raises(RuntimeError, match="cut short"):
        F.load_film(path, w, h, spp)
    open(path, "wb").write(b"X2RG" + data[4:])
    with pytest.raises(RuntimeError, match="not an XPU2 save"):
        F.load_film(path, w, h, spp)
    with pytest.raises(RuntimeError, match="more samples done"):
        F.save_film(path, w, h, spp, spp + 1, sums, counts)
