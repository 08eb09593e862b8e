"""The oracle for the on-disk formats — test infrastructure only (imported by tests/ and nothing shipped).

1. A checkpoint render's film save, "X2RS" version 1, restated from the reference's StagedFilm::save / load
   (xpu2/interactive/xpu2_session.cc:1416-1513): eight u32 (magic 0x53523258, version 1, width, height, spp, done,
   AOV channels, closest filters), then the arrays sums (f64), counts (u32), odd (f64), AOVs (f64), depths (f32),
   each a u64 element count and its elements, then a u32 count of Cryptomatte categories (and their fragments).
2. A scanline OpenEXR reader written from the published file layout (openexr.com, "The OpenEXR File Layout"): the
   magic number 20000630, the version word (2: single part, scanlines), the header's attributes (name, type, size,
   value) up to an empty name, the line offset table (one u64 a chunk), and the chunks (y, data size, data), each
   scanline's channels in the header's order, each channel's values for the data window's width; NONE, ZIPS (one
   line a chunk) and ZIP (16 lines) compression, the latter two deflated by zlib after OpenEXR's byte reordering
   (even bytes, then odd bytes) and delta predictor (each byte minus the one before plus 128). The reference writes
   its outputs through OIIO 3.1.17 (aov_write.h), which is not present here: these files are pinned by the format's
   published layout, a known-answer file written byte for byte from it, IEEE half rounding (numpy's float16) and
   zlib, not by files the reference wrote."""
import struct
import zlib

import numpy as np

FILM_MAGIC = 0x53523258
FILM_VERSION = 1
EXR_MAGIC = 20000630
NONE, RLE, ZIPS, ZIP = 0, 1, 2, 3
HALF, FLOAT, UINT = 1, 2, 0


def write_film(path, width, height, spp, done, sums, counts, odd=None):
    with open(path, "wb") as f:
        f.write(struct.pack("<8I", FILM_MAGIC, FILM_VERSION, width, height, spp, done, 0, 0))
        for arr, dt in ((sums, "<f8"), (counts, "<u4"), (odd, "<f8"), (None, "<f8"), (None, "<f4")):
            a = np.zeros(0, dt) if arr is None else np.ascontiguousarray(arr, dt).ravel()
            f.write(struct.pack("<Q", a.size))
            f.write(a.tobytes())
        f.write(struct.pack("<I", 0))


def read_film(path):
    with open(path, "rb") as f:
        b = f.read()
    magic, version, width, height, spp, done, channels, closest = struct.unpack_from("<8I", b, 0)
    at = 32
    arrays = []
    for dt in ("<f8", "<u4", "<f8", "<f8", "<f4"):
        (n,) = struct.unpack_from("<Q", b, at)
        at += 8
        size = np.dtype(dt).itemsize * n
        if at + size > len(b):
            raise ValueError("cut short")
        arrays.append(np.frombuffer(b, dt, n, at).copy())
        at += size
    (categories,) = struct.unpack_from("<I", b, at)
    at += 4
    return dict(magic=magic, version=version, width=width, height=height, spp=spp, done=done, channels=channels,
                closest=closest, sums=arrays[0].reshape(-1, 4), counts=arrays[1],
                odd=arrays[2].reshape(-1, 4) if arrays[2].size else None, aovs=arrays[3], depths=arrays[4],
                categories=categories, end=at, size=len(b))


def _cstring(b, at):
    end = b.index(b"\0", at)
    return b[at:end].decode(), end + 1


def unzip_block(data, expected):
    """zlib, then the predictor undone (each byte plus the one before minus 128), then the halves interleaved."""
    if len(data) >= expected:
        return data  # stored uncompressed
    t = bytearray(zlib.decompress(data))
    assert len(t) == expected
    for i in range(1, len(t)):
        t[i] = (t[i - 1] + t[i] - 128) & 0xFF
    out = bytearray(expected)
    half = (expected + 1) // 2
    out[0::2] = t[:half]
    out[1::2] = t[half:]
    return bytes(out)


def read_exr(path):
    with open(path, "rb") as f:
        b = f.read()
    magic, version = struct.unpack_from("<II", b, 0)
    assert magic == EXR_MAGIC, "not an OpenEXR file"
    assert version & 0xFF == 2 and not version & 0x200 and not version & 0x1000, "a single-part scanline file"
    at = 8
    attrs = {}
    while True:
        name, at = _cstring(b, at)
        if not name:
            break
        typ, at = _cstring(b, at)
        (size,) = struct.unpack_from("<i", b, at)
        at += 4
        attrs[name] = (typ, b[at:at + size])
        at += size
    # channels
    typ, v = attrs["channels"]
    assert typ == "chlist"
    channels = []
    p = 0
    while v[p] != 0:
        name, p = _cstring(v, p)
        ptype, linear, xs, ys = struct.unpack_from("<iB3xii", v, p)
        p += 16
        channels.append((name, ptype, linear, xs, ys))
    assert p == len(v) - 1
    assert [c[0] for c in channels] == sorted(c[0] for c in channels), "channels sorted by name"
    assert attrs["compression"][0] == "compression"
    compression = attrs["compression"][1][0]
    dw = struct.unpack("<4i", attrs["dataWindow"][1])
    disp = struct.unpack("<4i", attrs["displayWindow"][1])
    assert attrs["lineOrder"] == ("lineOrder", b"\0")
    for req in ("pixelAspectRatio", "screenWindowCenter", "screenWindowWidth"):
        assert req in attrs, req
    width, lines = dw[2] - dw[0] + 1, dw[3] - dw[1] + 1
    per = {NONE: 1, ZIPS: 1, ZIP: 16}[compression]
    chunks = (lines + per - 1) // per
    offsets = struct.unpack_from(f"<{chunks}Q", b, at)
    assert offsets[0] == at + 8 * chunks and list(offsets) == sorted(offsets)
    sizes = {HALF: 2, FLOAT: 4, UINT: 4}
    dtypes = {HALF: "<f2", FLOAT: "<f4", UINT: "<u4"}
    out = {c[0]: np.zeros((lines, width), dtypes[c[1]]) for c in channels}
    for k, off in enumerate(offsets):
        y, n = struct.unpack_from("<ii", b, off)
        assert y == dw[1] + k * per
        rows = min(per, dw[3] - y + 1)
        expected = rows * width * sum(sizes[c[1]] for c in channels)
        raw = b[off + 8:off + 8 + n] if compression == NONE else unzip_block(b[off + 8:off + 8 + n], expected)
        assert len(raw) == expected
        q = 0
        for r in range(rows):
            for name, ptype, *_ in channels:
                m = width * sizes[ptype]
                out[name][y - dw[1] + r] = np.frombuffer(raw, dtypes[ptype], width, q)
                q += m
    return dict(attrs=attrs, channels=channels, compression=compression, data_window=dw, display_window=disp,
                offsets=offsets, pixels=out)


def beauty(sums, counts):
    """Each pixel's mean as the reference's result() makes it: float(sum / count), 0 without a sample."""
    c = counts.astype(np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        m = np.where(c[:, None] > 0, sums / np.maximum(c, 1)[:, None], 0.0)
    return m.astype(np.float32)
