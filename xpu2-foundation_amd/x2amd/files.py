"""The on-disk formats either side of a frame, over the device library (x2_film_save / x2_film_load / x2_write_exr in
include/x2dev.h): a checkpoint render's film save ("X2RS", StagedFilm in xpu2/interactive/xpu2_session.cc) and the
beauty output as a scanline OpenEXR file (aov_write.h's RenderOutputWriter). Host code; no GPU is needed."""
import ctypes

import numpy as np

from . import abi

P64 = ctypes.POINTER(ctypes.c_double)
P32 = ctypes.POINTER(ctypes.c_uint32)


def save_film(path, width, height, spp, done, sums, counts, odd=None):
    """StagedFilm::save: the film (sums [pixels, 4] float64, counts [pixels] uint32, optional odd sums) to path."""
    s = np.ascontiguousarray(sums, np.float64)
    c = np.ascontiguousarray(counts, np.uint32)
    o = None if odd is None else np.ascontiguousarray(odd, np.float64)
    assert s.size == width * height * 4 and c.size == width * height
    f = abi.FilmFile(width, height, spp, done, s.ctypes.data_as(P64), c.ctypes.data_as(P32),
                     o.ctypes.data_as(P64) if o is not None else None)
    abi.check(abi.load().x2_film_save(path.encode(), ctypes.byref(f)), "x2_film_save")


def load_film(path, width, height, spp):
    """StagedFilm::load: (sums, counts, odd or None, done) of a save of this render; raises when there is none, it is
    not one, it is another render's or it is cut short."""
    n = width * height
    sums = np.zeros((n, 4), np.float64)
    counts = np.zeros(n, np.uint32)
    odd = np.zeros((n, 4), np.float64)
    done, has_odd = ctypes.c_uint32(), ctypes.c_uint32()
    abi.check(abi.load().x2_film_load(path.encode(), width, height, spp, sums.ctypes.data_as(P64),
                                      counts.ctypes.data_as(P32), odd.ctypes.data_as(P64), ctypes.byref(done),
                                      ctypes.byref(has_odd)), "x2_film_load")
    return sums, counts, (odd if has_odd.value else None), done.value


def write_exr(path, width, height, sums, counts, region=None, half=True, compression=abi.X2_EXR_ZIP, weight=False):
    """The beauty (A, B, G, R: each pixel's mean) and optionally its weight as a scanline OpenEXR file."""
    s = np.ascontiguousarray(sums, np.float64)
    c = np.ascontiguousarray(counts, np.uint32)
    im = abi.ExrImage(width, height, s.ctypes.data_as(P64), c.ctypes.data_as(P32),
                      (abi.i32 * 4)(*(region or (0, 0, 0, 0))), 1 if half else 0, compression, 1 if weight else 0, 0)
    abi.check(abi.load().x2_write_exr(path.encode(), ctypes.byref(im)), "x2_write_exr")
