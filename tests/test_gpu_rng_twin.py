"""The device's sample-identity RNG against the oracle's (which is pinned bit for bit against the reference's
sample_stream.h), evaluated on the device through x2dev_rng_twin: samplingWord over every policy-dimension class,
paths::Stream's first 64 draws at several bounces, and childBranch — all bit-exact on random identities."""
import ctypes

import numpy as np
import pytest

import oracle
from x2amd import abi
from x2amd.lane import GpuLane

pytestmark = pytest.mark.gpu


def test_device_rng_is_bit_exact():
    rng = np.random.default_rng(20261203)
    n, draws = 4096, 64
    q = (abi.RngQuery * n)()
    dims = np.concatenate([np.arange(0, 32), rng.integers(0, 4096, n - 32)])
    for i in range(n):
        q[i].pixel = int(rng.integers(0, 1 << 31))
        q[i].index = int(rng.integers(0, 1 << 63, dtype=np.uint64)) if i % 3 == 0 else int(rng.integers(0, 1 << 20))
        q[i].stream = int(rng.integers(0, 1 << 63, dtype=np.uint64))
        q[i].branch = int(rng.integers(0, 1 << 63, dtype=np.uint64)) if i % 2 else 0
        q[i].dimension = int(dims[i])
        q[i].policy = int(i % 16)
        q[i].bounce = int(i % 13)
        q[i].a, q[i].b = int(rng.integers(0, 1 << 16)), int(rng.integers(0, 1 << 16))
    words = (ctypes.c_uint32 * n)()
    out = (ctypes.c_float * (n * draws))()
    branches = (ctypes.c_uint64 * n)()
    lane = GpuLane(0)
    try:
        abi.check(lane.lib.x2dev_rng_twin(lane.h, q, n, draws, words, out, branches), "x2dev_rng_twin")
    finally:
        lane.close()
    lib = oracle.load()
    got = np.frombuffer(out, np.float32).reshape(n, draws)
    buf = (ctypes.c_float * draws)()
    for i in range(n):
        r = q[i]
        assert words[i] == lib.oracle_sampling_word(r.pixel, r.index, r.stream, r.branch, r.dimension, r.policy), i
        lib.oracle_stream(r.pixel, r.index, r.stream, r.branch, r.bounce, draws, buf)
        assert np.array_equal(got[i].view(np.uint32), np.frombuffer(buf, np.float32).view(np.uint32)), i
        assert branches[i] == lib.oracle_child_branch(r.branch, r.a, r.b), i


def test_rng_twin_argument_errors():
    lane = GpuLane(0)
    try:
        q = (abi.RngQuery * 1)()
        w = (ctypes.c_uint32 * 1)()
        b = (ctypes.c_uint64 * 1)()
        assert lane.lib.x2dev_rng_twin(lane.h, q, 0, 0, w, None, b) < 0
        assert lane.lib.x2dev_rng_twin(lane.h, q, 1, 300, w, None, b) < 0
    finally:
        lane.close()
