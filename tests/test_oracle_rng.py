"""The oracle's RNG against the reference's known-answer tests and invariants (sample_fixture.h, qmc_fixture.h),
published known answers (Random123 Philox4x32-10, SplitMix64) and, where the reference tree was present at build
time, the reference's own sample_stream.h compiled in place (oracle/_ref)."""
import ctypes

import numpy as np
import pytest

import oracle
import reflib

L = oracle.load()
u32a = ctypes.c_uint32 * 4
STREAM = 0x8E1076D9E24A326B


def philox(c, k):
    out = u32a()
    L.oracle_philox(u32a(*c), (ctypes.c_uint32 * 2)(*k), out)
    return list(out)


def test_philox_known_answers():  # sample_fixture.h:17-20 (Random123's KATs)
    assert philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    m = 0xFFFFFFFF
    assert philox([m, m, m, m], [m, m]) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_unit_exact():  # sample_fixture.h:50
    assert L.oracle_unit(0) == 0.0
    assert L.oracle_unit(0xFFFFFFFF) == np.float32(1.0) - np.float32(2.0 ** -24)
    for bits in (1 << 8, 0x12345678, 0x80000000):
        assert L.oracle_unit(bits) == (bits >> 8) * 2.0 ** -24


def test_word_is_block_word():  # sample_fixture.h:51-53
    out = u32a()
    L.oracle_block(7, 0xFEDCBA9876543210, 11, 9, 0x3FFFFFFF, out)
    for i in range(4):
        assert L.oracle_word(7, 0xFEDCBA9876543210, 11, 9, 0xFFFFFFFC + i) == out[i]


def test_sobol_first_values():  # qmc_fixture.h:71-72
    expected = [[0, 4, 6, 2, 3, 7, 5, 1], [0, 4, 2, 6, 3, 7, 1, 5], [0, 4, 2, 6, 5, 1, 7, 3], [0, 4, 2, 6, 7, 3, 5, 1]]
    for d in range(4):
        for i in range(8):
            assert L.oracle_sobol(i, d) == expected[d][i] << 29


def test_sobol_directions_recurrence_properties():
    """Direction number k (0-based) is m * 2^(31-k) with m odd: bit 31-k set and every bit below it clear -- the
    defining property of a Sobol direction table -- and row 0 is the van der Corput sequence."""
    out = (ctypes.c_uint32 * (17 * 32))()
    L.oracle_directions(out)
    d = np.array(out, dtype=np.uint64).reshape(17, 32)
    for k in range(32):
        assert d[0, k] == 1 << (31 - k)
        for dim in range(17):
            assert (int(d[dim, k]) >> (31 - k)) & 1 == 1
            assert int(d[dim, k]) & ((1 << (31 - k)) - 1) == 0


def test_prefix_stratification():  # qmc_fixture.h:73-77: every power-of-two prefix fills each 1D stratum once
    for pixel in range(4):
        for m in range(1, 11):
            for dim in (0, 1, 16, 17):
                cells = {L.oracle_sampling_word(pixel, i, 19, 3, dim, 3) >> (32 - m) for i in range(1 << m)}
                assert len(cells) == 1 << m


def test_light_dimensions_stratified():  # policy bit 4: dimensions 18/19 stratified the same way
    for m in range(1, 9):
        for dim in (18, 19):
            cells = {L.oracle_sampling_word(3, i, 19, 5, dim, 15) >> (32 - m) for i in range(1 << m)}
            assert len(cells) == 1 << m


def test_policy_zero_is_philox():  # qmc_fixture.h:80
    for d in range(32):
        assert L.oracle_sampling_word(12, 0xFEDCBA9876543210, 713, 17, d, 0) == \
            L.oracle_word(12, 0xFEDCBA9876543210, 713, 17, d)


def test_policy_three_dimensions():
    """Policy 3 takes Sobol for the camera's dimensions 0-4 (pixel, lens, time) and the first BSDF sample's 16/17,
    Philox elsewhere (sample_stream.h:128-130). qmc_fixture.h:81 still lists only 0, 1, 16, 17: written before the
    lens and shutter dimensions took Sobol (sample_stream.h:78-81); the code is what this checks."""
    for d in range(32):
        w = L.oracle_sampling_word(12, 0xFEDCBA9876543210, 713, 17, d, 3)
        p = L.oracle_word(12, 0xFEDCBA9876543210, 713, 17, d)
        if d < 5 or d in (16, 17):
            assert w != p
        else:
            assert w == p
        assert 0.0 <= L.oracle_unit(w) < 1.0


def test_high_index_bits_change_scramble():  # qmc_fixture.h:84
    i = 0xFEDCBA9876543210
    assert L.oracle_sampling_word(12, i, 713, 17, 16, 3) != L.oracle_sampling_word(12, i ^ (1 << 32), 713, 17, 16, 3)


def test_camera_net():  # qmc_fixture.h:86-88: every elementary rectangle of the 256-point camera net holds one point
    for k in range(9):
        seen = set()
        for i in range(256):
            a = L.oracle_sampling_word(19, i, 713, 0, 0, 3) >> (32 - k) if k else 0
            b = L.oracle_sampling_word(19, i, 713, 0, 1, 3) >> (24 + k) if k < 8 else 0
            seen.add((a << (8 - k)) | b)
        assert len(seen) == 256


def test_owen_scramble_is_nested():
    """Burley's nested uniform scramble: bit k of the output depends only on bits >= k of the input (a prefix tree
    of flips), so it is a bijection that maps each dyadic interval onto one of the same size."""
    rng = np.random.default_rng(3)
    for seed in rng.integers(0, 2 ** 32, 8, dtype=np.uint64):
        outs = {L.oracle_owen_scramble(x << 24, int(seed)) >> 24 for x in range(256)}
        assert len(outs) == 256  # the top 8 bits of 256 distinct top-8-bit inputs stay distinct
        for x in rng.integers(0, 2 ** 32, 16, dtype=np.uint64):
            x = int(x)
            a = L.oracle_owen_scramble(x, int(seed))
            b = L.oracle_owen_scramble(x ^ 1, int(seed))  # flipping the lowest bit changes only the lowest bit
            assert (a ^ b) == 1


def test_child_branch_is_splitmix64():
    """childBranch(parent, a, b) is SplitMix64's output for state parent + golden * (a * 0x10001 + b + 1): from state
    0 the published sequence begins 0xe220a8397b1dcdaf, 0x6e789e6aa1b965f4, 0x06c45d188009454f."""
    assert L.oracle_child_branch(0, 0, 0) == 0xE220A8397B1DCDAF
    assert L.oracle_child_branch(0, 0, 1) == 0x6E789E6AA1B965F4
    assert L.oracle_child_branch(0, 0, 2) == 0x06C45D188009454F


def test_stream_layout():
    """paths::Stream draws dimension 16 + bounce * 256 + k (async_integrator.h:156-176)."""
    out = (ctypes.c_float * 9)()
    L.oracle_stream(5, 77, STREAM, 0x1234, 3, 9, out)
    for k in range(9):
        assert out[k] == L.oracle_unit(L.oracle_word(5, 77, STREAM, 0x1234, 16 + 3 * 256 + k))


# ---- against the reference's own sample_stream.h (oracle/_ref) ------------------------------------------------
def test_directions_match_reference():
    R = reflib.load()
    a = (ctypes.c_uint32 * (17 * 32))()
    b = (ctypes.c_uint32 * (17 * 32))()
    L.oracle_directions(a)
    R.ref_directions(b)
    assert list(a) == list(b)


def test_rng_matches_reference():
    R = reflib.load()
    rng = np.random.default_rng(11)
    for _ in range(3000):
        pixel = int(rng.integers(0, 2 ** 32))
        index = int(rng.integers(0, 2 ** 63)) if rng.random() < 0.3 else int(rng.integers(0, 4096))
        stream = int(rng.integers(0, 2 ** 63))
        branch = int(rng.integers(0, 2 ** 63))
        dim = int(rng.integers(0, 40))
        policy = int(rng.integers(0, 16))
        assert L.oracle_sampling_word(pixel, index, stream, branch, dim, policy) == \
            R.ref_sampling_word(pixel, index, stream, branch, dim, policy)
        group = int(rng.integers(0, 2 ** 32))
        a, b = u32a(), u32a()
        L.oracle_block(pixel, index, stream, branch, group, a)
        R.ref_block(pixel, index, stream, branch, group, b)
        assert list(a) == list(b)
        x, s = int(rng.integers(0, 2 ** 32)), int(rng.integers(0, 2 ** 32))
        assert L.oracle_owen_scramble(x, s) == R.ref_owen_scramble(x, s)
        assert L.oracle_sobol(x, dim % 17) == R.ref_sobol(x, dim % 17)
