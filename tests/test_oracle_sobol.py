"""Pins the oracle's Owen-scrambled Sobol / ZSobol sampler.

Mirrors the reference's tests/unit/sobol.test.ts:15-90 (elementary-interval
nets, Owen bijectivity, ZSobol bijection, blue-noise-first policy, hashed
dimensions) and cross-checks the C restatement of sobol.wgsl.ts against a
Python restatement of the reference's independent test oracle
(tests/unit/sobolReference.ts:1-69).
"""
import numpy as np
import pytest

M32 = 0xFFFFFFFF


# -- Python restatement of tests/unit/sobolReference.ts (independent form) --
def _imul(a, b):
    return (a * b) & M32


def ref_hash(x):
    x = _imul(x ^ (x >> 16), 0x7FEB352D)
    x = _imul(x ^ (x >> 15), 0x846CA68B)
    return (x ^ (x >> 16)) & M32


def ref_seed(a, b):
    return ref_hash(a ^ ref_hash((b + 0x9E3779B9) & M32))


def ref_reverse(x):
    return int("{:032b}".format(x & M32)[::-1], 2)


def ref_owen(x, s):
    x = (ref_reverse(x) + s) & M32
    for c in (0x6C50B47C, 0xB82F1E52, 0xC7AFE638, 0x8D22F6E6):
        x ^= _imul(x, c)
    return ref_reverse(x)


def ref_sobol2d(index, s):
    i = ref_owen(index, ref_seed(s, 0))
    x = ref_owen(ref_reverse(i), ref_seed(s, 1))
    y, v = 0, 0x80000000
    while i:
        if i & 1:
            y ^= v
        i >>= 1
        v ^= v >> 1
    return (x >> 8, ref_owen(y, ref_seed(s, 2)) >> 8)


def ref_morton2(x, y, bits=8):
    r = 0
    for b in range(bits):
        r |= ((x >> b) & 1) << (2 * b) | ((y >> b) & 1) << (2 * b + 1)
    return r


ZPERM = [[0, 1, 2, 3], [0, 1, 3, 2], [0, 2, 1, 3], [0, 2, 3, 1], [0, 3, 2, 1], [0, 3, 1, 2],
         [1, 0, 2, 3], [1, 0, 3, 2], [1, 2, 0, 3], [1, 2, 3, 0], [1, 3, 2, 0], [1, 3, 0, 2],
         [2, 1, 0, 3], [2, 1, 3, 0], [2, 0, 1, 3], [2, 0, 3, 1], [2, 3, 0, 1], [2, 3, 1, 0],
         [3, 1, 2, 0], [3, 1, 0, 2], [3, 2, 1, 0], [3, 2, 0, 1], [3, 0, 2, 1], [3, 0, 1, 2]]
ZPACKED = [sum(d << (2 * i) for i, d in enumerate(p)) for p in ZPERM]


def ref_zindex(m, dim, log_spp, tile_bits=8):
    odd = log_spp & 1
    digits = tile_bits + ((log_spp + 1) >> 1)
    r = 0
    for d in range(digits - 1, odd - 1, -1):
        shift = 2 * d - odd
        higher = m >> (shift + 2) if shift < 30 else 0
        p = (ref_hash(higher ^ _imul(0x55555555, dim)) >> 24) % 24
        r |= ((ZPACKED[p] >> (2 * ((m >> shift) & 3))) & 3) << shift
    if odd:
        r |= (m & 1) ^ (ref_hash((m >> 1) ^ _imul(0x55555555, dim)) & 1)
    return r & M32


def ref_sample(pattern, px, py, width, index, dim, remaining_log, seq_seed):
    s = ref_seed(ref_seed((px + py * width) & M32, dim), seq_seed)
    if pattern == "zsobol":
        first = index == 0
        log = 0 if first else remaining_log
        ordinal = 0 if first else (index - 1) & M32
        block = (ordinal >> log) + (0 if first else 1)
        index = ref_zindex(((ref_morton2(px & 255, py & 255) << log) | (ordinal & ((1 << log) - 1))) & M32, dim, log)
        s = ref_seed(ref_seed(dim, seq_seed), ref_seed(px >> 8, py >> 8))
        s = ref_seed(s, block)
    x, y = ref_sobol2d(index, s)
    return [np.float32(0.0005) + np.float32(0.999) * np.float32(v / 16777216.0) for v in (x, y)]


def _net(points, k):
    pts = np.asarray(points, dtype=np.float64)
    assert np.all((pts >= 0) & (pts < 1))
    for a in range(k + 1):
        cell = np.floor(pts[:, 0] * 2 ** a).astype(np.int64) + \
            np.floor(pts[:, 1] * 2 ** (k - a)).astype(np.int64) * 2 ** a
        counts = np.bincount(cell, minlength=2 ** k)
        assert np.all(counts == 1), (a, k)


@pytest.mark.parametrize("seed", [0, 1, 12345, 0xFFFFFFFF])
def test_progressive_nets(orc, seed):
    for k in range(11):
        pts = orc.sobol2d(np.arange(2 ** k), seed)
        _net(pts, k)
    assert not np.array_equal(orc.sobol2d([0], 1), orc.sobol2d([0], 2))


def test_owen_bijective_nested_prefixes(orc):
    vals = [orc.owen(i, 77) for i in range(4096)]
    assert len(set(vals)) == 4096
    assert len({v >> 12 for v in vals}) == 1


@pytest.mark.parametrize("k", range(8))
def test_zsobol_bijection_and_pixel_nets(orc, k):
    for dim in (0, 1, 9, 0x80000001):
        indices = []
        for y in range(4):
            for x in range(4):
                px = [ref_zindex((ref_morton2(x, y, 2) << k) | i, dim, k, 2) for i in range(2 ** k)]
                indices += px
                _net(orc.sobol2d(px, 123), k)
        assert len(set(indices)) == 16 * 2 ** k
        assert max(indices) < 16 * 2 ** k


def test_zsobol_c_matches_reference_oracle(orc):
    rng = np.random.default_rng(5)
    for _ in range(400):
        m = int(rng.integers(0, 2 ** 32))
        dim = int(rng.integers(0, 2 ** 32))
        log = int(rng.integers(0, 17))
        assert orc.zsobol_index(m, dim, log) == ref_zindex(m, dim, log)
    assert orc.zsobol_index(0xFFFFFFFF, 0, 16) >= 0


def test_sobol_c_matches_reference_oracle(orc):
    rng = np.random.default_rng(7)
    idx = rng.integers(0, 2 ** 32, size=300, dtype=np.uint64).astype(np.uint32)
    for seed in (0, 11, 0xFFFFFFFF, 0x80001234):
        got = orc.sobol2d_bits(idx, seed)
        exp = np.array([ref_sobol2d(int(i), seed) for i in idx], dtype=np.uint32)
        assert np.array_equal(got, exp)


def test_direction_recurrence_on_basis_vectors(orc):
    # Sobol's second dimension is linear over GF(2): D1(a ^ b) = D1(a) ^ D1(b),
    # and D1(2^k) is the k-th direction number v_k = v_{k-1} ^ (v_{k-1} >> 1).
    v = 0x80000000
    for k in range(32):
        assert orc.sobol_d1(1 << k) == v
        v ^= v >> 1
    rng = np.random.default_rng(3)
    for _ in range(200):
        a, b = (int(x) for x in rng.integers(0, 2 ** 32, size=2))
        assert orc.sobol_d1(a ^ b) == orc.sobol_d1(a) ^ orc.sobol_d1(b)


def test_structured_samples_match_reference_policy(orc):
    # sobol.test.ts:52-75 blue-noise-first and per-pixel Sobol policies.
    for seed in (0, 11, 0xFFFFFFFF):
        for px, py in ((0, 0), (255, 256), (512, 91)):
            first = orc.structured_sample(1, px, py, 1024, 0, 14, 0, seed)
            for log in (0, 1, 4, 16):
                assert np.array_equal(orc.structured_sample(1, px, py, 1024, 0, 14, log, seed), first)
                assert not np.array_equal(orc.structured_sample(1, px, py, 1024, 1, 14, log, seed), first)
            for index in (0, 1, 5, 77):
                for log in (0, 3, 7):
                    got = orc.structured_sample(1, px, py, 1024, index, 14, log, seed)
                    exp = ref_sample("zsobol", px, py, 1024, index, 14, log, seed)
                    assert np.array_equal(got, np.array(exp, np.float32))
    for i in (0, 1, 17, 65536):
        got = orc.structured_sample(0, 256, 9, 1024, i, 0x80001234, 8, 12)
        exp = ref_sample("sobol", 256, 9, 1024, i, 0x80001234, 8, 12)
        assert np.array_equal(got, np.array(exp, np.float32))
