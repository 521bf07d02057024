"""Pins of the oracle's spec math: Philox4x32-10 against the published
known-answer vectors and an independent pure-Python restatement; pow / log
/ exp / sin / cos against numpy's double-precision libm rounded to float
(<= 1 ulp) and against closed forms."""
import math

import numpy as np
import pytest

import pyoracle as O

# Random123 known-answer vectors for philox4x32 with 10 rounds
# (Salmon et al., SC'11): counter (4 words), key (2 words), expected output
PHILOX_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


def philox_py(ctr, key):
    """Independent restatement of Philox4x32-10 in Python integers."""
    M = 0xFFFFFFFF
    c = list(ctr)
    k0, k1 = key
    for r in range(10):
        if r:
            k0 = (k0 + 0x9E3779B9) & M
            k1 = (k1 + 0xBB67AE85) & M
        p0 = 0xD2511F53 * c[0]
        p1 = 0xCD9E8D57 * c[2]
        c = [((p1 >> 32) ^ c[1] ^ k0) & M, p1 & M, ((p0 >> 32) ^ c[3] ^ k1) & M, p0 & M]
    return tuple(c)


@pytest.mark.parametrize("ctr,key,expected", PHILOX_KAT)
def test_philox_known_answers(oracle, ctr, key, expected):
    assert tuple(int(x) for x in oracle.philox(ctr, key)) == expected
    assert philox_py(ctr, key) == expected


def test_philox_matches_python(oracle):
    rng = np.random.default_rng(1)
    for _ in range(200):
        ctr = tuple(int(x) for x in rng.integers(0, 2**32, 4, dtype=np.uint64))
        key = tuple(int(x) for x in rng.integers(0, 2**32, 2, dtype=np.uint64))
        assert tuple(int(x) for x in oracle.philox(ctr, key)) == philox_py(ctr, key)


def test_u01_closed_form(oracle):
    for x in (0, 1, 255, 256, 0x12345678, 0xFFFFFFFF):
        assert oracle.u01(x) == np.float32((x >> 8) / 16777216.0)
    assert oracle.u01(0xFFFFFFFF) < 1.0


def ulp_diff(a, b):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    ia = a.view(np.int32).astype(np.int64)
    ib = b.view(np.int32).astype(np.int64)
    ia = np.where(ia < 0, -(ia & 0x7FFFFFFF), ia)
    ib = np.where(ib < 0, -(ib & 0x7FFFFFFF), ib)
    return np.abs(ia - ib)


def test_pow_against_libm(oracle):
    rng = np.random.default_rng(2)
    xs = np.exp(rng.uniform(-20, 20, 20000)).astype(np.float32)
    ys = rng.uniform(-3, 3, 20000).astype(np.float32)
    ys[:5000] = np.float32(0.3333333)
    got = np.array([oracle.powf(float(x), float(y)) for x, y in zip(xs, ys)], np.float32)
    ref = np.power(xs.astype(np.float64), ys.astype(np.float64)).astype(np.float32)
    ok = np.isfinite(ref)
    assert ulp_diff(got[ok], ref[ok]).max() <= 1


def test_pow_closed_forms(oracle):
    assert oracle.powf(8.0, np.float32(1.0 / 3.0)) == np.float32(8.0 ** float(np.float32(1.0 / 3.0)))
    assert oracle.powf(2.0, 10.0) == 1024.0
    assert oracle.powf(0.25, 0.5) == 0.5
    assert oracle.powf(0.0, 0.3333333) == 0.0
    assert oracle.powf(5.0, 0.0) == 1.0
    assert oracle.powf(1.0, 7.5) == 1.0


def test_log_exp_against_libm(oracle):
    rng = np.random.default_rng(3)
    xs = np.exp(rng.uniform(-80, 80, 20000)).astype(np.float32)
    got = np.array([oracle.logf(float(x)) for x in xs], np.float32)
    ref = np.log(xs.astype(np.float64)).astype(np.float32)
    assert ulp_diff(got, ref).max() <= 1
    zs = rng.uniform(-80, 80, 20000).astype(np.float32)
    got = np.array([oracle.expf(float(z)) for z in zs], np.float32)
    ref = np.exp(zs.astype(np.float64)).astype(np.float32)
    assert ulp_diff(got, ref).max() <= 1
    assert oracle.logf(1.0) == 0.0
    assert oracle.expf(0.0) == 1.0
    assert oracle.logf(0.0) == -math.inf


def test_sincos_against_libm(oracle):
    rng = np.random.default_rng(4)
    a = (np.float32(2.0) * np.float32(math.pi)) * rng.uniform(0, 1, 20000).astype(np.float32)
    a = a.astype(np.float32)
    res = np.array([oracle.sincosf(float(x)) for x in a], np.float32)
    ref_s = np.sin(a.astype(np.float64)).astype(np.float32)
    ref_c = np.cos(a.astype(np.float64)).astype(np.float32)
    # absolute error near the zeros of sin / cos, ulp error elsewhere
    for got, ref in ((res[:, 0], ref_s), (res[:, 1], ref_c)):
        big = np.abs(ref) > 1e-3
        assert ulp_diff(got[big], ref[big]).max() <= 1
        assert np.abs(got - ref).max() < 1e-7
    s, c = oracle.sincosf(0.0)
    assert (s, c) == (0.0, 1.0)
    s, c = oracle.sincosf(float(np.float32(math.pi / 6)))
    assert abs(s - 0.5) < 1e-7 and abs(c - math.sqrt(3) / 2) < 1e-7
