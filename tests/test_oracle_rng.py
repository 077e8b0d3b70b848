"""Pins of the oracle's noise: Philox4x32-10 against published known-answer
vectors and rocRAND, the transcendental kernels against libm in FP64, and
the Box-Muller normals against their closed form and distribution."""
import json
import math
import os

import numpy as np
import pytest
from scipy import stats

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def ulp32(x):
    return np.spacing(np.float32(abs(x)))


def test_philox_random123_kat(oracle, pkg):
    with open(os.path.join(GOLDEN, "philox_kat.json")) as f:
        kat = json.load(f)["vectors"]
    for v in kat:
        ctr = [int(x, 16) for x in v["ctr"]]
        key = [int(x, 16) for x in v["key"]]
        out = [int(x, 16) for x in v["out"]]
        assert oracle.philox(ctr, key) == out
        assert list(pkg.philox.philox4x32_10(np.array(ctr, dtype=np.uint32), key)) == out


def test_philox_matches_rocrand(oracle, pkg):
    with open(os.path.join(GOLDEN, "philox_rocrand.json")) as f:
        vecs = json.load(f)
    assert len(vecs) == 28
    ctr = np.array([v["ctr"] for v in vecs], dtype=np.uint32)
    for v in vecs:
        assert oracle.philox(v["ctr"], v["key"]) == v["out"]
        assert list(pkg.philox.philox4x32_10(np.array(v["ctr"], dtype=np.uint32), v["key"])) == v["out"]
    assert ctr.shape == (28, 4)


def test_log32_against_libm(oracle):
    # dense sweep over (0, 1]: every 24-bit uniform the generator can produce, subsampled
    ks = np.unique(np.concatenate([np.arange(1, 4097), np.linspace(1, 2 ** 24, 20001).astype(np.int64)]))
    worst = 0.0
    for k in ks:
        x = np.float32(k * 2.0 ** -24)
        got = oracle.log32(float(x))
        ref = math.log(float(x))
        if ref == 0.0:
            assert got == 0.0
            continue
        worst = max(worst, abs(got - ref) / float(ulp32(ref)))
    assert worst <= 2.0, worst


def test_exp32_against_libm(oracle):
    xs = np.concatenate([np.linspace(-86.0, 0.0, 40001), np.linspace(-1e-3, 1e-3, 2001), np.linspace(0, 88, 2001)])
    worst = 0.0
    for x in xs.astype(np.float32):
        got = oracle.exp32(float(x))
        ref = math.exp(float(x))
        worst = max(worst, abs(got - ref) / float(ulp32(ref)))
    assert worst <= 2.0, worst
    assert oracle.exp32(-86.5) == 0.0
    assert oracle.exp32(0.0) == 1.0
    assert math.isinf(oracle.exp32(89.0))


def test_sincos_against_libm(oracle):
    ks = np.concatenate([np.arange(0, 2 ** 24, 997), [0, 2 ** 21 - 1, 2 ** 21, 2 ** 22, 3 * 2 ** 22, 2 ** 24 - 1]])
    worst = 0.0
    for k in ks:
        s, c = oracle.sincos(int(k))
        a = 2.0 * math.pi * int(k) / 2 ** 24
        # absolute error in units of 2^-24 (values are O(1); near zeros relative ulps are meaningless)
        worst = max(worst, abs(s - math.sin(a)) * 2 ** 24, abs(c - math.cos(a)) * 2 ** 24)
    assert worst <= 2.0, worst


def test_normals_closed_form(oracle, pkg):
    # z0 = sqrt(-2 ln u1) cos(2 pi u2), z1 = ... sin, from the same Philox words
    key = (12345, 678)
    for i in range(2000):
        ctr = (i, i % 7, 3, 0)
        z = oracle.normals4(key, ctr)
        w = oracle.philox(list(ctr), list(key))
        for j in range(2):
            u1 = ((w[2 * j] >> 8) + 1) * 2.0 ** -24
            u2 = (w[2 * j + 1] >> 8) * 2.0 ** -24
            r = math.sqrt(-2.0 * math.log(u1))
            assert abs(z[2 * j] - r * math.cos(2 * math.pi * u2)) <= 4e-6 * max(1.0, r)
            assert abs(z[2 * j + 1] - r * math.sin(2 * math.pi * u2)) <= 4e-6 * max(1.0, r)


def test_normals_distribution(oracle):
    # 2^16 draws from the noise layout the engine uses (k, t, step, 0)
    zs = []
    for k in range(256):
        for t in range(64):
            zs.append(oracle.normals4((7, 11), (k, t, 5, 0)))
    z = np.concatenate(zs).astype(np.float64)
    assert abs(z.mean()) < 0.02
    assert abs(z.std() - 1.0) < 0.02
    assert abs(stats.skew(z)) < 0.05
    assert abs(stats.kurtosis(z)) < 0.1
    assert stats.kstest(z, "norm").pvalue > 1e-3
    # the four lanes of one block are uncorrelated
    b = np.stack(zs)
    cc = np.corrcoef(b.T)
    assert np.max(np.abs(cc - np.eye(4))) < 0.03
