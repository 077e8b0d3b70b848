"""The Clopper-Pearson bounds of the plan verification (dr_cp.h, restating
clopperPearsonLowerBound / UpperBound of state.h:134-151, which call
Boost.Math's beta quantile): the host build of the same header against
scipy.stats.beta.ppf over the verification's range of n and alpha, the
reference's special cases k = 0 and k = n, and monotonicity in k."""
import os
import subprocess

import numpy as np
import pytest
from scipy.stats import beta as beta_dist

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def cp_exe(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("cp") / "cp_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", os.path.join(ROOT, "tests", "cpp", "cp_check.cpp"), "-o", out],
                   check=True)
    return out


def run(exe, triples):
    inp = "\n".join(f"{k} {n} {a!r}" for k, n, a in triples) + "\n"
    out = subprocess.run([exe], input=inp, capture_output=True, text=True, check=True, timeout=120).stdout.split("\n")
    return np.array([[float(x) for x in l.split()] for l in out if l.strip()])


def test_bounds_match_scipy(cp_exe):
    rng = np.random.default_rng(0)
    triples = []
    for n in (1, 2, 10, 1000, 2 ** 17, 10 ** 6):
        ks = {0, 1, 2, n // 2, n - 1, n} | set(int(x) for x in rng.integers(0, n + 1, 6))
        for k in sorted(x for x in ks if 0 <= x <= n):
            for alpha in (0.5e-5, 1e-5, 1e-3, 0.05):
                triples.append((k, n, alpha))
    got = run(cp_exe, triples)
    for (k, n, a), (lo, hi) in zip(triples, got):
        want_lo = 0.0 if k == 0 else beta_dist.ppf(a, k, n - k + 1)
        want_hi = 1.0 if k == n else beta_dist.ppf(1 - a, k + 1, n - k)
        assert lo == pytest.approx(want_lo, rel=1e-9, abs=1e-300), (k, n, a)
        assert hi == pytest.approx(want_hi, rel=1e-9, abs=1e-300), (k, n, a)
        assert lo <= k / n <= hi


def test_monotone_in_failures(cp_exe):
    n = 2 ** 17
    got = run(cp_exe, [(k, n, 0.5e-5) for k in range(0, 400, 7)])
    assert np.all(np.diff(got[:, 0]) > 0) and np.all(np.diff(got[:, 1]) > 0)
