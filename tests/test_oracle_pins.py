"""Pins the CPU oracle's primitives against facts fixed outside this project.

* Philox4x32-10: Random123 known-answer vectors and rocRAND's open-source
  implementation (rocrand_philox4x32_10.h).
* Elementary functions: within 1 ulp of correctly rounded values (numpy
  double / numpy longdouble).
* Variates: exact special cases and chi-square goodness of fit against
  exact pmfs (scipy), with a planted-error check that the test rejects a
  wrong parameter.
* Parameter matrices: the reference's own invariants (tests/ParamsTest.cpp:57-66)
  and the closed form of genMatK / genP_i2j / genMatP.
"""
import ctypes
import math
import os
import shutil
import subprocess

import numpy as np
import pytest
from scipy import stats

from tests.conftest import ROOT

# Random123 kat_vectors, philox4x32 with 10 rounds: (counter, key) -> output
PHILOX_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,want", PHILOX_KAT)
def test_philox_known_answers(oracle, ctr, key, want):
    assert tuple(int(x) for x in oracle.philox(ctr, key)) == want


def test_philox_matches_rocrand(oracle, tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    exe = str(tmp_path / "rr_philox")
    subprocess.run([hipcc, "-x", "hip", "--offload-host-only", "-O1",
                    os.path.join(ROOT, "tests", "host", "rocrand_philox.cpp"), "-o", exe],
                   check=True, capture_output=True)
    rng = np.random.default_rng(1)
    q = rng.integers(0, 2**32, size=(500, 6), dtype=np.uint64)
    inp = "\n".join(" ".join(str(int(v)) for v in row) for row in q)
    res = subprocess.run([exe], input=inp, capture_output=True, text=True, check=True)
    got = np.array([[int(x) for x in line.split()] for line in res.stdout.strip().splitlines()])
    want = np.array([oracle.philox(r[:4], r[4:]) for r in q], dtype=np.int64)
    assert np.array_equal(got, want)


def test_uniform_conversions(oracle):
    L = oracle.lib()
    assert L.or_u01f(0) == 2.0 ** -24
    assert L.or_u01f(0xFFFFFFFF) == 1.0
    assert L.or_u01f(0xFF) == 2.0 ** -24      # low 8 bits dropped
    assert L.or_u53(0, 0) == 0.0
    assert L.or_u53(0xFFFFFFFF, 0xFFFFFFFF) == 1.0 - 2.0 ** -53
    assert L.or_u53(1 << 5, 0) == 2.0 ** -27


def _ulps32(got, ref):
    ref32 = np.float32(ref)
    return abs(float(got) - float(ref)) / float(np.spacing(abs(ref32)))


def test_logf_expf_within_one_ulp(oracle):
    L = oracle.lib()
    rng = np.random.default_rng(2)
    xs = np.concatenate([rng.uniform(0, 2, 4000), np.exp(rng.uniform(-85, 85, 4000)),
                         [1e-40, 1e-45, 3.4e38]]).astype(np.float32)
    worst = max(_ulps32(L.or_logf(float(x)), np.log(np.float64(x))) for x in xs
                if np.log(np.float64(x)) != 0)
    assert worst <= 1.0
    assert L.or_logf(1.0) == 0.0
    assert L.or_logf(0.0) == -math.inf
    assert math.isnan(L.or_logf(-1.0))
    xs = rng.uniform(-103, 88.7, 8000).astype(np.float32)
    worst = max(_ulps32(L.or_expf(float(x)), math.exp(float(x))) for x in xs
                if math.exp(float(x)) > 1.2e-38)
    assert worst <= 1.0
    assert L.or_expf(0.0) == 1.0
    assert L.or_expf(-200.0) == 0.0
    assert L.or_expf(100.0) == math.inf


def test_log_exp_double_within_one_ulp(oracle):
    L = oracle.lib()
    rng = np.random.default_rng(3)
    for x in np.concatenate([rng.uniform(0, 2, 3000), np.exp(rng.uniform(-700, 700, 3000))]):
        ref = np.log(np.longdouble(x))
        if ref == 0:
            continue
        assert abs(np.longdouble(L.or_log(float(x))) - ref) <= np.spacing(abs(np.float64(ref)))
    for x in rng.uniform(-700, 700, 6000):
        ref = np.exp(np.longdouble(x))
        assert abs(np.longdouble(L.or_exp(float(x))) - ref) <= np.spacing(np.float64(ref))
    assert L.or_log(1.0) == 0.0 and L.or_exp(0.0) == 1.0


def test_cos_turns_and_normal(oracle):
    L = oracle.lib()
    for u in np.linspace(0, 1, 4001, dtype=np.float32):
        assert abs(L.or_cos2pif(float(u)) - math.cos(2 * math.pi * float(u))) < 2e-7
    rng = np.random.default_rng(4)
    w = rng.integers(0, 2**32, size=(100_000, 2), dtype=np.uint64)
    z = np.array([L.or_normal(int(a), int(b)) for a, b in w])
    assert stats.kstest(z, "norm").pvalue > 1e-3
    # planted error: a unit shift of the mean is rejected
    assert stats.kstest(z + 0.05, "norm").pvalue < 1e-6


def _words(oracle, n, seed):
    """n Philox words of a fixed key, as the walk draws them."""
    out = []
    for i in range((n + 3) // 4):
        out.extend(int(v) for v in oracle.philox([i, 0, 0, seed], [seed, 0]))
    return np.array(out[:n], dtype=np.uint64)


def test_categorical_first_above(oracle):
    L = oracle.lib()
    p = (ctypes.c_float * 4)(0.1, 0.2, 0.3, 0.4)
    assert L.or_categorical(4, p, 0.05) == 0
    assert L.or_categorical(4, p, 0.1) == 1       # strict '>' (src/gen_rand.cuh:43)
    assert L.or_categorical(4, p, 0.65) == 3
    q = (ctypes.c_float * 3)(0.2, 0.2, 0.2)
    assert L.or_categorical(3, q, 0.9) == 2        # fallback n - 1 (src/gen_rand.cuh:49)
    w = _words(oracle, 200_000, 5)
    idx = np.array([L.or_categorical(4, p, L.or_u01f(int(x))) for x in w])
    cnt = np.bincount(idx, minlength=4)
    assert stats.chisquare(cnt, np.array([0.1, 0.2, 0.3, 0.4]) * len(w)).pvalue > 1e-3


def _binom_samples(oracle, n, p, count, seed):
    L = oracle.lib()
    rng = np.random.default_rng(seed)
    w = rng.integers(0, 2**32, size=(count, 4), dtype=np.uint64)
    return np.array([L.or_binomial(n, p, int(a), int(b), int(c), 0, 0, int(d) & 0xFFFFFFFF,
                                   int(d) >> 3 & 0xFFFFFFFF, None) for a, b, c, d in w])


def _gof_discrete(x, pmf, support):
    obs = np.array([(x == k).sum() for k in support], dtype=float)
    exp = pmf(np.array(support)) * len(x)
    keep = exp >= 5
    obs_k, exp_k = obs[keep], exp[keep]
    obs_k = np.append(obs_k, len(x) - obs_k.sum())
    exp_k = np.append(exp_k, len(x) - exp_k.sum())
    if exp_k[-1] < 1e-9:
        obs_k, exp_k = obs_k[:-1], exp_k[:-1]
    return stats.chisquare(obs_k, exp_k * obs_k.sum() / exp_k.sum()).pvalue


@pytest.mark.parametrize("n,p", [(12, 0.3), (40, 0.75), (100, 0.5), (127, 0.62), (200, 0.1), (500, 0.3), (3000, 0.5),
                                 (2000, 0.93)])
def test_binomial_goodness_of_fit(oracle, n, p):
    x = _binom_samples(oracle, n, p, 20_000, 6 + n)
    pv = _gof_discrete(x, lambda k: stats.binom.pmf(k, n, p), range(n + 1))
    assert pv > 1e-3, (n, p, pv)
    # planted error: the same samples are rejected for a slightly different p
    pv_bad = _gof_discrete(x, lambda k: stats.binom.pmf(k, n, p * 1.05), range(n + 1))
    assert pv_bad < 1e-4


def test_binomial_special_cases(oracle):
    L = oracle.lib()
    for s in range(50):
        assert L.or_binomial(0, 0.4, s, 0, 0, 0, 0, s, s, None) == 0
        assert L.or_binomial(17, 0.0, s, 0, 0, 0, 0, s, s, None) == 0
        assert L.or_binomial(17, 1.0, s, 0, 0, 0, 0, s, s, None) == 17
        assert L.or_binomial(5000, 1.0, s, 0, 0, 0, 0, s, s, None) == 5000
        assert L.or_binomial(-3, 0.5, s, 0, 0, 0, 0, s, s, None) == 0
        assert L.or_binomial(9, math.nan, s, 0, 0, 0, 0, s, s, None) == 0


def _bg_pmf(lam, cap):
    """Exact law of draw_P_B_Tr (src/gen_rand.cuh:92-107): k ~ Poisson(lam), k if k <= cap,
    else floor(cap * U) with U uniform."""
    kmax = int(math.floor(cap))
    p = np.zeros(kmax + 1)
    p[:] = stats.poisson.pmf(np.arange(kmax + 1), lam)
    over = 1.0 - p.sum()
    if cap > 0:
        # floor(cap U) = k for U in [k/cap, (k+1)/cap)
        for k in range(kmax + 1):
            p[k] += over * (min(k + 1, cap) - k) / cap
    else:
        p[0] += over
    return p


@pytest.mark.parametrize("lam,cap", [(0.7, 3.0), (2.5, 1.9), (4.0, 12.0), (150.0, 160.0),
                                     (300.0, 20.0)])
def test_background_count_law(oracle, lam, cap):
    L = oracle.lib()
    rng = np.random.default_rng(int(lam * 10 + cap))
    n = 20_000
    w = rng.integers(0, 2**32, size=(n, 4), dtype=np.uint64)
    x = np.array([L.or_bg_capped(lam, cap, L.or_u01f(int(a)), L.or_u01f(int(b)), int(c), 0,
                                 int(d) & 0xFFFFFFFF, 0, 0, 0x100000, None) for a, b, c, d in w])
    assert np.all(x == np.floor(x)) and x.min() >= 0 and x.max() <= cap
    pmf = _bg_pmf(lam, cap)
    pv = _gof_discrete(x, lambda k: pmf[k], range(len(pmf)))
    assert pv > 1e-3, pv
    assert L.or_bg_capped(0.0, 5.0, 0.5, 0.5, 0, 0, 0, 0, 0, 0, None) == 0.0


@pytest.mark.parametrize("n", [2, 3, 4, 5, 6])
def test_param_matrices_reference_invariants(oracle, n):
    """tests/ParamsTest.cpp:10-66 with seeded draws: E in [0,1), k in [0,1000), v in [0,10)."""
    rng = np.random.default_rng(100 + n)
    args = np.concatenate([rng.uniform(0, 1, n), rng.uniform(0, 1000, n * (n - 1)),
                           rng.uniform(0, 10, n)]).astype(np.float32)
    p = oracle.make_params(n, args)
    K = np.array(p.K[:n * n], np.float32).reshape(n, n).T      # K[i, j] = K(i, j)
    P = np.array(p.Pi2j[:n * n], np.float32).reshape(n, n).T
    flatK = np.array(p.K[:n * n], np.float32)
    for i in range(n):
        assert K[0, i] == flatK[i * n]                                # column-major layout
    for j in range(n):
        assert abs(P[:, j].sum() - 1) < 1e-5                          # columns sum to 1
        if n == 3:
            i, ii = (j + 1) % n, (j + 2) % n
            assert abs(P[i, j] / P[ii, j] - K[i, j] / K[ii, j]) < 1e-4
    # closed form: off-diagonal args row-major over the target state, diag = -column sum
    kargs = args[n:n * n]
    it = iter(kargs)
    for i in range(n):
        for j in range(n):
            if i != j:
                assert K[i, j] == next(it)
    for j in range(n):
        assert abs(K[j, j] + (K[:, j].sum() - K[j, j])) <= 1e-3
        assert P[j, j] == 0
    p0 = np.array(p.p0[:n])
    assert np.allclose(p0, K.sum(axis=1) - np.diag(K), rtol=1e-5)
    assert np.allclose(np.array(p.E[:n]), args[:n]) and np.allclose(np.array(p.v[:n]), args[n * n:])


def test_param_bad_counts(oracle):
    with pytest.raises(ValueError):
        oracle.make_params(3, np.zeros(11, np.float32))
    with pytest.raises(ValueError):
        oracle.make_params(1, np.zeros(2, np.float32))
