"""Pins of the oracle's Rao-Blackwellised (RB) form (CPU only).

* The walk summary is the plain walk's: same F bit for bit, acceptor draws
  never above the photon count the RB form convolves.
* The spread pmf is the exact law of the acceptor total: a unit's fixed-point
  histogram equals the convolution of scipy's binomial pmfs binned at a / F.
* The FP64 efficiency gradient equals central differences of the FP64
  expected histogram (the walk does not depend on E when v = 0).
* The score has mean zero under the walk's own law, and a planted mistake
  (1/K instead of 2/K for the chosen transition) is rejected.
* The likelihood-ratio rate gradient agrees, over independent seeds, with
  central differences of the expected histogram under common random numbers
  (the finite differences are the noisier side: they sum discrete jumps of
  photon counts; the many-seed comparison at scale runs on the GPU).
"""
import numpy as np
import pytest
from scipy import stats

from gsmfretda_amd import synth

PH = (1.0, 0.0, 0.05, 52.0)


def _data(n_bursts=60, seed=5, bg=True):
    args = synth.make_args([0.2, 0.55, 0.85], [900, 2500, 700, 3000, 1500, 5000])
    d = synth.generate(n_bursts, args, 3, seed=seed)
    if not bg:
        d.bg_ad = d.bg_dd = 0.0
    return d


ARGS3 = synth.make_args([0.25, 0.5, 0.8], [800, 2000, 600, 2500, 1200, 4000])


def test_rb_walk_is_the_plain_walk(oracle):
    d = _data()
    od = oracle.Data.from_bursts(d)
    p = oracle.make_params(3, ARGS3, *PH)
    o = oracle.run(od, p, R=4, seed=9, set_id=2)
    for b in range(d.nbursts):
        for r in range(4):
            cnt, F, _ = oracle.rb_unit(od, p, b, r, 9, 2)
            u = b * 4 + r
            assert np.float32(F) == o["F"][u]
            assert o["acc"][u] <= cnt.sum()


def test_rb_pmf_is_the_exact_binomial_convolution(oracle):
    d = _data(n_bursts=40)
    od = oracle.Data.from_bursts(d)
    p = oracle.make_params(3, ARGS3, *PH)
    B = 50
    checked = 0
    for b in range(d.nbursts):
        cnt, F, _ = oracle.rb_unit(od, p, b, 0, 3, 0)
        M = oracle.rb_run(od, p, 1, 3, 0, b0=b, b1=b + 1, B=B).astype(np.float64) * 2.0 ** -32
        pmf = np.array([1.0])
        for s in range(3):
            if cnt[s]:
                # the oracle's p_s is E_s after the distance round trip (equal
                # to E_s within float rounding)
                pmf = np.convolve(pmf, stats.binom.pmf(np.arange(cnt[s] + 1), cnt[s], ARGS3[s]))
        want = np.zeros(B)
        for a, w in enumerate(pmf):
            Ea = np.float32(a) / np.float32(F) if F > 0 else np.float32(0)
            if 0 <= Ea < 1:
                want[int(np.floor(np.float64(Ea) * B))] += w
        assert np.allclose(M, want, rtol=2e-5, atol=2e-7), f"burst {b}"
        checked += cnt.sum() > 0
    assert checked > 30


def test_rb_efficiency_gradient_matches_finite_differences(oracle):
    d = _data(n_bursts=80)
    od = oracle.Data.from_bursts(d)
    rng = np.random.default_rng(3)
    g = rng.normal(size=50)
    R, seed = 8, 11
    p0 = oracle.make_params(3, ARGS3, *PH)
    _, grad = oracle.rb_grad(od, p0, R, seed, g)
    for s in range(3):
        h = 1e-3 * ARGS3[s]
        vals = []
        for sgn in (1, -1):
            a = np.array(ARGS3, np.float64)
            a[s] += sgn * h
            a32 = a.astype(np.float32)
            hs = (np.float64(a32[s]) - np.float64(np.float32(ARGS3[s])))
            M, _ = oracle.rb_grad(od, oracle.make_params(3, a32, *PH), R, seed, g)
            vals.append((g @ M, hs))
        fd = (vals[0][0] - vals[1][0]) / (vals[0][1] - vals[1][1])
        assert abs(grad[s] - fd) <= 2e-3 * abs(fd) + 1e-6, f"E[{s}]: {grad[s]} vs {fd}"


def test_rb_scores_have_zero_mean(oracle):
    d = _data(n_bursts=150, seed=8)
    od = oracle.Data.from_bursts(d)
    p = oracle.make_params(3, ARGS3, *PH)
    out = [oracle.rb_unit(od, p, b, r, 21, 0, transitions=True)
           for b in range(d.nbursts) for r in range(16)]
    S = np.array([o[2] for o in out])
    N = np.array([o[3] for o in out], np.float64)
    z = S.mean(axis=0) / (S.std(axis=0) / np.sqrt(len(S)))
    assert np.all(np.abs(z) < 4.5), z
    assert N.sum() > 1000
    # planted mistake: 1/K instead of 2/K for the chosen transition (the
    # categorical part of the score dropped) -- rejected
    k = np.asarray(ARGS3[3:9], np.float64)
    bad = S - N / k
    zb = bad.mean(axis=0) / (bad.std(axis=0) / np.sqrt(len(bad)))
    assert np.any(np.abs(zb) > 6), zb


def test_rb_rate_gradient_matches_finite_differences_over_seeds(oracle):
    d = _data(n_bursts=40, seed=12)
    od = oracle.Data.from_bursts(d)
    rng = np.random.default_rng(4)
    g = rng.normal(size=50)
    R = 48
    lr, fd = [], []
    for seed in range(24):
        p0 = oracle.make_params(3, ARGS3, *PH)
        _, grad = oracle.rb_grad(od, p0, R, 100 + seed, g)
        lr.append(grad[3:])
        row = []
        for i in range(6):
            h = 0.05 * ARGS3[3 + i]
            v = []
            for sgn in (1, -1):
                a = np.array(ARGS3, np.float64)
                a[3 + i] += sgn * h
                M, _ = oracle.rb_grad(od, oracle.make_params(3, a.astype(np.float32), *PH), R,
                                      100 + seed, g)
                v.append(g @ M)
            row.append((v[0] - v[1]) / (2 * h))
        fd.append(row)
    lr, fd = np.array(lr), np.array(fd)
    diff = lr - fd
    z = diff.mean(axis=0) / (diff.std(axis=0, ddof=1) / np.sqrt(len(diff)))
    assert np.all(np.abs(z) < 4.0), (lr.mean(0), fd.mean(0), z)
    # the estimates carry signal: the mean gradient is resolved for most rates
    zs = lr.mean(axis=0) / (lr.std(axis=0, ddof=1) / np.sqrt(len(lr)))
    assert (np.abs(zs) > 2).sum() >= 2, zs
