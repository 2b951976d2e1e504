"""Distributional parity with an independent reference-semantics MC.

tests/refmc_numpy.py simulates mc_kernel with numpy's random generator (no
shared code or random stream with the oracle or the kernels).  Two-sample
chi-square homogeneity tests on the E histograms must not reject at p = 0.01:
oracle vs numpy on the CPU, and the HIP path vs numpy on the GPU.  A planted
error (every state's efficiency 10 % higher in the numpy model) must be
rejected, so the test has power.
"""
import numpy as np
import pytest
from scipy import stats

from gsmfretda_amd import synth
from tests import refmc_numpy

CASES = {
    "2state_bg": dict(n=2, args=synth.make_args([0.30, 0.70], [1e3, 1e3]), bg=True,
                      ph=(1.0, 0.0, 0.05, 52.0)),
    "3state_broad": dict(n=3, args=synth.make_args([0.15, 0.5, 0.85],
                                                   [300, 2000, 800, 5000, 150, 9000],
                                                   v=[3.0, 0.0, 5.0]),
                         bg=True, ph=(0.34, 0.08, 0.07, 52.0)),
}


def _data(bg):
    d = synth.generate(500, synth.make_args([0.3, 0.7], [1e3, 1e3]), 2, seed=31)
    if not bg:
        d.bg_dd = d.bg_ad = 0.0
    return d


def homogeneity_p(h1, h2):
    keep = (h1 + h2) > 0
    table = np.vstack([h1[keep], h2[keep]])
    return stats.chi2_contingency(table)[1]


def _numpy_hist(d, case, R, seed, shift=1.0):
    args = np.array(case["args"], np.float64)
    n = case["n"]
    args[:n] = np.minimum(args[:n] * shift, 1.0)
    E = refmc_numpy.simulate(d, case["n"], args, R, np.random.default_rng(seed),
                             *case["ph"])
    return np.histogram(E[(E >= 0) & (E < 1)], bins=50, range=(0, 1))[0]


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_matches_independent_mc(oracle, name):
    case = CASES[name]
    d = _data(case["bg"])
    R = 200
    od = oracle.Data.from_bursts(d)
    p = oracle.make_params(case["n"], case["args"], *case["ph"])
    h_or = oracle.run(od, p, R=R, seed=7, want_units=False)["hist"]
    h_np = _numpy_hist(d, case, R, 8)
    assert homogeneity_p(h_or, h_np) > 0.01
    h_bad = _numpy_hist(d, case, R, 9, shift=1.1)
    assert homogeneity_p(h_or, h_bad) < 1e-4


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CASES))
def test_gpu_matches_independent_mc(gpu_mc, name):
    case = CASES[name]
    d = _data(case["bg"])
    R = 200
    gpu_mc.load(d, SgDivSr=np.zeros(d.nbursts, np.float32), nbins=50, left_pad=2)
    gpu_mc.set_photophysics(0, *case["ph"])
    assert gpu_mc.set_params(case["n"], 0, case["args"])
    gpu_mc.set_burst_range(0, -1, 0)
    gpu_mc.reSampleTimes = R
    gpu_mc.run_kernel(sid=0, seed=1234)
    gpu_mc.synchronize(0)
    h_gpu = gpu_mc.get_hist_chi2(0)[0].astype(np.int64)
    h_np = _numpy_hist(d, case, R, 10)
    assert homogeneity_p(h_gpu, h_np) > 0.01
