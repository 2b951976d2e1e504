"""The synthetic workload generator follows the configured laws."""
import numpy as np
from scipy import stats

from gsmfretda_amd import synth


def test_c0_shape_and_laws():
    cfg = synth.config(0)
    d = cfg["data"]
    assert d.nbursts == 5000 and cfg["R"] == 20 and cfg["n"] == 2
    assert abs(d.phcount.mean() - 90) < 2.0                     # 50 + Geom(mean 40)
    assert d.phcount.min() >= 50
    dur_s = (d.stop - d.start + 1) * d.clk_p
    assert abs(np.median(dur_s) / 1e-3 - 1) < 0.05              # LogNormal(1 ms, 0.5)
    assert abs(np.std(np.log(dur_s)) - 0.5) < 0.03
    np.testing.assert_allclose(d.T, dur_s.astype(np.float32), rtol=1e-6)
    assert d.nphotons == int(d.phcount.sum())
    for b in range(0, 5000, 97):
        t = d.times[d.istart[b]:d.istart[b] + d.phcount[b]]
        assert np.all(np.diff(t) >= 0) and t[0] >= d.start[b] and t[-1] <= d.stop[b]
    assert np.all((d.mask_ad + d.mask_dd) == 1)
    # channel law: acceptor fraction = stationary mixture of the state efficiencies
    assert abs(d.mask_ad.mean() - 0.5) < 0.01


def test_trajectory_channel_law():
    # slow 2-state kinetics: a burst mostly stays in one state, so its
    # acceptor fraction is bimodal around E0 and E1
    args = synth.make_args([0.1, 0.9], [5.0, 5.0])
    d = synth.generate(2000, args, 2, seed=9)
    frac = np.array([d.mask_ad[d.istart[b]:d.istart[b] + d.phcount[b]].mean()
                     for b in range(d.nbursts)])
    lo, hi = (frac < 0.5).mean(), (frac >= 0.5).mean()
    assert 0.4 < lo < 0.6 and 0.4 < hi < 0.6
    assert abs(np.median(frac[frac < 0.5]) - 0.1) < 0.03
    assert abs(np.median(frac[frac >= 0.5]) - 0.9) < 0.03


def test_stationary_distribution():
    args = synth.make_args([0.2, 0.5, 0.8], [100, 200, 300, 400, 500, 600])
    K = synth.rate_matrix(3, args)
    pi = synth.stationary(K)
    Q = K - np.diag(K.sum(axis=0))
    assert np.allclose(Q @ pi, 0, atol=1e-9) and abs(pi.sum() - 1) < 1e-12


def test_geometric_photon_law():
    d = synth.generate(20000, synth.make_args([0.3, 0.7], [1e3, 1e3]), 2, base_photons=30,
                       mean_geom=80.0, dur_median=2e-3, dur_sigma=0.6, seed=4)
    x = d.phcount - 30                      # Geometric on {0,1,...}, p = 1/81
    k = np.arange(0, 400, 20)
    obs = np.histogram(x, bins=np.append(k, 10**6))[0]
    cdf = 1 - (80 / 81) ** np.append(k, 10**6)
    exp = np.diff(cdf) * len(x)
    assert stats.chisquare(obs, exp * obs.sum() / exp.sum()).pvalue > 1e-3
    assert abs(x.mean() - 80) < 2.5


def test_deterministic():
    a = synth.config(0, n_bursts=100)["data"]
    b = synth.config(0, n_bursts=100)["data"]
    assert np.array_equal(a.times, b.times) and np.array_equal(a.mask_ad, b.mask_ad)
