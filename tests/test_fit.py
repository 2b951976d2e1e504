"""Fitting helpers (CPU): parameter transforms, state relabelling, GA operators."""
import numpy as np

from gsmfretda_amd import fit, synth


def test_theta_round_trip():
    rng = np.random.default_rng(0)
    for n in (2, 3, 4):
        E = np.sort(rng.uniform(0.05, 0.95, n))
        args = synth.make_args(E, synth.loguniform_rates(rng, n * (n - 1), 1e3, 1e5))
        th = fit.args_to_theta(args, n)
        back = fit.theta_to_args(th, n)
        assert np.allclose(back, np.asarray(args, np.float32), rtol=1e-6)
        assert np.all(back[n * n:] == 0)


def test_rate_index_matches_rate_matrix():
    n = 4
    args = synth.make_args(np.linspace(0.1, 0.9, n), np.arange(1, n * (n - 1) + 1, dtype=float))
    K = synth.rate_matrix(n, args)      # K[j, s] = rate s -> j
    for s in range(n):
        for j in range(n):
            if s != j:
                assert K[j, s] == args[n + fit.rate_index(n, s, j)]


def test_sort_states_is_a_relabelling():
    rng = np.random.default_rng(1)
    n = 4
    E = rng.uniform(0.05, 0.95, n)
    th = fit.args_to_theta(synth.make_args(E, synth.loguniform_rates(rng, 12, 1e3, 1e5)), n)
    out = fit.sort_states(th, n)
    order = np.argsort(th[:n])
    assert np.all(np.diff(out[:n]) >= 0)
    K0 = synth.rate_matrix(n, fit.theta_to_args(th, n))
    K1 = synth.rate_matrix(n, fit.theta_to_args(out, n))
    assert np.allclose(K1, K0[np.ix_(order, order)])
    assert np.allclose(fit.sort_states(out, n), out)


def test_ga_children_stay_in_bounds_and_ordered():
    rng = np.random.default_rng(2)
    n = 3
    ga = fit.GeneticFitter(batched=True, rng=rng, sigma0=3.0)
    pop = [fit.clip_theta(rng.normal(0, 5, n * n), n) for _ in range(16)]
    fitv = rng.random(16)
    for _ in range(200):
        c = ga._child(pop, fitv, n)
        assert np.all(c[:n] >= fit.LOGIT_E_BOUNDS[0]) and np.all(c[:n] <= fit.LOGIT_E_BOUNDS[1])
        assert np.all(c[n:] >= fit.LOG_K_BOUNDS[0]) and np.all(c[n:] <= fit.LOG_K_BOUNDS[1])
        assert np.all(np.diff(c[:n]) >= 0)


def test_history_time_to_target():
    h = fit.History("x")
    for t, c in ((0.0, 5.0), (1.0, 3.0), (2.0, 4.0), (3.0, 1.5)):
        h.record(t, c, 1, np.zeros(2))
    assert h.chi2 == [5.0, 3.0, 3.0, 1.5]
    assert h.time_to(3.0) == 1.0 and h.time_to(2.0) == 3.0 and h.time_to(1.0) is None
