"""GPU: what the parameter-server counterpart sees equals the oracle.

The fitters of gsmfretda_amd.fit stand where the reference's external
parameter server stands: it sends parameter sets and reads back one chi2 per
set (src/streamWorker.cpp:197-223).  On a reduced fast-dynamics workload
(4 states, rates 1e3-1e5 s^-1, 1,500 bursts x 32 resamples) every chi2 the
per-set and the batched genetic algorithm receive for a sample of their
individuals, and the best chi2 Adam and Levenberg-Marquardt report, equals
the oracle's reduced chi2 (src/streamWorker.cpp:197-206) of the oracle's
histogram for that parameter set and the same random streams: the float
accumulation bit for bit, FP64 to 1e-12 relative.  The observed histogram the
problem fits is itself checked against the oracle's R = 1 run.
"""
import numpy as np
import pytest

from gsmfretda_amd import fit, fit_study, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small_problem(gpu_mc, oracle):
    cfg = synth.config(4, n_bursts=1500)
    prob = fit_study.make_problem(gpu_mc, cfg)
    od = oracle.Data.from_bursts(cfg["data"])
    n = cfg["n"]
    p = oracle.make_params(n, cfg["args"], *cfg["photophysics"])
    obs = oracle.run(od, p, R=1, seed=cfg["seeds"]["observed"])["E"]
    O = oracle.hist(obs).astype(np.float32)
    assert np.array_equal(O.astype(np.float64), prob.observed)
    pop0 = [fit.args_to_theta(a, n) for _, a in cfg["population"][:8]]
    return prob, cfg, od, O, pop0


def _oracle_chi2(oracle, cfg, od, O, prob, theta):
    n = prob.n
    p = oracle.make_params(n, fit.theta_to_args(theta, n), *cfg["photophysics"])
    h = oracle.run(od, p, R=prob.R, seed=prob.seed, set_id=0, want_units=False)["hist"]
    c32, c64, _ = oracle.chi2(O, h, prob.R, n)
    return c32, c64


def _check(oracle, small_problem, thetas, chi2s):
    prob, cfg, od, O, _ = small_problem
    for th, c in zip(thetas, chi2s):
        _, c64 = _oracle_chi2(oracle, cfg, od, O, prob, th)
        assert abs(c - c64) <= 1e-12 * abs(c64), (c, c64)


@pytest.mark.parametrize("batched", [False, True])
def test_genetic_algorithm_chi2_equals_oracle(oracle, small_problem, batched):
    prob, cfg, od, O, pop0 = small_problem
    seen = []
    ga = fit.GeneticFitter(batched=batched, pop_size=8, rng=np.random.default_rng(3))
    evaluate = ga._evaluate

    def recording(prob_, pop):
        f = evaluate(prob_, pop)
        seen.append((list(pop), np.array(f)))
        return f
    ga._evaluate = recording
    h = ga.run(prob, pop0, max_seconds=60.0, max_generations=2)
    assert len(seen) == 3 and all(len(p) == 8 for p, _ in seen)
    rng = np.random.default_rng(4)
    for pop, f in (seen[0], seen[-1]):
        pick = rng.choice(8, 3, replace=False)
        _check(oracle, small_problem, [pop[i] for i in pick], [f[i] for i in pick])
    # the reported best is the best individual's chi2
    assert h.chi2[-1] == min(float(f.min()) for _, f in seen)
    # per-set evaluation: the chi2 of the float accumulation too
    if not batched:
        th = seen[0][0][0]
        c32, _ = _oracle_chi2(oracle, cfg, od, O, prob, th)
        prob.mc.reSampleTimes = prob.R
        assert prob.mc.set_params(prob.n, 0, fit.theta_to_args(th, prob.n))
        prob.mc.run_kernel(sid=0, seed=prob.seed, set_id=0)
        prob.mc.synchronize(0)
        _, _, g32, _ = prob.mc.get_hist_chi2(0)
        assert np.float32(g32).view(np.uint32) == np.float32(c32).view(np.uint32)


@pytest.mark.parametrize("name", ["adam-rb", "lm-rb"])
def test_gradient_fitters_report_the_oracle_chi2(oracle, small_problem, name):
    prob, cfg, od, O, pop0 = small_problem
    fitter = (fit.AdamFitter(lr=0.1, R_grad=8, check_every=2) if name == "adam-rb"
              else fit.LMFitter(R_grad=8))
    h, _ = fit.fit_from_population(fitter, prob, pop0, 4.0, rank_resamples=4)
    assert h.best_theta is not None and len(h.chi2) >= 2
    _check(oracle, small_problem, [h.best_theta], [h.chi2[-1]])
