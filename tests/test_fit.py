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


class _FakeRB:
    """A smooth stand-in for the RB forward and its Jacobian (CPU): expected
    counts R N sum_s pi_s phi(x_b; E_s, 0.08) with pi the stationary
    distribution of the rate matrix, Jacobian by central differences."""

    def __init__(self, n, observed, B=50, N=5000.0, left_pad=2):
        self.n, self.B, self.N, self.left_pad = n, B, N, left_pad
        self.observed = observed
        self.x = (np.arange(B) + 0.5) / B

    def expected(self, args):
        n = self.n
        E = np.asarray(args[:n], np.float64)
        K = synth.rate_matrix(n, np.asarray(args, np.float64))
        Q = K - np.diag(K.sum(0))
        w, v = np.linalg.eig(Q)
        pi = np.real(v[:, np.argmin(np.abs(w))])
        pi = pi / pi.sum()
        h = sum(pi[s] * np.exp(-0.5 * ((self.x - E[s]) / 0.08) ** 2) for s in range(n))
        return self.N * h / h.sum()

    def set_params(self, n, sid, args):
        self.args = np.asarray(args, np.float64)
        return True

    def rb_forward(self, R, seed, set_id=0, scores=True):
        self.R = R
        self.M = R * self.expected(self.args)

    def rb_result(self):
        m = self.M / self.R
        keep = (np.arange(self.B) >= self.left_pad) & (self.M > 0)
        c2 = np.sum((self.observed[keep] - m[keep]) ** 2 / m[keep]) / (keep.sum() - self.n * (self.n + 1))
        return self.M, None, c2, int(keep.sum())

    def rb_jacobian(self, n, nbins):
        J = np.zeros((nbins, n * (n + 1)))
        for i in range(n * n):
            h = 1e-6 * max(abs(self.args[i]), 1e-3)
            ap, am = self.args.copy(), self.args.copy()
            ap[i] += h
            am[i] -= h
            J[:, i] = self.R * (self.expected(ap) - self.expected(am)) / (2 * h)
        return J


class _FakeProblem(fit.Problem):
    def chi2(self, theta):
        self.evaluations += 1
        self.mc.set_params(self.n, 0, fit.theta_to_args(theta, self.n))
        self.mc.rb_forward(self.R, self.seed)
        return self.mc.rb_result()[2]


def _fake(n=3):
    truth = synth.make_args([0.2, 0.5, 0.8][:n], [800, 2000, 600, 2500, 1200, 4000][:n * (n - 1)])
    mc = _FakeRB(n, None)
    mc.observed = mc.expected(np.asarray(truth, np.float64))
    return _FakeProblem(mc, n, 4, 1, mc.observed), fit.args_to_theta(truth, n)


def test_lm_normal_equations_are_the_chi2_gradient():
    prob, th_true = _fake()
    th = th_true + np.random.default_rng(3).normal(0, 0.2, th_true.shape)
    lm = fit.LMFitter(R_grad=4)
    M, _ = lm._forward(prob, th, 4)
    A, g = lm._system(prob, th, M, 4)
    dof = 48 - 12

    def sumsq(t):
        return lm._forward(prob, t, 4)[1] * dof
    # the parameters pass through float32 (mc::set_params): steps of 1e-3
    for i in range(len(th)):
        e = np.zeros_like(th)
        e[i] = 1e-3
        fd = (sumsq(th + e) - sumsq(th - e)) / 2e-3
        assert abs(2 * g[i] - fd) <= 1e-3 * max(abs(fd), 1e-3 * np.abs(g).max()), (i, 2 * g[i], fd)
    assert np.all(np.linalg.eigvalsh(A) >= -1e-9 * np.abs(A).max())


def test_lm_converges_on_a_smooth_problem():
    prob, th_true = _fake()
    start = th_true + np.array([0.3, -0.2, 0.25, 0.4, -0.3, 0.2, -0.4, 0.3, 0.1])
    h = fit.LMFitter(R_grad=4).run(prob, start, max_seconds=30, max_steps=200)
    assert h.chi2[-1] < 1e-6, h.chi2[-5:]
    assert h.chi2[0] > 1.0


def test_lm_restarts_after_a_converged_descent():
    prob, th_true = _fake()
    starts = [th_true + 0.3, th_true - 0.2, th_true + np.linspace(-0.2, 0.2, len(th_true))]
    # an unreachable target: every start is descended in turn
    h = fit.LMFitter(R_grad=4).run(prob, starts[0], max_seconds=60, target=-1.0, max_steps=3,
                                   restarts=starts[1:])
    assert len(h.chi2) >= 3 * 2          # a start record and at least one step per descent
    # a reachable one: the first descent gets there, no restart is started
    prob2, _ = _fake()
    h2 = fit.LMFitter(R_grad=4).run(prob2, starts[0], max_seconds=60, target=1e-3,
                                    restarts=starts[1:])
    assert h2.chi2[-1] <= 1e-3


def test_fit_from_population_ranks_then_runs_from_the_best():
    class Prob:
        evaluations = 0

        def __init__(self):
            self.calls = []

        def chi2_population(self, thetas, R=None):
            self.calls.append(R)
            return [float(np.sum(t ** 2)) for t in thetas]

    class Fitter:
        def run(self, prob, theta0, max_seconds, target=None, restarts=()):
            self.got = (theta0, list(restarts), max_seconds, target)
            h = fit.History("stub")
            h.record(0.5, 1.0, 1, theta0)
            return h

    pop = [np.full(3, v) for v in (3.0, -1.0, 2.0, 0.5)]
    prob, f = Prob(), Fitter()
    h, t = fit.fit_from_population(f, prob, pop, 10.0, target=2.0, rank_resamples=4,
                                   restarts=True)
    assert prob.calls == [4]
    assert np.array_equal(f.got[0], pop[3])
    assert [r[0] for r in f.got[1]] == [-1.0, 2.0, 3.0]
    assert f.got[2] == 10.0 - t and f.got[3] == 2.0 and t >= 0.0
    assert h.times == [0.5 + t]
    fit.fit_from_population(f, prob, pop, 10.0)
    assert f.got[1] == [] and prob.calls == [4, 4]
