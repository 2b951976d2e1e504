"""Fitting loops over the MI355X PDA path: the parameter server's side.

The reference's optimiser is pySMFRETda, an external parameter server that
hands a genetic algorithm's individuals to the GPU workers one at a time and
reads their chi2 back (src/streamWorker.cpp:83-182; protocol in
3rdparty/protobuf/args.proto).  This module provides that loop against the
C ABI in three forms:

* `GeneticFitter(batched=False)` -- one evaluation per individual through
  pda_run (the reference's one-set-per-call pattern);
* `GeneticFitter(batched=True)` -- a generation per pda_run_population call
  (one graph replay);
* `AdamFitter` -- Adam on (logit E, log K) with the gradient of the
  Rao-Blackwellised chi2 from the adjoint kernel (gsmfretda_amd.rb_torch);
* `LMFitter` -- Levenberg-Marquardt on the same chi2 written as a sum of
  squared Pearson residuals, with the Jacobian of the expected histogram
  from one pass of the Jacobian kernel (pda_rb_jacobian).

Every fitter reports progress in the same currency, the plain simulator's
chi2 at the current best parameters with fixed random streams (seed, set id
0), so the wall time each needs to reach a chi2 target can be compared.
"""
import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import pda

LOGIT_E_BOUNDS = (-6.0, 6.0)
LOG_K_BOUNDS = (math.log(1e1), math.log(1e7))


def theta_to_args(theta, n):
    """(logit E (n), log K (n (n - 1))) -> mc::set_params args with v = 0."""
    theta = np.asarray(theta, np.float64)
    E = 1.0 / (1.0 + np.exp(-theta[:n]))
    K = np.exp(theta[n:])
    return np.concatenate([E, K, np.zeros(n)]).astype(np.float32)


def args_to_theta(args, n):
    args = np.asarray(args, np.float64)
    E = np.clip(args[:n], 1e-6, 1 - 1e-6)
    return np.concatenate([np.log(E / (1 - E)), np.log(args[n:n * n])])


def rate_index(n, s, j):
    """Position in the off-diagonal rate list of K(j, s), the rate s -> j
    (genMatK's row-major order, src/eigenhelper.cpp:4-21)."""
    return j * (n - 1) + (s if s < j else s - 1)


def sort_states(theta, n):
    """Relabels the states by increasing E (the model is symmetric under a
    permutation of its states), carrying the rates along."""
    theta = np.asarray(theta, np.float64)
    order = np.argsort(theta[:n], kind="stable")
    if np.all(order == np.arange(n)):
        return theta.copy()
    inv = np.empty(n, np.int64)
    inv[order] = np.arange(n)
    out = theta.copy()
    out[:n] = theta[:n][order]
    for s in range(n):
        for j in range(n):
            if s != j:
                out[n + rate_index(n, inv[s], inv[j])] = theta[n + rate_index(n, s, j)]
    return out


def clip_theta(theta, n):
    theta = np.asarray(theta, np.float64).copy()
    theta[:n] = np.clip(theta[:n], *LOGIT_E_BOUNDS)
    theta[n:] = np.clip(theta[n:], *LOG_K_BOUNDS)
    return theta


@dataclass
class Problem:
    """A loaded data set, its observed histogram and the fixed random streams."""
    mc: pda.MC
    n: int
    R: int
    seed: int
    observed: np.ndarray          # observed histogram (float64, B bins)
    evaluations: int = 0

    def chi2(self, theta):
        """Plain-simulator chi2 of one parameter set (pda_run, set id 0)."""
        self.evaluations += 1
        _, c64, _, _ = pda.evaluate(self.mc, self.n, theta_to_args(theta, self.n), self.R,
                                    self.seed, set_id=0)
        return c64

    def chi2_sets(self, thetas):
        """Plain-simulator chi2 of each set, one pda_run per set (the
        reference's per-set evaluation) with the context's stream slots in
        flight in turn, as the reference's worker keeps them
        (src/gpuWorker.cpp:38-88); set id 0 everywhere, the streams of
        `chi2`.  Every slot needs the photophysics and burst range set."""
        m, k = self.mc, max(1, self.mc.streamNum)
        self.evaluations += len(thetas)
        out = [0.0] * len(thetas)
        m.reSampleTimes = self.R
        for i, t in enumerate(thetas):
            sid = i % k
            if i >= k:
                m.synchronize(sid)
                out[i - k] = m.get_hist_chi2(sid)[1]
            if not m.set_params(self.n, sid, theta_to_args(t, self.n)):
                raise pda.PdaError("bad parameter set")
            m.run_kernel(sid=sid, seed=self.seed, set_id=0)
        for i in range(max(0, len(thetas) - k), len(thetas)):
            m.synchronize(i % k)
            out[i] = m.get_hist_chi2(i % k)[1]
        return out

    def chi2_population(self, thetas, R=None):
        """Plain-simulator chi2 of a generation in one population run; every
        set uses set id 0, the streams of `chi2`.  R: resamples per burst
        (default the problem's)."""
        self.evaluations += len(thetas)
        self.mc.set_population([self.n] * len(thetas),
                               [theta_to_args(t, self.n) for t in thetas])
        self.mc.run_population(R=self.R if R is None else R, seed=self.seed, set_base=0,
                               set_stride=0)
        _, c64, _, _ = self.mc.get_population_result()
        return c64


def fit_from_population(fitter, prob, pop0, seconds, target=None, rank_resamples=4,
                        restarts=False):
    """Runs `fitter` from the best of the starting sets `pop0`.

    The sets are ranked by one batched plain evaluation with `rank_resamples`
    resamples per burst: random starting sets differ in chi2 by orders of
    magnitude, which a few resamples of every burst resolve, so the ranking
    costs a fraction of an evaluation at the problem's R.  Its wall time is
    part of the returned history (every chi2 in which is the plain chi2 at
    the problem's R).  restarts: hand the next best sets to the fitter
    (LMFitter).  Returns (History, seconds spent ranking)."""
    t0 = time.perf_counter()
    fit0 = prob.chi2_population(pop0, R=rank_resamples)
    ranked = [pop0[i] for i in np.argsort(fit0, kind="stable")]
    t_init = time.perf_counter() - t0
    extra = {"restarts": ranked[1:]} if restarts else {}
    h = fitter.run(prob, ranked[0], seconds - t_init, target=target, **extra)
    h.times = [t + t_init for t in h.times]
    return h, t_init


@dataclass
class History:
    method: str
    times: list = field(default_factory=list)     # wall seconds since start
    chi2: list = field(default_factory=list)      # best plain chi2 so far
    evals: list = field(default_factory=list)
    best_theta: np.ndarray = None

    def record(self, t, c, e, theta):
        if not self.chi2 or c < self.chi2[-1]:
            self.best_theta = np.array(theta)
        best = c if not self.chi2 else min(c, self.chi2[-1])
        self.times.append(t)
        self.chi2.append(best)
        self.evals.append(e)

    def time_to(self, target):
        for t, c in zip(self.times, self.chi2):
            if c <= target:
                return t
        return None

    def as_dict(self):
        return dict(method=self.method, times=self.times, chi2=self.chi2, evals=self.evals,
                    best_theta=None if self.best_theta is None else self.best_theta.tolist())


class GeneticFitter:
    """Real-coded genetic algorithm: tournament selection, blend crossover in
    (logit E, log K), Gaussian mutation with a decaying scale, elitism; the
    states of every individual kept ordered by E."""

    def __init__(self, batched, pop_size=64, elite=2, tournament=3, blend=0.3, sigma0=0.3,
                 sigma_decay=0.97, sigma_min=0.01, rng=None):
        self.batched = batched
        self.pop_size = pop_size
        self.elite = elite
        self.tournament = tournament
        self.blend = blend
        self.sigma = sigma0
        self.sigma_decay = sigma_decay
        self.sigma_min = sigma_min
        self.rng = rng if rng is not None else np.random.default_rng(0)

    def _evaluate(self, prob, pop):
        if self.batched:
            return np.asarray(prob.chi2_population(pop))
        return np.asarray(prob.chi2_sets(pop))

    def _child(self, pop, fit, n):
        rng = self.rng

        def pick():
            idx = rng.integers(0, len(pop), self.tournament)
            return pop[idx[np.argmin(fit[idx])]]
        a, b = pick(), pick()
        lo, hi = np.minimum(a, b), np.maximum(a, b)
        span = hi - lo
        child = rng.uniform(lo - self.blend * span, hi + self.blend * span)
        child = child + rng.normal(0.0, self.sigma, size=child.shape) * (rng.random(child.shape) < 0.5)
        return sort_states(clip_theta(child, n), n)

    def run(self, prob, pop0, max_seconds, target=None, max_generations=100000):
        n = prob.n
        hist = History("ga-batched" if self.batched else "ga-per-set")
        t0 = time.perf_counter()
        pop = [sort_states(clip_theta(t, n), n) for t in pop0]
        fit = self._evaluate(prob, pop)
        i = int(np.argmin(fit))
        hist.record(time.perf_counter() - t0, float(fit[i]), prob.evaluations, pop[i])
        for _ in range(max_generations):
            if target is not None and hist.chi2[-1] <= target:
                break
            if time.perf_counter() - t0 > max_seconds:
                break
            order = np.argsort(fit)
            new = [pop[k] for k in order[:self.elite]]
            while len(new) < self.pop_size:
                new.append(self._child(pop, fit, n))
            pop = new
            fit = self._evaluate(prob, pop)
            i = int(np.argmin(fit))
            hist.record(time.perf_counter() - t0, float(fit[i]), prob.evaluations, pop[i])
            self.sigma = max(self.sigma_min, self.sigma * self.sigma_decay)
        return hist


class AdamFitter:
    """Adam on (logit E, log K) with the adjoint gradient of the
    Rao-Blackwellised chi2; progress measured by the plain chi2 every
    `check_every` steps."""

    def __init__(self, lr=0.05, betas=(0.9, 0.999), check_every=1, lr_decay=1.0, R_grad=None,
                 plateau=0, plateau_factor=0.5, lr_min=1e-3):
        self.lr = lr
        self.betas = betas
        self.check_every = check_every
        self.lr_decay = lr_decay      # per-step multiplicative decay
        self.R_grad = R_grad  # resamples of the RB gradient (default: the problem's R)
        # reduce on plateau: after `plateau` steps without a 0.1 % improvement
        # of the RB chi2, multiply the learning rate by plateau_factor
        self.plateau = plateau
        self.plateau_factor = plateau_factor
        self.lr_min = lr_min

    def run(self, prob, theta0, max_seconds, target=None, max_steps=100000):
        import torch

        from . import rb_torch
        n = prob.n
        hist = History("adam-rb")
        t0 = time.perf_counter()
        theta = torch.tensor(clip_theta(theta0, n), dtype=torch.float64, requires_grad=True)
        opt = torch.optim.Adam([theta], lr=self.lr, betas=self.betas)
        sched = torch.optim.lr_scheduler.ExponentialLR(opt, self.lr_decay)
        obs = torch.tensor(prob.observed, dtype=torch.float64)
        c = prob.chi2(theta.detach().numpy())
        hist.record(time.perf_counter() - t0, c, prob.evaluations, theta.detach().numpy())
        best_loss, since = float("inf"), 0
        for step in range(max_steps):
            if target is not None and hist.chi2[-1] <= target:
                break
            if time.perf_counter() - t0 > max_seconds:
                break
            opt.zero_grad()
            E = torch.sigmoid(theta[:n])
            K = torch.exp(theta[n:])
            args = torch.cat([E, K, torch.zeros(n, dtype=torch.float64)])
            loss = rb_torch.rb_chi2(args, prob.mc, n, self.R_grad or prob.R, prob.seed, obs,
                                    set_id=0)
            loss.backward()
            prob.evaluations += 1
            opt.step()
            sched.step()
            if self.plateau:
                lv = float(loss.detach())
                if lv < best_loss * 0.999:
                    best_loss, since = lv, 0
                else:
                    since += 1
                    if since >= self.plateau:
                        for grp in opt.param_groups:
                            grp["lr"] = max(self.lr_min, grp["lr"] * self.plateau_factor)
                        since = 0
            with torch.no_grad():
                theta[:n].clamp_(*LOGIT_E_BOUNDS)
                theta[n:].clamp_(*LOG_K_BOUNDS)
            if (step + 1) % self.check_every == 0:
                th = theta.detach().numpy()
                hist.record(time.perf_counter() - t0, prob.chi2(th), prob.evaluations, th)
        return hist


class LMFitter:
    """Levenberg-Marquardt on (logit E, log K) for the Rao-Blackwellised chi2.

    With m_b = M_b / R the expected counts of the kept bins (b >= left_pad,
    M_b > 0), the chi2 is sum_b r_b^2 / dof with r_b = (O_b - m_b) / sqrt(m_b);
    dr_b / dM_b = -(O_b + m_b) / (2 R m_b^1.5) and dM / dargs comes from
    pda_rb_jacobian on the forward run at the current point, so one
    accepted step costs one forward (with scores) and one Jacobian pass; a
    rejected step costs one forward.  Marquardt damping: solve
    (A + lam D) delta = -J^T r with A = J^T J and D = diag(A) floored at
    1e-3 max diag(A) (rates the histogram barely sees would otherwise take
    huge steps), lam / 3 after an accepted step, lam x 4 after a rejected
    one; a step is also capped at `step_max` in every coordinate (logit E,
    log K), doubled after an accepted step and halved after a rejected one
    (within [step_min, 2]).  The RB chi2 with common random numbers is a
    step function of the rates at fine scales, so the fit stops when the cap
    falls below step_min.  Progress is reported as the plain chi2, like the
    other fitters."""

    def __init__(self, R_grad=None, lam0=1e-2, lam_max=1e8, step_max=0.5, step_min=1e-4,
                 check_every=1, left_pad=2):
        self.R_grad = R_grad
        self.lam0 = lam0
        self.lam_max = lam_max
        self.step_max = step_max
        self.step_min = step_min
        self.check_every = check_every
        self.left_pad = left_pad

    def _forward(self, prob, theta, R):
        """RB expected histogram (summed over R resamples) and its reduced chi2."""
        args = theta_to_args(theta, prob.n)
        if not prob.mc.set_params(prob.n, 0, args):
            raise ValueError("set_params rejected the parameters")
        prob.mc.rb_forward(R=R, seed=prob.seed, set_id=0, scores=True)
        M, _, c2, _ = prob.mc.rb_result()
        return M, c2

    def _system(self, prob, theta, M, R):
        """Normal equations of the residuals at the forward just run."""
        n = prob.n
        J = prob.mc.rb_jacobian(n, len(M))[:, :n * n]       # d M_b / d (E, K)
        args = theta_to_args(theta, n).astype(np.float64)
        E = args[:n]
        dargs = np.concatenate([E * (1.0 - E), args[n:n * n]])   # d args / d theta
        O = prob.observed
        m = M / R
        keep = (np.arange(len(M)) >= self.left_pad) & (M > 0)
        r = np.zeros(len(M))
        r[keep] = (O[keep] - m[keep]) / np.sqrt(m[keep])
        dr = np.zeros(len(M))
        dr[keep] = -(O[keep] + m[keep]) / (2.0 * R * m[keep] ** 1.5)
        Jr = (dr[:, None] * J) * dargs[None, :]
        return Jr.T @ Jr, Jr.T @ r

    def run(self, prob, theta0, max_seconds, target=None, max_steps=100000, restarts=()):
        """Descends from theta0; when a descent converges (the step cap
        falls below step_min) short of the target, starts again from the next
        point of `restarts` (e.g. the next best sets of an initial
        population) while time is left.  The history keeps the best plain
        chi2 over all descents."""
        n = prob.n
        R = self.R_grad or prob.R
        hist = History("lm-rb")
        t0 = time.perf_counter()
        for start in [theta0, *restarts]:
            self._descend(prob, clip_theta(start, n), R, hist, t0, max_seconds, target, max_steps)
            if target is not None and hist.chi2 and hist.chi2[-1] <= target:
                break
            if time.perf_counter() - t0 > max_seconds:
                break
        return hist

    def _descend(self, prob, theta, R, hist, t0, max_seconds, target, max_steps):
        n = prob.n
        hist.record(time.perf_counter() - t0, prob.chi2(theta), prob.evaluations, theta)
        M, c2 = self._forward(prob, theta, R)
        prob.evaluations += 1
        A, g = self._system(prob, theta, M, R)
        lam, cap = self.lam0, self.step_max
        last_recorded = theta
        for step in range(max_steps):
            if target is not None and hist.chi2[-1] <= target:
                break
            if time.perf_counter() - t0 > max_seconds or lam > self.lam_max or cap < self.step_min:
                break
            dA = np.diag(A)
            D = np.diag(np.maximum(dA, 1e-3 * max(dA.max(), 1e-300)))
            try:
                delta = np.linalg.solve(A + lam * D, -g)
            except np.linalg.LinAlgError:
                lam *= 4.0
                continue
            big = np.abs(delta).max()
            if big > cap:
                delta *= cap / big
            cand = clip_theta(theta + delta, n)
            Mc, cc = self._forward(prob, cand, R)
            prob.evaluations += 1
            if np.isfinite(cc) and cc < c2:
                theta, M, c2 = cand, Mc, cc
                A, g = self._system(prob, theta, M, R)
                lam = max(lam / 3.0, 1e-12)
                cap = min(2.0 * cap, 2.0)
                if (step + 1) % self.check_every == 0:
                    hist.record(time.perf_counter() - t0, prob.chi2(theta), prob.evaluations, theta)
                    last_recorded = theta
            else:
                lam *= 4.0
                cap *= 0.5
        if last_recorded is not theta:
            hist.record(time.perf_counter() - t0, prob.chi2(theta), prob.evaluations, theta)


def rb_cotangent(M, observed, R, n, left_pad=2):
    """d chi2 / d M_b of the reduced chi2 of the expected counts M (summed
    over R resamples): (1 - O_b^2 / m_b^2) / (R dof) on the kept bins
    (b >= left_pad, M_b > 0), m_b = M_b / R."""
    m = M / R
    keep = (np.arange(len(M)) >= left_pad) & (M > 0)
    dof = keep.sum() - n * (n + 1)
    g = np.zeros(len(M))
    g[keep] = (1.0 - observed[keep] ** 2 / m[keep] ** 2) / (R * dof)
    return g


class BatchedLMFitter(LMFitter):
    """Levenberg-Marquardt chains in lockstep, several starting points at
    once: every iteration evaluates the candidate step of every chain in one
    RB population forward (pda_rb_forward_population) and the Jacobians of
    the chains whose step was accepted in one population Jacobian pass
    (pda_rb_jacobian_population).  A chain whose descent converges short of
    the target restarts from the next point of `restarts`.  The history
    keeps the best plain chi2 of all chains (one batched plain evaluation of
    the improved chains per iteration)."""

    def __init__(self, chains=4, **kw):
        super().__init__(**kw)
        self.chains = chains

    def _forward_population(self, prob, thetas, R):
        n = prob.n
        prob.mc.set_population([n] * len(thetas), [theta_to_args(t, n) for t in thetas])
        prob.mc.rb_forward_population(R=R, seed=prob.seed, set_base=0, set_stride=0)
        M, _, c2, _ = prob.mc.rb_population_result()
        return M, c2

    def _system_from(self, prob, theta, M, J, R):
        n = prob.n
        J = J[:, :n * n]
        args = theta_to_args(theta, n).astype(np.float64)
        E = args[:n]
        dargs = np.concatenate([E * (1.0 - E), args[n:n * n]])
        O = prob.observed
        m = M / R
        keep = (np.arange(len(M)) >= self.left_pad) & (M > 0)
        r = np.zeros(len(M))
        r[keep] = (O[keep] - m[keep]) / np.sqrt(m[keep])
        dr = np.zeros(len(M))
        dr[keep] = -(O[keep] + m[keep]) / (2.0 * R * m[keep] ** 1.5)
        Jr = (dr[:, None] * J) * dargs[None, :]
        return Jr.T @ Jr, Jr.T @ r

    def run(self, prob, theta0, max_seconds, target=None, max_steps=100000, restarts=()):
        n = prob.n
        R = self.R_grad or prob.R
        hist = History("lm-rb-batched")
        t0 = time.perf_counter()
        pending = [clip_theta(t, n) for t in [theta0, *restarts]]
        K = min(self.chains, len(pending))
        theta = [pending.pop(0) for _ in range(K)]
        fresh = [True] * K                 # evaluated at theta, not yet a step
        c2 = [np.inf] * K
        M = [None] * K
        A = [None] * K
        g = [None] * K
        lam = [self.lam0] * K
        cap = [self.step_max] * K
        alive = [True] * K
        c0 = prob.chi2_population(theta)
        for k in range(K):
            hist.record(time.perf_counter() - t0, float(c0[k]), prob.evaluations, theta[k])
        for _ in range(max_steps):
            if target is not None and hist.chi2[-1] <= target:
                break
            if time.perf_counter() - t0 > max_seconds or not any(alive):
                break
            cand = []
            for k in range(K):
                if fresh[k] or not alive[k]:
                    cand.append(theta[k])
                    continue
                dA = np.diag(A[k])
                D = np.diag(np.maximum(dA, 1e-3 * max(dA.max(), 1e-300)))
                try:
                    delta = np.linalg.solve(A[k] + lam[k] * D, -g[k])
                except np.linalg.LinAlgError:
                    lam[k] *= 4.0
                    cand.append(theta[k])
                    continue
                big = np.abs(delta).max()
                if big > cap[k]:
                    delta *= cap[k] / big
                cand.append(clip_theta(theta[k] + delta, n))
            Mc, cc = self._forward_population(prob, cand, R)
            prob.evaluations += K
            accepted = np.zeros(K, np.int32)
            for k in range(K):
                if not alive[k]:
                    continue
                if fresh[k] or (np.isfinite(cc[k]) and cc[k] < c2[k]):
                    accepted[k] = 1
                    if not fresh[k]:
                        lam[k] = max(lam[k] / 3.0, 1e-12)
                        cap[k] = min(2.0 * cap[k], 2.0)
                    theta[k], M[k], c2[k] = cand[k], Mc[k], cc[k]
                    fresh[k] = False
                else:
                    lam[k] *= 4.0
                    cap[k] *= 0.5
                    if lam[k] > self.lam_max or cap[k] < self.step_min:
                        # converged short of the target: the next start
                        if pending:
                            theta[k] = pending.pop(0)
                            fresh[k], c2[k] = True, np.inf
                            lam[k], cap[k] = self.lam0, self.step_max
                        else:
                            alive[k] = False
            if accepted.any():
                J = prob.mc.rb_jacobian_population(n, len(Mc[0]), which=accepted)
                for k in np.flatnonzero(accepted):
                    A[k], g[k] = self._system_from(prob, theta[k], M[k], J[k], R)
                improved = [theta[k] for k in np.flatnonzero(accepted)]
                cp = prob.chi2_population(improved)
                for th, c in zip(improved, cp):
                    hist.record(time.perf_counter() - t0, float(c), prob.evaluations, th)
        return hist


class BatchedAdamFitter:
    """Adam chains in lockstep from several starting points: every step one
    RB population forward with scores and one population adjoint
    (pda_rb_adjoint_population) give the gradients of every chain; Adam on
    (logit E, log K) per chain; every `check_every` steps one batched plain
    evaluation of all chains records the best plain chi2."""

    def __init__(self, chains=4, lr=0.1, betas=(0.9, 0.999), eps=1e-8, check_every=5,
                 R_grad=None, left_pad=2):
        self.chains = chains
        self.lr = lr
        self.betas = betas
        self.eps = eps
        self.check_every = check_every
        self.R_grad = R_grad
        self.left_pad = left_pad

    def run(self, prob, theta0, max_seconds, target=None, max_steps=100000, restarts=()):
        n = prob.n
        R = self.R_grad or prob.R
        hist = History("adam-rb-batched")
        t0 = time.perf_counter()
        starts = [clip_theta(t, n) for t in [theta0, *restarts]]
        K = min(self.chains, len(starts))
        theta = np.array(starts[:K])
        m = np.zeros_like(theta)
        v = np.zeros_like(theta)
        b1, b2 = self.betas
        c0 = prob.chi2_population(list(theta))
        for k in range(K):
            hist.record(time.perf_counter() - t0, float(c0[k]), prob.evaluations, theta[k])
        for step in range(max_steps):
            if target is not None and hist.chi2[-1] <= target:
                break
            if time.perf_counter() - t0 > max_seconds:
                break
            prob.mc.set_population([n] * K, [theta_to_args(t, n) for t in theta])
            prob.mc.rb_forward_population(R=R, seed=prob.seed, set_base=0, set_stride=0)
            M, _, _, _ = prob.mc.rb_population_result()
            gM = np.array([rb_cotangent(M[k], prob.observed, R, n, self.left_pad)
                           for k in range(K)])
            ga = prob.mc.rb_adjoint_population(gM, n)
            prob.evaluations += K
            args = np.array([theta_to_args(t, n) for t in theta], np.float64)
            E = args[:, :n]
            grad = np.concatenate([ga[:, :n] * E * (1.0 - E), ga[:, n:n * n] * args[:, n:n * n]],
                                  axis=1)
            m = b1 * m + (1 - b1) * grad
            v = b2 * v + (1 - b2) * grad ** 2
            mh = m / (1 - b1 ** (step + 1))
            vh = v / (1 - b2 ** (step + 1))
            theta = np.array([clip_theta(t, n) for t in theta - self.lr * mh / (np.sqrt(vh) + self.eps)])
            if (step + 1) % self.check_every == 0:
                cp = prob.chi2_population(list(theta))
                for k in range(K):
                    hist.record(time.perf_counter() - t0, float(cp[k]), prob.evaluations, theta[k])
        return hist
