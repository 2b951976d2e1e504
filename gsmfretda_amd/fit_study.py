"""Fit wall-time study on one GPU: gradient + batch against plain evaluation.

Workload (fast dynamics): 4 states, true rates log-uniform in [1e3, 1e5]
s^-1, 1,000,000 synthetic bursts of 30 + Geom(mean 80) photons and
LogNormal(2 ms, 0.6) durations, 32 resamples per burst, 50 bins.  The
observed histogram is the simulator at the true parameters with R = 1 and a
held-out seed; chi2(truth) is the plain simulator's chi2 at the true
parameters with the evaluation streams.  From the same initial population of
64 random parameter sets (states from the fitted model's range), four
fitters run under the same wall-clock limit:

  ga-per-set   genetic algorithm, one pda_run per individual (the reference's
               one-set-per-call evaluation), three stream slots in flight as
               the reference's worker keeps several
  ga-batched   the same algorithm, one pda_run_population per generation
  adam-rb      Adam on (logit E, log K) with the adjoint gradient of the
               Rao-Blackwellised chi2 (8 resamples per burst), started from
               the best initial set, found by one batched evaluation with 4
               resamples per burst (timed)
  lm-rb        Levenberg-Marquardt on the same chi2 with the Jacobian of the
               expected histogram (pda_rb_jacobian), same start
  lm-rb-batched, adam-rb-batched
               the same with four chains in lockstep through the RB
               population passes (pda_rb_forward_population and its adjoint
               and Jacobian), started from the four best initial sets

Each is timed to reach chi2 <= 1.01 chi2(truth), and to reach within 1 % of
the chi2 the per-set genetic algorithm ends with.

    python -m gsmfretda_amd.fit_study [--bursts N] [--seconds S] [--fitters a,b] [--out FILE]
"""
import argparse
import json
import sys
import time

import numpy as np

from . import fit, pda, synth


def make_problem(m, cfg, nbins=50):
    """Loads a synth.config data set into the MC context m and returns the
    fit problem: the observed histogram is the simulator at the true
    parameters with R = 1 and the held-out seed."""
    d, n = cfg["data"], cfg["n"]
    truth = np.asarray(cfg["args"], np.float32)
    m.load(d, SgDivSr=None)
    m.set_observed_hist(np.zeros(nbins, np.float32))
    for sid in range(m.streamNum):
        m.set_photophysics(sid, *cfg["photophysics"])
        m.set_burst_range(0, -1, sid)
    assert m.set_params(n, 0, truth)
    m.reSampleTimes = 1
    m.run_kernel(sid=0, seed=cfg["seeds"]["observed"], samples=True)
    m.synchronize(0)
    obs_E, _, _ = m.get_samples(0, d.nbursts)
    m.set_observed(obs_E, nbins, 2)
    x = obs_E[(obs_E >= 0) & (obs_E < 1)].astype(np.float64)
    observed = np.bincount(np.floor(x * nbins).astype(np.int64),
                           minlength=nbins)[:nbins].astype(np.float64)
    return fit.Problem(m, n, cfg["R"], cfg["seeds"]["mc"], observed)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--bursts", type=int, default=1_000_000)
    ap.add_argument("--seconds", type=float, default=60.0, help="wall limit per fitter")
    ap.add_argument("--pop", type=int, default=64)
    ap.add_argument("--pop-seed", type=int, default=None,
                    help="draw the 64 starting sets with this seed instead of the workload's")
    ap.add_argument("--lr", type=float, default=0.1)
    ap.add_argument("--r-grad", type=int, default=8,
                    help="resamples per burst of the RB gradient (the plain chi2 keeps R = 32)")
    ap.add_argument("--plateau", type=int, default=0,
                    help="Adam: halve the learning rate after this many steps without progress")
    ap.add_argument("--lr-decay", type=float, default=1.0, help="Adam: per-step decay")
    ap.add_argument("--streams", type=int, default=3,
                    help="stream slots the per-set genetic algorithm keeps in flight")
    ap.add_argument("--rank-resamples", type=int, default=4,
                    help="resamples per burst of the batched evaluation that ranks the starting "
                         "sets for the gradient fitters")
    ap.add_argument("--fitters", default="ga-per-set,ga-batched,adam-rb,lm-rb",
                    help="comma-separated subset of ga-per-set, ga-batched, adam-rb, lm-rb, "
                         "lm-rb-batched, adam-rb-batched")
    ap.add_argument("--chains", type=int, default=4, help="chains of the batched fitters")
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)

    def log(*x):
        print(*x, file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    cfg = synth.config(4, n_bursts=a.bursts)
    d, n, R = cfg["data"], cfg["n"], cfg["R"]
    truth = np.asarray(cfg["args"], np.float32)
    log(f"data: {d.nbursts} bursts, {d.nphotons} photons, {time.perf_counter() - t0:.1f} s")
    m = pda.MC(0, a.streams)
    prob = make_problem(m, cfg)
    chi2_truth = prob.chi2(fit.args_to_theta(truth, n))
    target = 1.01 * chi2_truth
    log(f"chi2(truth) = {chi2_truth:.5f}, target {target:.5f}")
    start_sets = cfg["population"]
    if a.pop_seed is not None:
        start_sets = synth.population(a.pop, np.random.default_rng(a.pop_seed), lo=1e3, hi=1e5,
                                      states=(n,))
    pop0 = [fit.args_to_theta(args, n) for _, args in start_sets[:a.pop]]

    results = {}
    fitters = a.fitters.split(",")
    ga_specs = [s for s in [
        ("ga-per-set", lambda: fit.GeneticFitter(batched=False, pop_size=a.pop,
                                                 rng=np.random.default_rng(7))),
        ("ga-batched", lambda: fit.GeneticFitter(batched=True, pop_size=a.pop,
                                                 rng=np.random.default_rng(7)))] if s[0] in fitters]
    for name, make in ga_specs:
        prob.evaluations = 0
        h = make().run(prob, pop0, a.seconds, target=target)
        results[name] = h
        log(f"{name}: best chi2 {h.chi2[-1]:.5f} after {h.times[-1]:.1f} s, "
            f"{prob.evaluations} evaluations, {len(h.times)} generations")

    # the gradient fitters start from the best initial set, found by one
    # timed batched evaluation with rank_resamples resamples per burst
    # (Levenberg-Marquardt restarts from the next best sets when a descent
    # converges short of the target)
    grad_specs = [s for s in [
        ("adam-rb", lambda: fit.AdamFitter(lr=a.lr, R_grad=a.r_grad, check_every=5,
                                           plateau=a.plateau, lr_decay=a.lr_decay)),
        ("lm-rb", lambda: fit.LMFitter(R_grad=a.r_grad)),
        ("lm-rb-batched", lambda: fit.BatchedLMFitter(chains=a.chains, R_grad=a.r_grad)),
        ("adam-rb-batched", lambda: fit.BatchedAdamFitter(chains=a.chains, lr=a.lr,
                                                          R_grad=a.r_grad, check_every=5))]
        if s[0] in fitters]
    ranking_seconds = {}
    for name, make in grad_specs:
        prob.evaluations = 0
        h, t_init = fit.fit_from_population(make(), prob, pop0, a.seconds, target=target,
                                            rank_resamples=a.rank_resamples,
                                            restarts=name != "adam-rb")
        results[name] = h
        ranking_seconds[name] = t_init
        log(f"{name}: best chi2 {h.chi2[-1]:.5f} after {h.times[-1]:.1f} s "
            f"(ranking {t_init:.3f} s), {prob.evaluations} evaluations")

    ga_final = results["ga-per-set"].chi2[-1] if "ga-per-set" in results else float("nan")
    out = {"workload": "4 states, rates 1e3-1e5, %d bursts x %d resamples, 50 bins" % (d.nbursts, R),
           "pop_seed": a.pop_seed, "rank_resamples": a.rank_resamples, "streams": a.streams,
           "ranking_seconds": ranking_seconds,
           "chi2_truth": chi2_truth, "target_truth": target, "ga_per_set_final": ga_final,
           "seconds_limit": a.seconds,
           "time_to_truth_target": {k: v.time_to(target) for k, v in results.items()},
           "time_to_ga_final": {k: v.time_to(1.01 * ga_final) for k, v in results.items()},
           "final_chi2": {k: v.chi2[-1] for k, v in results.items()},
           "rb_ms": dict(zip(("walk", "forward", "last_derivative_pass"), m.rb_stats())),
           "truth": truth.tolist(),
           "histories": {k: v.as_dict() for k, v in results.items()}}
    if a.out:
        with open(a.out, "w") as f:
            json.dump(out, f)
    short = {k: v for k, v in out.items() if k != "histories"}
    print(json.dumps(short), flush=True)
    m.close()


if __name__ == "__main__":
    main()
