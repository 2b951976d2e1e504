"""Benchmark: burst-resamples/sec of the Monte-Carlo PDA hot path on one MI355X.

Workload: 3-state plain MC PDA at scale -- 100,000 synthetic bursts (50 +
Geom(mean 40) photons, LogNormal(1 ms, 0.5) durations, 12.5 ns ticks), the
reference's full evaluation per step: 100 resamples per burst, the 50-bin E
histogram and its chi2, all on the GPU, bursts already resident in HBM.
One step = one parameter-set evaluation = 1e7 burst-resamples, its histogram
and chi2 read back on the host; the timed steps take turns on three streams
(the reference's several streams per GPU; --streams 1 times them one at a
time).  The kernel durations behind the
rooflines come from a separate run of the same evaluations one at a time.

Besides the headline number the line carries the VALU roofline of the walk
kernel on two bases -- the algorithm's events (counted by the oracle on the
same random streams) times their microbenchmarked weights, and the kernel's
own VALU lane-cycle counter -- and its HBM roofline; a replay leg (the same
bursts under a 30 kHz background, where a few per cent of units leave the
fast form); the batched-population leg (256 sets in one graph replay); the
fast-dynamics legs at 1,000,000 bursts (64 sets x 32 resamples in one graph
replay, and the fit wall time to 1.01 x chi2(truth) of the per-set genetic
algorithm against the gradient fitters); the Rao-Blackwellised leg
(forward, adjoint and Jacobian GPU times) and the CPU baseline (the oracle
on the host's cores).

Prints one JSON line (rank 0).  --gpus above 1 is refused: this build runs
on one GPU (the multi-GPU design is parameter-set or burst sharding with a
histogram all-reduce).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "burst-resamples/sec on 1 MI355X; fit wall-time to chi2 tol (grad+batch vs plain)"
COST_MODEL = os.path.join(ROOT, "gsmfretda_amd", "cost_model.json")


_T0 = time.perf_counter()


def note(msg):
    """Progress line on stderr (the JSON result line is the only stdout)."""
    print(f"[bench {time.perf_counter() - _T0:7.1f} s] {msg}", file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(cfg, threads, sample_bursts, repeats=5, single_bursts=2000):
    """The oracle (REF-semantics CPU MC on the same streams) on host cores:
    median of `repeats` evaluations of the first `sample_bursts` bursts on
    `threads` threads, and the single-thread oracle on the first
    `single_bursts` bursts."""
    from oracle import oracle as O
    d = cfg["data"]
    od = O.Data.from_bursts(d)
    p = O.make_params(cfg["n"], cfg["args"], *cfg["photophysics"])
    R = cfg["R"]
    seed = cfg["seeds"]["mc"]
    b1 = min(sample_bursts, d.nbursts)
    O.run_threaded(od, p, R=R, seed=seed, b1=min(b1, 2000), nthreads=threads)
    times = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        O.run_threaded(od, p, R=R, seed=seed, b1=b1, nthreads=threads)
        times.append(time.perf_counter() - t0)
    dt = float(np.median(times))
    s1 = min(single_bursts, d.nbursts)
    t0 = time.perf_counter()
    O.run_threaded(od, p, R=R, seed=seed, b1=s1, nthreads=1)
    dt1 = time.perf_counter() - t0
    return dict(value=b1 * R / dt, unit="burst-resamples/s", cores=threads, kind="port",
                sample=f"{b1} of {d.nbursts} bursts x {R} resamples (the whole evaluation), "
                       f"oracle/pda_oracle.c literal photon scan, median of {repeats} runs, "
                       f"{threads} threads, {dt:.2f} s wall each",
                cpu=cpu_model(), host_cpus=os.cpu_count(),
                single_thread={"value": s1 * R / dt1, "sample": f"{s1} bursts x {R} resamples, "
                               f"1 thread, {dt1:.2f} s"})


def valu_roofline(workload, units, ms, kernel, traffic=None):
    """VALU roofline of a walk launch priced by the algorithmic cost model:
    the workload's events per unit x their weights, over the launch time."""
    from gsmfretda_amd import cost_model
    cm = cost_model.load()
    opu = cost_model.ops_per_unit(workload, cm)
    achieved = opu * units / (ms * 1e-3) / 1e12
    peak = cm["valu_peak_lane_ops_per_s"] / 1e12
    return {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TOP/s",
            "frac": achieved / peak, "traffic": traffic, "kernel": kernel, "kernel_ms_avg": ms,
            "ops_per_unit": opu, "basis": "algorithm: oracle events x microbenchmarked weights",
            "events": workload}


def roofline(units, walk_ms):
    """The walk kernel of the headline evaluation: VALU on the algorithmic
    basis, VALU on the counter basis (its own frozen lane-cycle count), HBM."""
    from gsmfretda_amd import cost_model
    cm = cost_model.load()
    t = walk_ms * 1e-3
    traffic = cm["hbm_traffic_bytes_per_launch"]
    valu = valu_roofline("plain_3state", units, walk_ms, "walk_kernel<0>", traffic)
    ops = cm["valu_lane_ops_per_unit"] * units
    achieved = ops / t / 1e12
    peak = cm["valu_peak_lane_ops_per_s"] / 1e12
    counter = {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TOP/s",
               "frac": achieved / peak, "traffic": traffic, "kernel": "walk_kernel<0>",
               "kernel_ms_avg": walk_ms, "ops_per_unit": cm["valu_lane_ops_per_unit"],
               "basis": "SQ_THREAD_CYCLES_VALU per unit of an earlier build, frozen"}
    hb = cm["hbm_algorithmic_bytes_per_unit"] * units / t / 1e9
    hbm = {"bound": "hbm", "achieved": hb, "peak": cm["hbm_peak_bytes_per_s"] / 1e9,
           "unit": "GB/s", "frac": hb * 1e9 / cm["hbm_peak_bytes_per_s"], "traffic": traffic}
    return valu, counter, hbm


def replay_leg(pda, cfg, steps):
    """The headline bursts under a 30 kHz donor / 15 kHz acceptor background:
    windows longer than ~1 ms have background means above 64, so a few per
    cent of units leave the fast form for the grid-wide replay pass.  Reports
    the replayed units and the GPU time of the walk and of replay + finish."""
    d = cfg["data"]
    bg = (d.bg_ad, d.bg_dd)
    d.bg_dd, d.bg_ad = 3e4, 1.5e4
    try:
        m = pda.MC(0, 1)
        m.load(d, SgDivSr=None)
        m.set_observed_hist(np.ones(50, np.float32))
        m.set_photophysics(0, *cfg["photophysics"])
        assert m.set_params(cfg["n"], 0, cfg["args"])
        m.set_burst_range(0, -1, 0)
        m.reSampleTimes = cfg["R"]
        walk, rest = [], []
        for i in range(steps + 2):
            m.run_kernel(sid=0, seed=cfg["seeds"]["mc"], set_id=0)
            m.synchronize(0)
            w, f = m.last_pass_ms(0)
            if i >= 2:
                walk.append(w)
                rest.append(f)
        units, replays = m.run_stats(0)
        m.close()
    finally:
        d.bg_ad, d.bg_dd = bg
    w, r = float(np.median(walk)), float(np.median(rest))
    return {"workload": "3 states, 100000 bursts x 100 resamples, background 30 kHz donor / "
                        "15 kHz acceptor",
            "units": units, "replayed_units": replays, "replayed_fraction": replays / units,
            "walk_gpu_ms": w, "replay_and_finish_gpu_ms": r, "total_over_walk": (w + r) / w}


def population_leg(pda, synth, steps, warmup):
    """Batched population evaluation: 256 parameter sets (2-4 states) over
    50,000 bursts x 32 resamples per generation, one graph replay each,
    launch shape picked by the autotuner first."""
    cfg = synth.config(2)
    d, R, pop = cfg["data"], cfg["R"], cfg["population"]
    m = pda.MC(0, 1)
    m.load(d, SgDivSr=None)
    m.set_observed_hist(np.zeros(50, np.float32))
    m.set_photophysics(0, *cfg["photophysics"])
    assert m.set_params(cfg["n"], 0, cfg["args"])
    m.set_burst_range(0, -1, 0)
    m.reSampleTimes = 1
    m.run_kernel(sid=0, seed=cfg["seeds"]["observed"], samples=True)
    m.synchronize(0)
    obs_E, _, _ = m.get_samples(0, d.nbursts)
    m.set_observed(obs_E, 50, 2)
    m.set_population([n for n, _ in pop], [a for _, a in pop])
    tuned = m.autotune(population=True, R=R, reps=2)
    seed = cfg["seeds"]["mc"]
    for _ in range(warmup):
        m.run_population(R=R, seed=seed)
        m.get_population_result()
    gpu_ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        m.run_population(R=R, seed=seed)
        _, c64, _, _ = m.get_population_result()
        gpu_ms.append(m.population_stats()[3])
    dt = time.perf_counter() - t0
    units = d.nbursts * R * len(pop)
    st = m.population_stats()
    m.close()
    # VALU roofline of the whole generation (graph replay timed with HIP
    # events): algorithmic basis, and the frozen lane-cycle counter
    cm = json.load(open(COST_MODEL))
    g_ms = float(np.mean(gpu_ms))
    ops = cm["population_valu_lane_ops_per_unit"] * units
    achieved = ops / (g_ms * 1e-3) / 1e12
    peak = cm["valu_peak_lane_ops_per_s"] / 1e12
    return {"workload": "batched population: 256 parameter sets (2-4 states, rates 1e2-1e4), "
                        "50000 bursts x 32 resamples, 50 bins, per-set chi2, one hipGraph replay "
                        "per generation",
            "value": units * steps / dt, "unit": "burst-resamples/s",
            "roofline": valu_roofline("population_256", units, g_ms,
                                      "pop_walk_kernel<0> (whole generation)"),
            "roofline_counter": {"bound": "valu", "achieved": achieved, "peak": peak,
                                 "unit": "TOP/s", "frac": achieved / peak,
                                 "ops_per_unit": cm["population_valu_lane_ops_per_unit"],
                                 "basis": "SQ_THREAD_CYCLES_VALU per unit, frozen"},
            "ms_per_generation": dt / steps * 1e3, "gpu_ms_per_generation": float(np.mean(gpu_ms)),
            "units_per_generation": units, "graph_captures": st[2], "replayed_units": st[1],
            "autotuned_shape": {"batch_tiles": tuned[0], "waves_per_cu": tuned[1],
                                "queues": tuned[2]},
            "chi2_median": float(np.median(c64))}


def rb_leg(pda, synth, steps):
    """Rao-Blackwellised differentiable form: 3 states (6 rates + 3 E),
    100,000 bursts x 64 resamples per evaluation, expected histogram and chi2
    (forward), then d chi2 / d(E, K) (adjoint)."""
    cfg = synth.config(3)
    d, R, n = cfg["data"], cfg["R"], cfg["n"]
    m = pda.MC(0, 1)
    m.load(d, SgDivSr=None)
    m.set_observed_hist(np.zeros(50, np.float32))
    m.set_photophysics(0, *cfg["photophysics"])
    assert m.set_params(n, 0, cfg["args"])
    m.set_burst_range(0, -1, 0)
    m.reSampleTimes = 1
    m.run_kernel(sid=0, seed=cfg["seeds"]["observed"], samples=True)
    m.synchronize(0)
    obs_E, _, _ = m.get_samples(0, d.nbursts)
    m.set_observed(obs_E, 50, 2)
    x = obs_E[(obs_E >= 0) & (obs_E < 1)].astype(np.float64)
    O = np.bincount(np.floor(x * 50).astype(np.int64), minlength=50)[:50].astype(np.float64)
    seed = cfg["seeds"]["mc"]
    fwd, adj, walk, jac = [], [], [], []
    for i in range(steps + 1):
        m.rb_forward(R=R, seed=seed, scores=True)
        M, _, c2, _ = m.rb_result()
        mm = M / R
        keep = (np.arange(50) >= 2) & (M > 0)
        g = np.zeros(50)
        g[keep] = (1.0 - O[keep] ** 2 / mm[keep] ** 2) / (R * (keep.sum() - n * (n + 1)))
        m.rb_adjoint(g, n)
        w, f, a = m.rb_stats()
        m.rb_jacobian(n, 50)
        j = m.rb_stats()[2]
        if i:
            fwd.append(f)
            adj.append(a)
            walk.append(w)
            jac.append(j)
    m.close()
    units = d.nbursts * R
    out = {"workload": "Rao-Blackwellised PDA, 3 states, 100000 bursts x 64 resamples, 50 bins: "
                       "expected histogram + chi2, then d chi2 / d(E, K)",
           "forward_units_per_s": units / (np.mean(fwd) * 1e-3),
           "forward_gpu_ms": float(np.mean(fwd)), "walk_gpu_ms": float(np.mean(walk)),
           "adjoint_gpu_ms": float(np.mean(adj)), "jacobian_gpu_ms": float(np.mean(jac)),
           "forward_plus_adjoint_units_per_s": units / ((np.mean(fwd) + np.mean(adj)) * 1e-3),
           "chi2": c2}
    # rooflines of the pmf passes, in the algorithm's work per unit (frozen
    # in the cost model from each unit's photon counts per state): every
    # pass bins and weights each of the unit's Mt + 1 outcomes (the
    # microbenchmarked outcome weight); the adjoint adds a double
    # multiply-add per leave-one-out term and per outcome and rate, the
    # Jacobian an LDS update per outcome and rate and two per leave-one-out
    # term (bin boundaries, at most); a multiply-add is two lane-ops (no
    # contraction).  The pmfs themselves come from the per-count-vector
    # memo, built beside the walk: the forward's pmf time is the forward's
    # time after the walk.
    from gsmfretda_amd import cost_model
    cm = cost_model.load()
    w = cm["rb_work_per_unit"]
    wo = cm["weights_lane_cycles"].get("rb_outcome")
    peak = cm["valu_peak_lane_ops_per_s"]
    lds_peak = 256 * 128 * 2.4e9  # bytes/s: 128 B per clock per CU
    K = n * (n - 1)
    o, dt = w["outcomes"], w["deriv_terms"]
    pmf_fwd_ms = float(np.mean(fwd)) - float(np.mean(walk))
    if wo is not None:
        for name, ms, ops, lds_bytes in (
                ("forward_pmf", pmf_fwd_ms, o * wo, 8.0 * o),
                ("adjoint", float(np.mean(adj)), o * wo + 2.0 * (dt + o * K), 8.0 * o * (K + 2)),
                ("jacobian", float(np.mean(jac)), o * wo + 2.0 * (o * K + 2.0 * dt),
                 8.0 * (o * K + 2.0 * dt))):
            t = ms * 1e-3
            out[f"roofline_{name}"] = {
                "bound": "valu", "achieved": ops * units / t / 1e12, "peak": peak / 1e12,
                "unit": "TOP/s", "frac": ops * units / t / peak, "traffic": None,
                "lds_frac": lds_bytes * units / t / lds_peak, "ops_per_unit": ops,
                "outcomes_per_unit": o, "kernel_ms": ms}
    return out


def fast_population_leg(pda, cfg, steps):
    """Fast-dynamics population: the 64 starting sets of a fit (4 states,
    rates 1e3-1e5 s^-1) over 1,000,000 bursts x 32 resamples in one graph
    replay per generation (2.05e9 units)."""
    d, R, pop = cfg["data"], cfg["R"], cfg["population"]
    m = pda.MC(0, 1)
    m.load(d, SgDivSr=None)
    m.set_observed_hist(np.ones(50, np.float32))
    m.set_photophysics(0, *cfg["photophysics"])
    m.set_burst_range(0, -1, 0)
    m.set_population([n for n, _ in pop], [a for _, a in pop])
    seed = cfg["seeds"]["mc"]
    m.run_population(R=R, seed=seed)
    m.get_population_result()
    gpu_ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        m.run_population(R=R, seed=seed)
        m.get_population_result()
        gpu_ms.append(m.population_stats()[3])
    dt = time.perf_counter() - t0
    st = m.population_stats()
    m.close()
    units = d.nbursts * R * len(pop)
    g_ms = float(np.mean(gpu_ms))
    return {"workload": "fast-dynamics population: 64 parameter sets (4 states, rates 1e3-1e5), "
                        "1000000 bursts x 32 resamples, 50 bins, per-set chi2, one hipGraph "
                        "replay per generation",
            "value": units * steps / dt, "unit": "burst-resamples/s",
            "units_per_generation": units, "ms_per_generation": dt / steps * 1e3,
            "gpu_ms_per_generation": g_ms, "replayed_units": st[1],
            "roofline": valu_roofline("fast_dynamics_population_64", units, g_ms,
                                      "pop_walk_kernel<0> (whole generation)")}


def fit_leg(pda, cfg, limits):
    """Fit wall time to 1.01 x chi2(truth) on the fast-dynamics workload at
    its full size (4 states, rates 1e3-1e5 s^-1, 1,000,000 bursts x 32
    resamples): the per-set genetic algorithm (one pda_run per individual on
    three stream slots in flight, the reference's pattern) against the
    gradient fitters on the RB form, single-chain and batched (four chains
    in lockstep through the RB population passes), from the same 64 starting
    sets (the gradient fitters start from the best of them, ranked by one
    timed batched evaluation with 4 resamples per burst).  Each fitter stops at the
    target or at its wall limit; a speed-up over a fitter that did not reach
    the target within its limit is a lower bound."""
    from gsmfretda_amd import fit, fit_study
    n = cfg["n"]
    m = pda.MC(0, 3)
    prob = fit_study.make_problem(m, cfg)
    c_truth = prob.chi2(fit.args_to_theta(np.asarray(cfg["args"], np.float32), n))
    target = 1.01 * c_truth
    pop0 = [fit.args_to_theta(a, n) for _, a in cfg["population"][:64]]
    res = {}
    note(f"fit: chi2(truth) {c_truth:.5g}, per-set genetic algorithm (limit {limits['ga-per-set']} s)")
    h = fit.GeneticFitter(batched=False, pop_size=64, rng=np.random.default_rng(7)).run(
        prob, pop0, limits["ga-per-set"], target=target)
    res["ga-per-set"] = h
    for name, make in (("adam-rb", lambda: fit.AdamFitter(lr=0.1, R_grad=8, check_every=5)),
                       ("lm-rb", lambda: fit.LMFitter(R_grad=8)),
                       ("lm-rb-batched", lambda: fit.BatchedLMFitter(chains=4, R_grad=8)),
                       ("adam-rb-batched",
                        lambda: fit.BatchedAdamFitter(chains=4, lr=0.1, R_grad=8, check_every=5))):
        note(f"fit: {name} (limit {limits[name]} s)")
        h, _ = fit.fit_from_population(make(), prob, pop0, limits[name], target=target,
                                       rank_resamples=4, restarts=name != "adam-rb")
        res[name] = h
    m.close()
    t = {k: v.time_to(target) for k, v in res.items()}
    out = {"workload": f"fit to 1.01 x chi2(truth): 4 states, rates 1e3-1e5 s^-1, "
                       f"{cfg['data'].nbursts} bursts x 32 resamples, 50 bins, 64 starting sets; "
                       f"wall seconds per fitter",
           "bursts": cfg["data"].nbursts, "chi2_truth": c_truth, "seconds_limit": limits,
           "seconds_to_target": t, "final_chi2": {k: v.chi2[-1] for k, v in res.items()},
           "evaluations": {k: v.evals[-1] for k, v in res.items()}}
    ga = t["ga-per-set"]
    for name in ("lm-rb", "adam-rb", "lm-rb-batched", "adam-rb-batched"):
        if t[name]:
            out[f"speedup_{name.replace('-rb', '').replace('-', '_')}_vs_ga_per_set"] = (
                ga if ga else limits["ga-per-set"]) / t[name]
    out["speedup_is_lower_bound"] = ga is None
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-population", action="store_true",
                    help="skip the batched-population leg")
    ap.add_argument("--no-fit", action="store_true", help="skip the fit wall-time leg")
    ap.add_argument("--fit-seconds", type=float, default=45.0,
                    help="wall limit of the genetic algorithm and Adam (LM: half)")
    ap.add_argument("--fit-bursts", type=int, default=1_000_000,
                    help="bursts of the fast-dynamics workload (fit and population legs)")
    ap.add_argument("--no-replay", action="store_true", help="skip the replay leg")
    ap.add_argument("--no-fast", action="store_true",
                    help="skip the fast-dynamics population leg")
    ap.add_argument("--no-rb", action="store_true",
                    help="skip the Rao-Blackwellised (differentiable) leg")
    ap.add_argument("--cpu-bursts", type=int, default=100_000)
    ap.add_argument("--streams", type=int, default=3,
                    help="streams the timed evaluations take turns on (1: one at a time)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain copy kernel of MIB MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU path is "
              "not part of this build", file=sys.stderr)
        sys.exit(2)

    from gsmfretda_amd import pda, synth

    if args.copy_calibration:
        n = args.copy_calibration * (1 << 20) // 4
        with pda.MC(0, 1) as m:
            x = np.arange(n, dtype=np.uint32)
            y = m.debug_primitive(9, x, np.uint32, n)
            assert np.array_equal(x, y)
        print(json.dumps({"copy_calibration_bytes": n * 4}), flush=True)
        return

    cfg = synth.config(1)
    d = cfg["data"]
    n, R = cfg["n"], cfg["R"]
    S = max(1, args.streams)
    m = pda.MC(0, S)
    m.load(d, SgDivSr=None)
    # observed histogram: the simulator at the true parameters, R = 1, held-out seed
    m.set_observed_hist(np.zeros(50, np.float32))
    for sid in range(S):
        m.set_photophysics(sid, *cfg["photophysics"])
        assert m.set_params(n, sid, cfg["args"])
        m.set_burst_range(0, -1, sid)
    m.reSampleTimes = 1
    m.run_kernel(sid=0, seed=cfg["seeds"]["observed"], samples=True)
    m.synchronize(0)
    obs_E, _, _ = m.get_samples(0, d.nbursts)
    m.set_observed(obs_E, 50, 2)

    m.reSampleTimes = R
    seed = cfg["seeds"]["mc"]
    # kernel durations (HIP events): one evaluation at a time on stream 0
    for _ in range(args.warmup):
        m.run_kernel(sid=0, seed=seed, set_id=0)
    m.synchronize(0)
    kernel_ms, walk_ms = [], []
    t_one = time.perf_counter()
    for _ in range(args.steps):
        m.run_kernel(sid=0, seed=seed, set_id=0)
        m.synchronize(0)
        kernel_ms.append(m.last_run_ms(0))
        walk_ms.append(m.last_pass_ms(0)[0])
    t_one = time.perf_counter() - t_one
    hist, chi2_64, chi2_ref, effN = m.get_hist_chi2(0)

    def evaluations(k):
        """k evaluations in turn over S streams, as the reference keeps
        several streams of a GPU in flight (src/gpuWorker.cpp:38-88): each
        evaluation's histogram and chi2 are read back before its stream
        takes the next one."""
        out = []
        for i in range(k):
            if i >= S:
                m.synchronize(i % S)
                out.append(m.get_hist_chi2(i % S))
                walk_in_flight.append(m.last_pass_ms(i % S)[0])
            m.run_kernel(sid=i % S, seed=seed, set_id=0)
        for i in range(max(0, k - S), k):
            m.synchronize(i % S)
            out.append(m.get_hist_chi2(i % S))
            walk_in_flight.append(m.last_pass_ms(i % S)[0])
        return out

    walk_in_flight = []

    # whole-job throughput (timed)
    evaluations(args.warmup)
    walk_in_flight.clear()
    t0 = time.perf_counter()
    res = evaluations(args.steps)
    dt = time.perf_counter() - t0
    for h, c64, _, _ in res:  # same streams, same parameters: the same result
        assert np.array_equal(h, hist) and c64 == chi2_64
    units = d.nbursts * R
    value = units * args.steps / dt

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "burst-resamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "plain MC PDA at scale: 3 states, 100000 bursts x 100 resamples, "
                               "50 E bins, chi2 on device",
                   "bursts": d.nbursts, "photons": int(d.nphotons), "resamples": R,
                   "bins": 50, "states": n, "parallelism": "single GPU", "streams": S},
        "gpu_ms_per_eval_avg": float(np.mean(kernel_ms)),
        "one_at_a_time": {"value": units * args.steps / t_one,
                          "ms_per_step": t_one / args.steps * 1e3},
        "chi2": chi2_64,
    }
    out["roofline"], out["roofline_counter"], out["roofline_hbm"] = roofline(
        units, float(np.mean(walk_ms)))
    out["walk_ms_in_flight_avg"] = float(np.mean(walk_in_flight))
    m.close()
    note(f"headline: {value:.4g} burst-resamples/s")
    if not args.no_replay:
        out["replay"] = replay_leg(pda, cfg, max(3, args.steps // 4))
        note("replay leg done")
    if not args.no_population:
        out["population"] = population_leg(pda, synth, max(1, args.steps // 4), 1)
        note("population leg done")
    if not (args.no_fast and args.no_fit):
        note(f"generating the fast-dynamics workload ({args.fit_bursts} bursts)")
        cfg4 = synth.config(4, n_bursts=args.fit_bursts)
        if not args.no_fast:
            out["fast_dynamics_population"] = fast_population_leg(pda, cfg4,
                                                                  max(2, args.steps // 4))
            note("fast-dynamics population leg done")
        if not args.no_fit:
            lim = args.fit_seconds
            out["fit"] = fit_leg(pda, cfg4, {"ga-per-set": lim, "adam-rb": lim,
                                             "lm-rb": lim / 2, "lm-rb-batched": lim / 2,
                                             "adam-rb-batched": lim})
            note("fit leg done")
        del cfg4
    if not args.no_rb:
        out["rao_blackwellised"] = rb_leg(pda, synth, max(2, args.steps // 4))
        note("Rao-Blackwellised leg done")
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        out["cpu_baseline"] = cpu_baseline(cfg, threads, args.cpu_bursts)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
