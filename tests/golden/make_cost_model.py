"""Writes the algorithmic cost model of the walk into gsmfretda_amd/cost_model.json.

The roofline of a walk kernel is priced in the algorithm's own events, not in
the instructions a build happens to issue:

    ops per unit = sum over events e of (events e per unit) x (weight of e)

* Events per unit (`--events`, on the CPU): the oracle's counters of the walk
  algorithm (or_stats in oracle/pda_oracle.h, from `units` on) on samples of
  the workloads bench.py times.  The oracle consumes the same keyed random
  streams as the kernels, so these are the events the GPU's units have.
* Weights (`--weights`, on the GPU): 32-bit lane-issue-cycles per event from
  the microbenchmarks of pda_debug_microbench (every lane of a full-
  occupancy grid repeats one event with a loop-carried dependence, so the
  loop is bound by VALU issue): GPU time x 256 CU x 64 lanes x 2.4 GHz per
  event, from the difference of two launch lengths (launch overhead
  cancels); the Poisson and binomial draws are split into a per-draw and a
  per-iteration weight from two means.

The weights are measured once and frozen with the event counts; bench.py
then reports ops per unit x units / kernel time against the lane-op peak.
This script only writes the model; run from the repository root:

    python tests/golden/make_cost_model.py --events            # CPU
    python tests/golden/make_cost_model.py --weights           # MI355X
    python tests/golden/make_cost_model.py --rb-weight         # MI355X
"""
import argparse
import json
import os
import sys
from multiprocessing import Pool

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from gsmfretda_amd.cost_model import PATH as COST_MODEL  # noqa: E402
PEAK = 256 * 64 * 2.4e9
EVENTS = ("units", "live_windows", "live_philox_blocks", "search_steps", "binom_draws",
          "live_binom_inv_iters", "live_btpe_iters", "bg_draws", "live_poisson_iters")

_G = {}


def _count(job):
    from oracle import oracle as O
    n, args, ph, R, seed, set_id, b1 = job
    p = O.make_params(n, args, *ph)
    o = O.run(_G["od"], p, R=R, seed=seed, set_id=set_id, b1=b1, want_units=False, stats=True)
    return o["stats"]


def _events(cfg, sets, R, seed, b1, procs):
    """Per-unit means of the algorithm's events over `sets` (list of
    (n, args, set_id)) on bursts [0, b1)."""
    from oracle import oracle as O
    _G["od"] = O.Data.from_bursts(cfg["data"])
    jobs = [(n, np.asarray(a, np.float32), cfg["photophysics"], R, seed, sid, b1)
            for n, a, sid in sets]
    with Pool(procs) as pool:
        stats = pool.map(_count, jobs)
    tot = {k: sum(s[k] for s in stats) for k in EVENTS}
    return {k: tot[k] / tot["units"] for k in EVENTS}, int(tot["units"])


def events_model(procs):
    from gsmfretda_amd import synth
    out = {}
    cfg = synth.config(1)
    ev, u = _events(cfg, [(cfg["n"], cfg["args"], 0)], cfg["R"], cfg["seeds"]["mc"], 20_000, procs)
    out["plain_3state"] = dict(events=ev, sample=f"bursts [0, 20000) of 100000 x {cfg['R']}, "
                               f"the true 3-state set, {u} units")
    cfg = synth.config(2)
    pop = cfg["population"]
    ev, u = _events(cfg, [(n, a, i) for i, (n, a) in enumerate(pop)], cfg["R"],
                    cfg["seeds"]["mc"], 1000, procs)
    out["population_256"] = dict(events=ev, sample=f"bursts [0, 1000) of 50000 x {cfg['R']}, "
                                 f"all 256 sets (set id = index), {u} units")
    cfg = synth.config(4)
    ev, u = _events(cfg, [(cfg["n"], cfg["args"], 0)], cfg["R"], cfg["seeds"]["mc"], 20_000, procs)
    out["fast_dynamics"] = dict(events=ev, sample=f"bursts [0, 20000) of 1000000 x {cfg['R']}, "
                                f"the true 4-state set, {u} units")
    pop = cfg["population"]
    ev, u = _events(cfg, [(n, a, i) for i, (n, a) in enumerate(pop)], cfg["R"],
                    cfg["seeds"]["mc"], 1000, procs)
    out["fast_dynamics_population_64"] = dict(
        events=ev, sample=f"bursts [0, 1000) of 1000000 x {cfg['R']}, all 64 starting sets "
                          f"(set id = index), {u} units")
    return out


def rb_model(bursts=2000):
    """Per-unit work of the Rao-Blackwellised passes on the 3-state workload
    (100,000 bursts x 64; the first `bursts` bursts), from each unit's
    photon counts per state m_s (the oracle's rb_unit), Mt = sum m_s:
      outcomes: Mt + 1 per unit -- each is binned (E = a / F) and weighted
        (its probability in 2^-32 fixed point) by every pass;
      deriv_terms: Mt per state with m_s > 0 -- the leave-one-trial-out
        probabilities P_s(c), c < Mt, each a multiply-add against the
        cotangent differences (adjoint) or a pair of updates at a bin
        boundary (Jacobian);
      conv_macs: the multiply-adds of building the unit's pmf by convolution
        (the first state a copy, each later one l (m_s + 1) for a pmf of
        length l): the work a per-unit build would do, which the memo of
        per-count-vector pmfs replaces by one build per vector;
      memo_share: the units whose total is below the memo's cap (48)."""
    from gsmfretda_amd import synth
    from oracle import oracle as O
    cfg = synth.config(3)
    od = O.Data.from_bursts(cfg["data"])
    p = O.make_params(cfg["n"], cfg["args"], *cfg["photophysics"])
    R, seed = cfg["R"], cfg["seeds"]["mc"]
    macs = outcomes = dterms = memo = 0
    units = 0
    for b in range(bursts):
        for r in range(R):
            cnt, F, _ = O.rb_unit(od, p, b, r, seed, 0)
            ms = [int(m) for m in cnt if m > 0]
            mt = sum(ms)
            length = 1
            for i, m in enumerate(ms):
                if i > 0:
                    macs += length * (m + 1)
                length += m
            outcomes += mt + 1
            dterms += mt * len(ms)
            memo += 1 if mt < 48 else 0
            units += 1
    return {"outcomes": outcomes / units, "deriv_terms": dterms / units,
            "conv_macs": macs / units, "memo_share": memo / units,
            "sample": f"bursts [0, {bursts}) of 100000 x {R}, the true 3-state set, {units} units"}


def _slope(m, op, iters, param=0.0):
    """(lane-issue-cycles per event, mean inner iterations per event)."""
    ms1, e1, i1 = m.debug_microbench(op, iters, param)
    ms2, e2, i2 = m.debug_microbench(op, 2 * iters, param)
    return (ms2 - ms1) * 1e-3 * PEAK / (e2 - e1), (i2 - i1) / (e2 - e1), ms2


def weights_model():
    from gsmfretda_amd import pda
    w, raw = {}, {}
    with pda.MC(0, 1) as m:
        for name, op, iters in (("philox_block", 0, 4000), ("window", 1, 4000),
                                ("search_step", 2, 6000), ("unit", 5, 8000)):
            s, _, ms = _slope(m, op, iters)
            w[name] = s
            raw[name] = dict(op=op, iters=[iters, 2 * iters], ms_long=ms)
        # draw + iterations: slopes at two means give the per-draw and
        # per-iteration weights
        for name, op, iters, lo, hi in (("poisson", 3, 4000, 1e-4, 8.0),
                                        ("binom", 4, 4000, 2.0, 60.0)):
            s_lo, i_lo, _ = _slope(m, op, iters, lo)
            s_hi, i_hi, _ = _slope(m, op, iters, hi)
            per_iter = (s_hi - s_lo) / (i_hi - i_lo)
            w[name + "_iter"] = per_iter
            w["bg_draw" if name == "poisson" else "binom_draw"] = s_lo - per_iter * i_lo
            raw[name] = dict(op=op, params=[lo, hi], iterations_per_draw=[i_lo, i_hi],
                             cycles_per_draw=[s_lo, s_hi])
    w["poisson_iter"] = w.pop("poisson_iter")
    w["binom_iter"] = w.pop("binom_iter")
    return w, raw


def rb_weight_model():
    """Lane-issue-cycles of one Rao-Blackwellised outcome (op 6)."""
    from gsmfretda_amd import pda
    with pda.MC(0, 1) as m:
        s, _, ms = _slope(m, 6, 8000)
    return s, dict(op=6, iters=[8000, 16000], ms_long=ms)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--events", action="store_true")
    ap.add_argument("--weights", action="store_true")
    ap.add_argument("--rb", action="store_true")
    ap.add_argument("--rb-weight", action="store_true",
                    help="measure only the Rao-Blackwellised outcome weight (MI355X)")
    ap.add_argument("--procs", type=int, default=8)
    ap.add_argument("--print", action="store_true", help="print instead of writing the model")
    a = ap.parse_args()
    cm = json.load(open(COST_MODEL))
    if a.events:
        cm["events_per_unit"] = events_model(a.procs)
        cm["events_source"] = ("oracle event counters (or_stats, from `units` on) on the "
                               "workloads' first bursts: python tests/golden/make_cost_model.py "
                               "--events")
    if a.rb:
        cm["rb_work_per_unit"] = rb_model()
        cm["rb_work_source"] = ("per-state photon counts of each unit (oracle rb_unit): python "
                                "tests/golden/make_cost_model.py --rb")
    if a.weights:
        w, raw = weights_model()
        cm["weights_lane_cycles"] = w
        cm["weights_raw"] = raw
        cm["weights_source"] = ("pda_debug_microbench on one MI355X, lane-issue-cycles per "
                                "event = GPU time x 256 CU x 64 lanes x 2.4 GHz / events, "
                                "difference of two launch lengths: python "
                                "tests/golden/make_cost_model.py --weights")
    if a.rb_weight:
        s, raw = rb_weight_model()
        cm["weights_lane_cycles"]["rb_outcome"] = s
        cm["weights_raw"]["rb_outcome"] = raw
    if a.print:
        print(json.dumps(cm, indent=2))
    else:
        with open(COST_MODEL, "w") as f:
            json.dump(cm, f, indent=2)
            f.write("\n")


if __name__ == "__main__":
    main()
