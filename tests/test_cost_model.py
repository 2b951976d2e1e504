"""The algorithmic cost model (gsmfretda_amd/cost_model.json) matches the oracle.

The roofline of bench.py prices a walk launch as events per unit x their
weights.  The frozen event counts must be the oracle's on the same workloads
and random streams: recounted here on other bursts of the same workloads
(the frozen counts come from the first bursts), their per-unit means agree
within the sampling spread.  The counters themselves are pinned on a hand-
sized case: windows up to the one closing at the last photon, and the
comparisons of a binary search for each window's end.
"""
import numpy as np
import pytest

from gsmfretda_amd import cost_model, synth

EVENTS = ("live_windows", "live_philox_blocks", "search_steps", "binom_draws", "bg_draws",
          "live_binom_inv_iters", "live_poisson_iters")


@pytest.mark.parametrize("c,workload,b0,b1", [(1, "plain_3state", 50_000, 53_000),
                                               (4, "fast_dynamics", 500, 1500)])
def test_frozen_events_match_the_oracle(oracle, c, workload, b0, b1):
    cfg = synth.config(c, n_bursts=b1 if c == 1 else 2000)
    od = oracle.Data.from_bursts(cfg["data"])
    p = oracle.make_params(cfg["n"], cfg["args"], *cfg["photophysics"])
    o = oracle.run(od, p, R=cfg["R"], seed=cfg["seeds"]["mc"], b0=b0, b1=b1, want_units=False,
                   stats=True)["stats"]
    frozen = cost_model.load()["events_per_unit"][workload]["events"]
    for e in EVENTS:
        got = o[e] / o["units"]
        assert abs(got - frozen[e]) <= 0.05 * frozen[e] + 0.02, (e, got, frozen[e])
    assert o["live_btpe_iters"] == 0 and frozen["live_btpe_iters"] == 0


def test_counters_on_a_hand_sized_case(oracle):
    """Single window per unit (rates 1e-6 /s outlast T): one live window whose
    end search runs over all P photons: ceil(log2(P + 1)) comparisons; no
    background draws without background; the oracle still walks every window
    until the dwell sum reaches T, the algorithm only the live ones."""
    from tests.test_oracle_walk import _single_window_data
    d = _single_window_data(["DDDDDDDDDDDDA"] * 10 + ["D"] * 5)  # P = 13 and P = 1
    od = oracle.Data.from_bursts(d)
    p = oracle.make_params(2, synth.make_args([0.5, 0.5], [1e-6, 1e-6]), 1.0, 0.0, 0.0, 52.0)
    st = oracle.run(od, p, R=4, seed=3, stats=True)["stats"]
    assert st["units"] == 15 * 4
    assert st["live_windows"] == 10 * 4                  # P = 1 bursts have no live window
    assert st["search_steps"] == 10 * 4 * 4              # ceil(log2(14)) = 4
    assert st["bg_draws"] == 0 and st["binom_draws"] == 10 * 4
    # the window's block and the initial state's: the inflow of state 0 is
    # 1e-6 /s (at most 1), so the initial-state draw is open and counts
    assert st["live_philox_blocks"] == 10 * 4 * 2
    p = oracle.make_params(2, synth.make_args([0.5, 0.5], [2.0, 2.0]), 1.0, 0.0, 0.0, 52.0)
    st = oracle.run(od, p, R=4, seed=3, stats=True)["stats"]
    assert st["live_philox_blocks"] - st["live_windows"] == 0  # inflow 2 /s: state 0 is fixed
    assert st["windows"] >= st["live_windows"]


def test_ops_per_unit():
    cm = cost_model.load()
    if "weights_lane_cycles" not in cm:
        pytest.skip("weights not measured yet")
    for w in cm["events_per_unit"]:
        b = cost_model.breakdown(w, cm)
        assert cost_model.ops_per_unit(w, cm) == pytest.approx(sum(b.values()))
        assert all(v > 0 for v in b.values())


def test_frozen_rb_work_matches_the_oracle(oracle):
    """The Rao-Blackwellised work counts (outcomes Mt + 1 and leave-one-out
    terms Mt per counted state, per unit) recounted on other bursts of the
    same workload agree with the frozen ones within the sampling spread, and
    almost every unit's pmf comes from the memo (Mt < 48)."""
    cfg = synth.config(3, n_bursts=60_000)
    od = oracle.Data.from_bursts(cfg["data"])
    p = oracle.make_params(cfg["n"], cfg["args"], *cfg["photophysics"])
    o = dt = units = 0
    for b in range(50_000, 50_150):
        for r in range(cfg["R"]):
            cnt, _, _ = oracle.rb_unit(od, p, b, r, cfg["seeds"]["mc"], 0)
            mt = int(np.sum(cnt))
            o += mt + 1
            dt += mt * int(np.count_nonzero(cnt))
            units += 1
    w = cost_model.load()["rb_work_per_unit"]
    assert abs(o / units - w["outcomes"]) <= 0.05 * w["outcomes"], (o / units, w["outcomes"])
    assert abs(dt / units - w["deriv_terms"]) <= 0.08 * w["deriv_terms"], (dt / units, w)
    assert w["memo_share"] > 0.99
    assert cost_model.load()["weights_lane_cycles"]["rb_outcome"] > 0
