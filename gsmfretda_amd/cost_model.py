"""The algorithmic cost model of the walk kernels (frozen in cost_model.json).

A walk launch is priced in the events of the algorithm it runs, per
(burst, resample) unit:

    lane-ops per unit = sum over events e of events_e per unit x weight_e

events_per_unit[workload] holds the event counts of a workload (counted by
the CPU oracle on the same keyed random streams as the kernels, so they are
the GPU's), weights_lane_cycles the 32-bit lane-issue-cycles of one event
(microbenchmarked on gfx950).  The roofline fraction of a launch is then
lane-ops per unit x units / (kernel time x the lane-op peak).
"""
import json
import os

PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cost_model.json")

# event counter -> weight
WEIGHT_OF = {"units": "unit", "live_windows": "window", "live_philox_blocks": "philox_block",
             "search_steps": "search_step", "binom_draws": "binom_draw",
             "live_binom_inv_iters": "binom_iter", "bg_draws": "bg_draw",
             "live_poisson_iters": "poisson_iter", "live_btpe_iters": "btpe_iter"}


def load(path=PATH):
    with open(path) as f:
        return json.load(f)


def ops_per_unit(workload, cm=None):
    """Lane-operations per unit of `workload` (a key of events_per_unit)."""
    cm = load() if cm is None else cm
    ev = cm["events_per_unit"][workload]["events"]
    w = cm["weights_lane_cycles"]
    total = 0.0
    for e, count in ev.items():
        if count == 0:
            continue
        name = WEIGHT_OF[e]
        if name not in w:
            raise KeyError(f"no weight for event {e} ({count} per unit) in the cost model")
        total += count * w[name]
    return total


def breakdown(workload, cm=None):
    """Lane-operations per unit of each event of `workload`."""
    cm = load() if cm is None else cm
    ev = cm["events_per_unit"][workload]["events"]
    w = cm["weights_lane_cycles"]
    return {e: c * w[WEIGHT_OF[e]] for e, c in ev.items() if c}
