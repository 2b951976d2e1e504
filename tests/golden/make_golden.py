"""Writes tests/golden/walk_small.npz: a small, self-contained regression
fixture of the oracle's per-sample outputs (acc, F, E) and histograms.

Inputs: 120 synthetic bursts (C0-like photon and duration laws, 10 % of
photons in neither channel) and three parameter sets (2, 3 and 4 states, one
with distance broadening), background on.  Run from the repository root:
    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from gsmfretda_amd import synth  # noqa: E402
from oracle import oracle as O  # noqa: E402


def main():
    args2 = synth.make_args([0.3, 0.7], [1e3, 1e3])
    d = synth.generate(120, args2, 2, seed=1234, p_neither=0.1)
    sets = [
        (2, args2, (1.0, 0.0, 0.05, 52.0)),
        (3, synth.make_args([0.15, 0.5, 0.85], [300, 2000, 800, 5000, 150, 9000]),
         (0.34, 0.08, 0.07, 52.0)),
        (4, synth.make_args([0.1, 0.4, 0.6, 0.9], np.linspace(500, 6000, 12), [1.5, 0, 3, 0]),
         (1.0, 0.0, 0.05, 52.0)),
    ]
    R, seed = 8, 424242
    od = O.Data.from_bursts(d)
    out = dict(start=d.start, stop=d.stop, istart=d.istart, phcount=d.phcount, times=d.times,
               mask_ad=d.mask_ad, mask_dd=d.mask_dd, T=d.T, clk_p=np.float64(d.clk_p),
               bg_ad=np.float64(d.bg_ad), bg_dd=np.float64(d.bg_dd), R=np.int64(R),
               seed=np.int64(seed), nsets=np.int64(len(sets)))
    for k, (n, args, ph) in enumerate(sets):
        p = O.make_params(n, args, *ph)
        r = O.run(od, p, R=R, seed=seed, set_id=k)
        out.update({f"n{k}": np.int64(n), f"args{k}": args, f"ph{k}": np.array(ph, np.float32),
                    f"acc{k}": r["acc"], f"F{k}": r["F"], f"E{k}": r["E"], f"hist{k}": r["hist"]})
    np.savez_compressed(os.path.join(ROOT, "tests", "golden", "walk_small.npz"), **out)


if __name__ == "__main__":
    main()
