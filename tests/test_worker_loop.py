"""The reference's worker loop, unchanged in structure, against the MI355X
`class mc` (tests/host/ps_loop.cpp): one GPU thread and one network thread per
stream in the dataready state machine of src/gpuWorker.cpp:38-88 and
src/streamWorker.cpp:39-248, talking args.proto messages (protobuf wire
format, "c"/"p"/"k"/"r" prefixes, -1/-2 sentinels) to a fake in-process
parameter server over an in-process REQ/REP channel.

Every parameter set of the server's population must come back exactly once,
with the chi2 the reference's own float formula gives on the oracle's
histogram for the stream's random streams (set id = stream id, common random
numbers), the observed histogram the server receives must be the binning of
SgDivSr, and the simulated histogram must be sent whenever a chi2 improves on
the best so far.
"""
import subprocess

import numpy as np
import pytest

from gsmfretda_amd import synth
from tests.conftest import host_program

pytestmark = pytest.mark.gpu

PH = (0.34, 0.08, 0.07, 52.0)   # the mc shim's photophysics (src/mc.cu:89-91)


def test_worker_loop_with_fake_parameter_server(oracle, tmp_path):
    n, R, seed, streams, nsets = 3, 8, 4242, 2, 6
    truth = synth.make_args([0.2, 0.5, 0.8], [800, 2000, 600, 2500, 1200, 4000])
    d = synth.generate(600, truth, 3, seed=91)
    od = oracle.Data.from_bursts(d)
    obs = oracle.run(od, oracle.make_params(n, truth, *PH), R=1, seed=5)["E"]
    rng = np.random.default_rng(3)
    pop = [synth.population(1, rng, states=(3,))[0][1] for _ in range(nsets)]
    inp = tmp_path / "in"
    inp.mkdir()
    arrays = dict(start=d.start, stop=d.stop, istart=d.istart, phcount=d.phcount,
                  burst_ad=d.burst_ad(), burst_dd=d.burst_dd(), T=d.T, SgDivSr=obs,
                  scalars=np.array([d.clk_p, d.bg_ad, d.bg_dd], np.float32),
                  pop_args=np.concatenate([np.asarray(a, np.float32) for a in pop]))
    for k, v in arrays.items():
        np.ascontiguousarray(v).tofile(str(inp / k))
    out = tmp_path / "out.txt"
    subprocess.run([host_program("ps_loop"), str(inp), str(out), str(n), str(R), str(seed),
                    str(streams), str(nsets)], check=True, capture_output=True, timeout=120)
    lines = out.read_text().splitlines()
    ohist = np.array(lines[0].split()[1:], np.float32)
    assert np.array_equal(ohist, oracle.hist(obs, 200).astype(np.float32))
    res = {}
    for ln in lines[1:]:
        f = ln.split()
        idx, sid, chi2, sent = int(f[1]), int(f[2]), np.float32(f[3]), int(f[4])
        assert idx not in res, f"set {idx} returned twice"
        res[idx] = (sid, chi2, sent, np.array(f[5:], np.float32))
    assert sorted(res) == list(range(nsets))
    assert {sid for sid, _, _, _ in res.values()} <= set(range(streams))
    best = np.float32(3.1e32)
    order = [int(ln.split()[1]) for ln in lines[1:]]      # the server's arrival order
    for idx in order:
        sid, chi2, sent, shist = res[idx]
        o = oracle.run(od, oracle.make_params(n, pop[idx], *PH), R=R, seed=seed, set_id=sid,
                       B=200, want_units=False)
        c32, _, _ = oracle.chi2(ohist, o["hist"], R, n)
        assert chi2 == np.float32(c32), (idx, chi2, c32)
        # the best-so-far the worker compares against is the server's value at
        # hand-out time, never better than the best of the results before it
        if chi2 < best:
            assert sent == 1
        if sent:
            assert np.array_equal(shist, o["hist"].astype(np.float32) / np.float32(R))
        best = min(best, chi2)
