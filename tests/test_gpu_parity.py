"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Bit-exact: Philox blocks, the shared elementary functions and variates, and
every per-sample acceptor count / corrected count / E and the integer
histograms.  chi2: equal to the oracle's float-accumulated value and to a
FP64 host recomputation from the returned histogram within 1e-5 relative.
"""
import numpy as np
import pytest

from gsmfretda_amd import synth
from tests.conftest import bits

pytestmark = pytest.mark.gpu

BINOM_DT = np.dtype([("n", "<i4"), ("pad0", "<u4"), ("p", "<f8"), ("seed", "<u4"), ("set", "<u4"),
                     ("r", "<u4"), ("b", "<u4"), ("w", "<u4"), ("hi", "<u4"), ("lo", "<u4"),
                     ("pad1", "<u4")])
BG_DT = np.dtype([("lam", "<f4"), ("cap", "<f4"), ("u1", "<f4"), ("u2", "<f4"), ("seed", "<u4"),
                  ("set", "<u4"), ("r", "<u4"), ("b", "<u4"), ("w", "<u4"), ("sub", "<u4")])


def _gpu_run(m, cfg_n, args, R, seed, set_id=0, begin=0, end=-1, ph=None, sid=0):
    if ph is not None:
        m.set_photophysics(sid, *ph)
    assert m.set_params(cfg_n, sid, args)
    n_b = m.set_burst_range(begin, end, sid)
    m.reSampleTimes = R
    m.run_kernel(sid=sid, seed=seed, set_id=set_id, samples=True)
    m.synchronize(sid)
    hist, c64, c32, eff = m.get_hist_chi2(sid)
    E, acc, F = m.get_samples(sid, n_b * R)
    return dict(hist=hist, chi2_64=c64, chi2_ref=c32, effN=eff, E=E, acc=acc, F=F)


def _compare_units(g, o):
    assert np.array_equal(g["acc"], o["acc"]), f"acc differs in {(g['acc'] != o['acc']).sum()} units"
    assert np.array_equal(bits(g["F"]), bits(o["F"])), "F differs"
    assert np.array_equal(bits(g["E"]), bits(o["E"])), "E differs"
    assert np.array_equal(g["hist"].astype(np.int64), o["hist"]), "histogram differs"


def test_primitives_bit_exact(gpu_mc, oracle):
    rng = np.random.default_rng(11)
    m = gpu_mc
    # Philox4x32-10
    q = rng.integers(0, 2**32, size=(4096, 6), dtype=np.uint64).astype(np.uint32)
    q[0] = 0
    q[1] = 0xFFFFFFFF
    got = m.debug_primitive(0, q, np.uint32, len(q)).reshape(-1, 4)
    want = np.array([oracle.philox(r[:4], r[4:]) for r in q])
    assert np.array_equal(got, want)
    # elementary functions
    xf = np.concatenate([rng.uniform(0, 1, 20000), np.exp(rng.uniform(-87, 88, 20000)),
                         [0.0, 1.0, np.inf, 1e-40, 3e-39, 0.5, 2.0]]).astype(np.float32)
    L = oracle.lib()
    for op, fn, x in ((1, L.or_logf, xf), (2, L.or_expf, rng.uniform(-110, 95, 20000).astype(np.float32)),
                      (5, L.or_cos2pif, rng.uniform(0, 1, 20000).astype(np.float32))):
        got = m.debug_primitive(op, x, np.float32, len(x))
        want = np.array([fn(float(v)) for v in x], np.float32)
        assert np.array_equal(bits(got), bits(want)), f"op {op}"
    xd = np.concatenate([rng.uniform(0, 1, 20000), np.exp(rng.uniform(-700, 700, 20000)),
                         [0.0, 1.0, 1e-310, np.inf]])
    got = m.debug_primitive(3, xd, np.float64, len(xd))
    want = np.array([L.or_log(float(v)) for v in xd])
    assert np.array_equal(bits(got), bits(want))
    xe = rng.uniform(-750, 712, 20000)
    got = m.debug_primitive(4, xe, np.float64, len(xe))
    want = np.array([L.or_exp(float(v)) for v in xe])
    assert np.array_equal(bits(got), bits(want))
    g2 = rng.integers(0, 2**32, size=(20000, 2), dtype=np.uint64).astype(np.uint32)
    got = m.debug_primitive(6, g2, np.float32, len(g2))
    want = np.array([L.or_normal(int(a), int(b)) for a, b in g2], np.float32)
    assert np.array_equal(bits(got), bits(want))


def test_variates_bit_exact(gpu_mc, oracle):
    rng = np.random.default_rng(12)
    L = oracle.lib()
    k = 6000
    it = np.zeros(k, BINOM_DT)
    it["n"] = np.concatenate([rng.integers(0, 40, k // 3), rng.integers(60, 2000, k // 3),
                              rng.integers(0, 400, k - 2 * (k // 3))])
    it["p"] = np.concatenate([rng.uniform(0, 1, k - 6), [0.0, 1.0, 0.5, 0.5, 1e-9, 1 - 1e-9]])
    for f in ("seed", "set", "r", "b", "w", "hi", "lo"):
        it[f] = rng.integers(0, 2**32, k, dtype=np.uint64).astype(np.uint32)
    got = gpu_mc.debug_primitive(7, it, np.int64, k)
    want = np.array([L.or_binomial(int(x["n"]), float(x["p"]), int(x["seed"]), int(x["set"]),
                                   int(x["b"]), int(x["r"]), int(x["w"]), int(x["hi"]),
                                   int(x["lo"]), None) for x in it])
    assert np.array_equal(got, want)
    bg = np.zeros(k, BG_DT)
    bg["lam"] = np.concatenate([rng.uniform(0, 3, k // 2), rng.uniform(50, 600, k // 4),
                                [0.0] * (k - k // 2 - k // 4)])
    bg["cap"] = np.concatenate([rng.uniform(0, 8, k // 2), rng.uniform(0, 700, k - k // 2)])
    bg["u1"] = rng.uniform(1e-7, 1, k)
    bg["u2"] = rng.uniform(1e-7, 1, k)
    for f in ("seed", "set", "r", "b", "w"):
        bg[f] = rng.integers(0, 2**32, k, dtype=np.uint64).astype(np.uint32)
    bg["sub"] = 0x100000
    got = gpu_mc.debug_primitive(8, bg, np.float32, k)
    want = np.array([L.or_bg_capped(float(x["lam"]), float(x["cap"]), float(x["u1"]),
                                    float(x["u2"]), int(x["seed"]), int(x["set"]), int(x["b"]),
                                    int(x["r"]), int(x["w"]), int(x["sub"]), None) for x in bg],
                    np.float32)
    assert np.array_equal(bits(got), bits(want))


def _oracle_run(oracle, data, n, args, ph, R, seed, set_id=0, b0=0, b1=None, B=50):
    od = oracle.Data.from_bursts(data)
    p = oracle.make_params(n, args, *ph)
    return oracle.run(od, p, R=R, seed=seed, set_id=set_id, b0=b0, b1=b1, B=B)


def _load(m, data, obs_E, B=50):
    m.load(data, SgDivSr=obs_E, nbins=B, left_pad=2)


def test_c0_full_bit_exact(gpu_mc, oracle):
    cfg = synth.config(0)
    d = cfg["data"]
    obs = _oracle_run(oracle, d, cfg["n"], cfg["args"], cfg["photophysics"], 1, 2000)["E"]
    _load(gpu_mc, d, obs)
    g = _gpu_run(gpu_mc, cfg["n"], cfg["args"], cfg["R"], 3000, ph=cfg["photophysics"])
    o = _oracle_run(oracle, d, cfg["n"], cfg["args"], cfg["photophysics"], cfg["R"], 3000)
    _compare_units(g, o)
    O = oracle.hist(obs).astype(np.float32)
    c32, c64, eff = oracle.chi2(O, o["hist"], cfg["R"], cfg["n"])
    assert g["effN"] == eff
    assert np.float32(g["chi2_ref"]).view(np.uint32) == np.float32(c32).view(np.uint32)
    assert abs(g["chi2_64"] - c64) <= 1e-5 * abs(c64)


def test_c1_subset_bit_exact(gpu_mc, oracle):
    cfg = synth.config(1, n_bursts=3000)
    d = cfg["data"]
    obs = _oracle_run(oracle, d, cfg["n"], cfg["args"], cfg["photophysics"], 1, 2001)["E"]
    _load(gpu_mc, d, obs)
    for seed, set_id in ((3001, 0), (17, 5)):
        g = _gpu_run(gpu_mc, cfg["n"], cfg["args"], cfg["R"], seed, set_id=set_id,
                     ph=cfg["photophysics"])
        o = _oracle_run(oracle, d, cfg["n"], cfg["args"], cfg["photophysics"], cfg["R"], seed,
                        set_id=set_id)
        _compare_units(g, o)
        # chi2 in FP64 recomputed on the host from the returned histogram
        O = oracle.hist(obs).astype(np.float64)
        M = g["hist"].astype(np.float64) / cfg["R"]
        keep = (np.arange(50) >= 2) & (g["hist"] > 0)
        dof = keep.sum() - cfg["n"] * (cfg["n"] + 1)
        c64 = float(np.sum((O[keep] - M[keep]) ** 2 / M[keep]) / dof)
        assert abs(g["chi2_64"] - c64) <= 1e-5 * abs(c64)


def test_ranges_neither_photons_and_no_background(gpu_mc, oracle):
    cfg = synth.config(1, n_bursts=1)
    d = synth.generate(1500, cfg["args"], 3, seed=91, p_neither=0.25)
    d.mask_ad[::97] = 1  # photons in both channels
    d.mask_dd[::97] = 1
    d.phcount[10] = 0    # empty and single-photon bursts
    d.phcount[11] = 1
    d.T[12] = 0.0
    d.bg_dd = 0.0        # reference's no-background branch (bg_dd < 1e-4)
    obs = np.linspace(0, 1, 1500, endpoint=False).astype(np.float32)
    _load(gpu_mc, d, obs)
    for begin, end, R in ((0, -1, 7), (100, 900, 33), (1499, 1500, 64)):
        g = _gpu_run(gpu_mc, 3, cfg["args"], R, 99, set_id=3, begin=begin, end=end,
                     ph=(0.34, 0.08, 0.07, 52.0))
        o = _oracle_run(oracle, d, 3, cfg["args"], (0.34, 0.08, 0.07, 52.0), R, 99, set_id=3,
                        b0=begin, b1=(d.nbursts if end < 0 else end))
        _compare_units(g, o)


def test_empty_burst_range(gpu_mc, oracle):
    """An empty [begin, end) runs nothing: a zero histogram and no bins kept;
    the next run on the slot equals the oracle."""
    cfg = synth.config(0, n_bursts=300)
    d = cfg["data"]
    _load(gpu_mc, d, np.zeros(300, np.float32))
    gpu_mc.set_photophysics(0, *cfg["photophysics"])
    assert gpu_mc.set_params(cfg["n"], 0, cfg["args"])
    assert gpu_mc.set_burst_range(120, 120, 0) == 0
    gpu_mc.reSampleTimes = 9
    gpu_mc.run_kernel(sid=0, seed=5, set_id=0)
    gpu_mc.synchronize(0)
    hist, _, _, eff = gpu_mc.get_hist_chi2(0)
    assert not hist.any() and eff == 0
    g = _gpu_run(gpu_mc, cfg["n"], cfg["args"], 9, 5, ph=cfg["photophysics"])
    o = _oracle_run(oracle, d, cfg["n"], cfg["args"], cfg["photophysics"], 9, 5)
    _compare_units(g, o)


def test_cumulative_window_ends_closed_form(gpu_mc):
    """The HIP walk on 20,000 evenly spaced single-colour bursts x 100: the
    mean acceptor count equals the closed form of the reference's cumulative
    window ends (src/mc.cu:132, 184) and rejects window ends at burst start +
    dwell sum (tests/test_oracle_walk.py derives both)."""
    from tests.test_oracle_walk import (WINDOW_ARGS, WINDOW_P, WINDOW_PH, window_end_data,
                                        window_end_prediction)
    d = window_end_data(20000)
    _load(gpu_mc, d, np.zeros(d.nbursts, np.float32))
    g = _gpu_run(gpu_mc, 2, WINDOW_ARGS, 100, 12, ph=WINDOW_PH)
    assert np.all(g["F"] == WINDOW_P - 1)
    acc = g["acc"].astype(np.float64)
    se = acc.std(ddof=1) / np.sqrt(acc.size)
    want = window_end_prediction(cumulative=True)
    wrong = window_end_prediction(cumulative=False, kmax=61)
    assert abs(acc.mean() - want) < 4.5 * se, (acc.mean(), want, se)
    assert abs(acc.mean() - wrong) > 50 * se, (acc.mean(), wrong, se)


def test_photophysics_giving_negative_counts_are_refused(gpu_mc):
    """gamma < crosstalk or DexDirAem > 1 would make the corrected photon count
    (the binomial's trial count, src/mc.cu:143) negative: PDA_E_ARG, and the
    slot keeps its constants.  The boundary cases gamma == crosstalk and
    DexDirAem == 1 (a channel that counts nothing) are accepted."""
    from gsmfretda_amd import pda
    for ph in ((0.05, 0.08, 0.07, 52.0), (0.34, 1.01, 0.07, 52.0)):
        with pytest.raises(pda.PdaError, match="negative"):
            gpu_mc.set_photophysics(1, *ph)
    gpu_mc.set_photophysics(1, 0.07, 1.0, 0.07, 52.0)
    gpu_mc.set_photophysics(1, *synth.PHOTOPHYSICS)


def test_broadening_and_four_states(gpu_mc, oracle):
    args = synth.make_args([0.1, 0.4, 0.6, 0.95], np.linspace(200, 3000, 12), v=[2.0, 0.0, 5.0, 1.0])
    d = synth.generate(800, args, 4, seed=5)
    obs = np.zeros(800, np.float32)
    _load(gpu_mc, d, obs)
    g = _gpu_run(gpu_mc, 4, args, 40, 7, ph=(1.0, 0.0, 0.05, 52.0))
    o = _oracle_run(oracle, d, 4, args, (1.0, 0.0, 0.05, 52.0), 40, 7)
    _compare_units(g, o)


def test_c1_full_histogram(gpu_mc, oracle):
    """Full C1 size: histogram identical to the oracle's (threaded), and runs repeat."""
    cfg = synth.config(1)
    d = cfg["data"]
    od = oracle.Data.from_bursts(d)
    p = oracle.make_params(cfg["n"], cfg["args"], *cfg["photophysics"])
    obs = oracle.run(od, p, R=1, seed=2001, want_units=True)["E"]
    _load(gpu_mc, d, obs)
    assert gpu_mc.set_params(cfg["n"], 0, cfg["args"])
    gpu_mc.set_photophysics(0, *cfg["photophysics"])
    gpu_mc.set_burst_range(0, -1, 0)
    gpu_mc.reSampleTimes = cfg["R"]
    hists = []
    for _ in range(2):
        gpu_mc.run_kernel(sid=0, seed=3001, set_id=0)
        gpu_mc.synchronize(0)
        hists.append(gpu_mc.get_hist_chi2(0))
    assert np.array_equal(hists[0][0], hists[1][0])
    assert hists[0][1] == hists[1][1]
    want = oracle.run_threaded(od, p, R=cfg["R"], seed=3001, nthreads=16)
    assert np.array_equal(hists[0][0].astype(np.int64), want)


def test_replayed_units(gpu_mc, oracle):
    """Units the fast pass abandons (BTPE draws, background means > 64) are
    replayed by the general pass with identical results."""
    args = synth.make_args([0.05, 0.2, 0.35, 0.5, 0.65, 0.8, 0.9, 0.97], np.full(56, 150.0),
                           np.zeros(8))
    d = synth.generate(200, synth.make_args([0.97, 0.97], [10.0, 10.0]), 2, base_photons=300,
                       mean_geom=200, seed=3)
    _load(gpu_mc, d, np.zeros(200, np.float32))
    g = _gpu_run(gpu_mc, 8, args, 12, 5, ph=(1.0, 0.0, 0.0, 52.0))
    o = _oracle_run(oracle, d, 8, args, (1.0, 0.0, 0.0, 52.0), 12, 5)
    _compare_units(g, o)
    units, replays = gpu_mc.run_stats(0)
    assert units == 200 * 12 and replays > 0, (units, replays)
    from tests.test_oracle_walk import tiny_bursts
    d = tiny_bursts(np.random.default_rng(52), 60)
    d.bg_dd, d.bg_ad = 1e6, 3e6
    _load(gpu_mc, d, np.zeros(60, np.float32))
    args = synth.make_args([0.1, 0.3, 0.6, 0.9], np.full(12, 1e4))
    for R in (9, 70):
        g = _gpu_run(gpu_mc, 4, args, R, 1002, ph=(0.34, 0.08, 0.07, 52.0))
        o = _oracle_run(oracle, d, 4, args, (0.34, 0.08, 0.07, 52.0), R, 1002)
        _compare_units(g, o)
        assert gpu_mc.run_stats(0)[1] > 0


def test_many_replayed_units(gpu_mc, oracle):
    """Workloads where a few per cent of units leave the fast form and go to
    the grid-wide replay pass: the 3-state workload's bursts under a 30 kHz
    donor / 15 kHz acceptor background (long windows have background means
    above 64), and long bursts (150 + Geom(100) photons) of a near-1
    efficiency (windows with 128 or more acceptor photons in one run).  Every
    unit equals the oracle's, twice in a row (the finish kernel leaves the
    replay counters zero)."""
    cfg = synth.config(1, n_bursts=1)
    d1 = synth.generate(2000, cfg["args"], 3, seed=93)
    d1.bg_dd, d1.bg_ad = 3e4, 1.5e4
    d2 = synth.generate(2000, synth.make_args([0.98, 0.98], [200.0, 200.0]), 2, base_photons=150,
                        mean_geom=100.0, seed=94)
    for d in (d1, d2):
        _load(gpu_mc, d, np.zeros(d.nbursts, np.float32))
        o = _oracle_run(oracle, d, 3, cfg["args"], cfg["photophysics"], 20, 44)
        for _ in range(2):
            g = _gpu_run(gpu_mc, 3, cfg["args"], 20, 44, ph=cfg["photophysics"])
            _compare_units(g, o)
            units, replays = gpu_mc.run_stats(0)
            assert units == d.nbursts * 20 and replays >= units // 100, (units, replays)


def test_launch_shape_does_not_change_results(gpu_mc, oracle):
    cfg = synth.config(0, n_bursts=700)
    d = cfg["data"]
    _load(gpu_mc, d, np.zeros(700, np.float32))
    o = _oracle_run(oracle, d, cfg["n"], cfg["args"], cfg["photophysics"], 13, 21)
    saved = gpu_mc.get_launch_shape()
    try:
        for tiles, waves in ((1, 0), (3, 1), (16, 2), (64, 5), (0, 3)):
            gpu_mc.set_launch_shape(batch_tiles=tiles, waves_per_cu=waves)
            g = _gpu_run(gpu_mc, cfg["n"], cfg["args"], 13, 21, ph=cfg["photophysics"])
            _compare_units(g, o)
    finally:
        gpu_mc.set_launch_shape(*saved)


def test_work_counters_do_not_change_results(gpu_mc, oracle):
    """Every queue count up to 64 (each counter on its own line) and batch
    size gives the oracle's units, twice in a row (the finish kernel leaves
    all counters zero for the next run)."""
    cfg = synth.config(0, n_bursts=900)
    d = cfg["data"]
    _load(gpu_mc, d, np.zeros(900, np.float32))
    o = _oracle_run(oracle, d, cfg["n"], cfg["args"], cfg["photophysics"], 11, 23)
    saved = gpu_mc.get_launch_shape()
    try:
        for tiles, q in ((1, 64), (2, 3), (4, 32), (0, 1), (16, 64), (0, 0)):
            gpu_mc.set_launch_shape(batch_tiles=tiles, queues=q)
            for _ in range(2):
                g = _gpu_run(gpu_mc, cfg["n"], cfg["args"], 11, 23, ph=cfg["photophysics"])
                _compare_units(g, o)
    finally:
        gpu_mc.set_launch_shape(*saved)


def test_stream_slots_in_flight(gpu_mc, oracle):
    """Both stream slots running at once, each with its own parameter set,
    burst range, resample count and seed: every unit equals the oracle's,
    on two rounds in a row."""
    cfg = synth.config(0, n_bursts=1200)
    d = cfg["data"]
    _load(gpu_mc, d, np.zeros(1200, np.float32))
    ph = cfg["photophysics"]
    args3 = synth.make_args([0.15, 0.5, 0.85], [300.0, 2000.0, 800.0, 5000.0, 150.0, 1200.0])
    runs = [(cfg["n"], cfg["args"], 0, 1200, 9, 31, 4), (3, args3, 200, 1100, 14, 32, 5)]
    for sid, (n, args, b0, b1, R, seed, set_id) in enumerate(runs):
        gpu_mc.set_photophysics(sid, *ph)
        assert gpu_mc.set_params(n, sid, args)
        assert gpu_mc.set_burst_range(b0, b1, sid) == b1 - b0
    want = [_oracle_run(oracle, d, n, args, ph, R, seed, set_id=set_id, b0=b0, b1=b1)
            for n, args, b0, b1, R, seed, set_id in runs]
    for _ in range(2):
        for sid, (n, args, b0, b1, R, seed, set_id) in enumerate(runs):
            gpu_mc.reSampleTimes = R
            gpu_mc.run_kernel(sid=sid, seed=seed, set_id=set_id, samples=True)
        for sid, (n, args, b0, b1, R, seed, set_id) in enumerate(runs):
            gpu_mc.synchronize(sid)
            hist = gpu_mc.get_hist_chi2(sid)[0]
            E, acc, F = gpu_mc.get_samples(sid, (b1 - b0) * R)
            _compare_units(dict(hist=hist, E=E, acc=acc, F=F), want[sid])


def test_mc_shim_dropin(oracle, tmp_path):
    """The reference's call sequence through the C++ `class mc` shim
    (tests/host/mc_dropin.cpp): hmcE bit-exact with the oracle, and the
    device histogram / chi2 equal to the caller's own binning of hmcE."""
    import subprocess
    from tests.conftest import host_program
    cfg = synth.config(1, n_bursts=1200)
    d = cfg["data"]
    obs = _oracle_run(oracle, d, cfg["n"], cfg["args"], (0.34, 0.08, 0.07, 52.0), 1, 5)["E"]
    inp = tmp_path / "in"
    inp.mkdir()
    arrays = dict(start=d.start, stop=d.stop, istart=d.istart, phcount=d.phcount,
                  burst_ad=d.burst_ad(), burst_dd=d.burst_dd(), T=d.T, SgDivSr=obs,
                  scalars=np.array([d.clk_p, d.bg_ad, d.bg_dd], np.float32), args=cfg["args"])
    for k, v in arrays.items():
        np.ascontiguousarray(v).tofile(str(inp / k))
    exe = host_program("mc_dropin")
    R, seed, sid, begin, end = 16, 777, 1, 100, 1100
    out = str(tmp_path / "out.bin")
    subprocess.run([exe, str(inp), out, str(cfg["n"]), str(R), str(seed), str(begin), str(end),
                    str(sid)], check=True, capture_output=True, timeout=120)
    raw = open(out, "rb").read()
    nE = (end - begin) * R
    E = np.frombuffer(raw[:4 * nE], np.float32)
    hist = np.frombuffer(raw[4 * nE:4 * nE + 800], np.int32)
    c64, = np.frombuffer(raw[4 * nE + 800:4 * nE + 808], np.float64)
    c32, chisqr = np.frombuffer(raw[4 * nE + 808:], np.float32)
    # common random numbers: set id = the stream id
    o = _oracle_run(oracle, d, cfg["n"], cfg["args"], (0.34, 0.08, 0.07, 52.0), R, seed,
                    set_id=sid, b0=begin, b1=end, B=200)
    assert np.array_equal(bits(E), bits(o["E"]))
    assert np.array_equal(hist.astype(np.int64), o["hist"])
    assert c32 == chisqr                      # same float accumulation as streamWorker
    assert abs(c64 - chisqr) <= 1e-5 * abs(c64)
