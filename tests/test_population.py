"""Population runs: a generation of parameter sets of uneven state counts in
one graph-captured launch sequence.

Each set's integer histogram must equal the oracle's for that set's random
stream key (seed, set_base + i * set_stride) bit for bit, and its chi2 the
oracle's float value; the same holds against single-set runs through pda_run.
Graph replays are checked against direct launches with eager work in
between, the launch shape and the autotuner against the default shape.
"""
import numpy as np
import pytest

from gsmfretda_amd import pda, synth

pytestmark = pytest.mark.gpu

PH = synth.PHOTOPHYSICS


@pytest.fixture(scope="module")
def data():
    args = synth.make_args([0.2, 0.6, 0.85], [800, 3000, 500, 2500, 1200, 6000])
    d = synth.generate(700, args, 3, seed=77)
    rng = np.random.default_rng(78)
    obs = rng.uniform(0, 1, d.nbursts).astype(np.float32)
    return d, obs


def _population(count, seed, states=(2, 3, 4), v=False):
    rng = np.random.default_rng(seed)
    sets = synth.population(count, rng, states=states)
    if v:  # distance broadening on some states
        out = []
        for n, a in sets:
            a = np.array(a, np.float32)
            a[n * n:] = rng.uniform(0, 4, n) * (rng.uniform(size=n) < 0.5)
            out.append((n, a))
        sets = out
    return sets


def _load(m, d, obs, B=50):
    m.load(d, SgDivSr=obs, nbins=B, left_pad=2)
    m.set_photophysics(0, *PH)
    m.set_burst_range(0, -1, 0)


def _oracle_set(oracle, od, obs_hist, n, args, R, seed, set_id, B=50):
    """Oracle histogram and float-accumulated chi2 of one set."""
    p = oracle.make_params(n, args, *PH)
    o = oracle.run(od, p, R=R, seed=seed, set_id=set_id, B=B, want_units=False)
    c32, _, _ = oracle.chi2(obs_hist, o["hist"], R, n)
    return o["hist"], c32


def _obs_hist(oracle, obs, B=50):
    return oracle.hist(obs, B).astype(np.float32)


def test_population_matches_oracle_per_set(gpu_mc, oracle, data):
    d, obs = data
    m = gpu_mc
    _load(m, d, obs)
    sets = _population(12, 5)
    assert {n for n, _ in sets} == {2, 3, 4}
    m.set_population([n for n, _ in sets], [a for _, a in sets])
    R, seed, base, stride = 16, 901, 40, 3
    m.run_population(R=R, seed=seed, set_base=base, set_stride=stride)
    hist, c64, c32, eff = m.get_population_result()
    od = oracle.Data.from_bursts(d)
    oh = _obs_hist(oracle, obs)
    for i, (n, a) in enumerate(sets):
        h_o, c_o = _oracle_set(oracle, od, oh, n, a, R, seed, base + i * stride)
        assert np.array_equal(hist[i].astype(np.int64), h_o), f"set {i} (n={n}) histogram"
        assert c32[i] == np.float32(c_o), f"set {i} chi2_ref {c32[i]} vs {c_o}"
        # FP64 recomputation from the returned histogram
        Mb = hist[i].astype(np.float64) / R
        keep = np.arange(50) >= 2
        keep &= hist[i] > 0
        want = np.sum((oh[keep] - Mb[keep]) ** 2 / Mb[keep]) / (keep.sum() - n * (n + 1))
        assert abs(c64[i] - want) <= 1e-9 * abs(want)
        assert eff[i] == keep.sum()
    units, replays, captures, ms = m.population_stats()
    assert units == d.nbursts * R and captures >= 1 and ms > 0


def test_population_equals_single_set_runs(gpu_mc, data):
    d, obs = data
    m = gpu_mc
    _load(m, d, obs)
    sets = _population(6, 6)
    m.set_population([n for n, _ in sets], [a for _, a in sets])
    R, seed = 24, 17
    m.run_population(R=R, seed=seed, set_base=100, set_stride=1)
    hist, c64, c32, _ = m.get_population_result()
    for i, (n, a) in enumerate(sets):
        h1, s64, s32, _ = pda.evaluate(m, n, a, R, seed, set_id=100 + i)
        assert np.array_equal(hist[i], h1), f"set {i}"
        assert c64[i] == s64 and c32[i] == s32


def test_graph_replays_match_direct_launches(gpu_mc, data):
    d, obs = data
    m = gpu_mc
    _load(m, d, obs)
    sets = _population(9, 7)
    m.set_population([n for n, _ in sets], [a for _, a in sets])
    R = 12
    m.run_population(R=R, seed=5, graph=False)
    want5 = m.get_population_result()
    m.run_population(R=R, seed=6, graph=False)
    want6 = m.get_population_result()
    _, _, cap0, _ = m.population_stats()
    for rep in range(4):
        seed = 5 if rep % 2 == 0 else 6
        m.run_population(R=R, seed=seed)
        got = m.get_population_result()
        want = want5 if seed == 5 else want6
        for g, w in zip(got, want):
            assert np.array_equal(g, w), f"replay {rep}"
        # eager work between replays: a single-set run on the same stream
        pda.evaluate(m, 2, synth.make_args([0.3, 0.7], [1e3, 1e3]), 4, 99)
    _, _, cap1, _ = m.population_stats()
    assert cap1 == cap0 + 1, "seed changes must replay the captured graph"
    # new parameters of the same shape: same graph, new results
    sets2 = _population(9, 8)
    m.set_population([n for n, _ in sets2], [a for _, a in sets2])
    m.run_population(R=R, seed=5)
    g2 = m.get_population_result()
    m.run_population(R=R, seed=5, graph=False)
    w2 = m.get_population_result()
    for g, w in zip(g2, w2):
        assert np.array_equal(g, w)
    assert m.population_stats()[2] == cap1
    # a different R is a different graph
    m.run_population(R=R + 1, seed=5)
    m.get_population_result()
    assert m.population_stats()[2] == cap1 + 1


def test_graph_after_reload_reads_the_new_bursts(gpu_mc, oracle, data):
    """A captured graph holds the addresses of the resident bursts.  Loading
    other bursts of the same count (the context frees and reallocates its
    buffers, and may get an address back) must not replay the old graph:
    the result equals the oracle on the new bursts, across two reloads."""
    d, obs = data
    m = gpu_mc
    sets = _population(4, 31)
    R, seed = 8, 313
    args = synth.make_args([0.25, 0.55, 0.9], [900, 2500, 600, 2000, 1500, 5000])
    for k, ds in enumerate((77, 79, 81)):
        dk = d if ds == 77 else synth.generate(d.nbursts, args, 3, seed=ds)
        _load(m, dk, obs)
        m.set_population([n for n, _ in sets], [a for _, a in sets])
        m.run_population(R=R, seed=seed, set_base=0, set_stride=1)
        hist, _, c32, _ = m.get_population_result()
        od = oracle.Data.from_bursts(dk)
        oh = _obs_hist(oracle, obs)
        for i, (n, a) in enumerate(sets):
            h_o, c_o = _oracle_set(oracle, od, oh, n, a, R, seed, i)
            assert np.array_equal(hist[i].astype(np.int64), h_o), f"load {k} set {i}"
            assert c32[i] == np.float32(c_o), f"load {k} set {i}"
    _load(m, d, obs)


def test_common_random_numbers(gpu_mc, data):
    d, obs = data
    m = gpu_mc
    _load(m, d, obs)
    n, a = 3, synth.make_args([0.2, 0.5, 0.8], [500, 900, 2000, 300, 700, 4000])
    m.set_population([n] * 4, [a] * 4)
    m.run_population(R=8, seed=3, set_base=7, set_stride=0)
    hist, c64, _, _ = m.get_population_result()
    assert all(np.array_equal(hist[0], hist[i]) for i in range(4))
    m.run_population(R=8, seed=3, set_base=7, set_stride=1)
    hist2, _, _, _ = m.get_population_result()
    assert np.array_equal(hist2[0], hist[0]) and not np.array_equal(hist2[1], hist2[0])


def test_population_with_broadening(gpu_mc, oracle, data):
    d, obs = data
    m = gpu_mc
    _load(m, d, obs)
    sets = _population(5, 9, v=True)
    m.set_population([n for n, _ in sets], [a for _, a in sets])
    R, seed = 8, 44
    m.run_population(R=R, seed=seed, set_base=0, set_stride=1)
    hist, _, c32, _ = m.get_population_result()
    od = oracle.Data.from_bursts(d)
    oh = _obs_hist(oracle, obs)
    for i, (n, a) in enumerate(sets):
        h_o, c_o = _oracle_set(oracle, od, oh, n, a, R, seed, i)
        assert np.array_equal(hist[i].astype(np.int64), h_o), f"set {i}"
        assert c32[i] == np.float32(c_o)


def test_population_up_to_eight_states(gpu_mc, oracle, data):
    """Sets of 5 to 8 states (the largest model the tables hold) mixed with a
    2-state set in one population: every set equal to the oracle."""
    d, obs = data
    m = gpu_mc
    _load(m, d, obs)
    sets = _population(6, 13, states=(5, 6, 7, 8)) + _population(1, 14, states=(2,))
    assert {n for n, _ in sets} == {2, 5, 6, 7, 8}
    m.set_population([n for n, _ in sets], [a for _, a in sets])
    R, seed = 6, 808
    m.run_population(R=R, seed=seed, set_base=50, set_stride=2)
    hist, _, c32, _ = m.get_population_result()
    od = oracle.Data.from_bursts(d)
    oh = _obs_hist(oracle, obs)
    for i, (n, a) in enumerate(sets):
        h_o, c_o = _oracle_set(oracle, od, oh, n, a, R, seed, 50 + 2 * i)
        assert np.array_equal(hist[i].astype(np.int64), h_o), f"set {i} (n={n})"
        assert c32[i] == np.float32(c_o), f"set {i}"


def test_population_beyond_the_guide_budget(gpu_mc, data):
    """A population whose binomial threshold table would exceed its 1 GiB
    budget (8,240 states of 134,144 bytes) walks with the inversion loop; every
    set checked equals its single-set run (which draws by the table)."""
    d, obs = data
    m = gpu_mc
    _load(m, d, obs)
    sets = _population(1030, 21, states=(8,))
    assert len(sets) * 8 * 8384 * 16 > 1 << 30
    m.set_population([n for n, _ in sets], [a for _, a in sets])
    R, seed = 2, 515
    m.run_population(R=R, seed=seed, set_base=7, set_stride=1)
    hist, c64, c32, _ = m.get_population_result()
    for i in (0, 1, 514, 1029):
        n, a = sets[i]
        h1, s64, s32, _ = pda.evaluate(m, n, a, R, seed, set_id=7 + i)
        assert np.array_equal(hist[i], h1), f"set {i}"
        assert c64[i] == s64 and c32[i] == s32


def test_population_at_200_bins(gpu_mc, oracle, data):
    """The reference's default of 200 E bins: every set equal to the oracle."""
    d, obs = data
    m = gpu_mc
    _load(m, d, obs, B=200)
    sets = _population(5, 15)
    m.set_population([n for n, _ in sets], [a for _, a in sets])
    R, seed = 8, 77
    m.run_population(R=R, seed=seed, set_base=0, set_stride=1)
    hist, _, c32, _ = m.get_population_result()
    assert hist.shape == (5, 200)
    od = oracle.Data.from_bursts(d)
    oh = _obs_hist(oracle, obs, 200)
    for i, (n, a) in enumerate(sets):
        h_o, c_o = _oracle_set(oracle, od, oh, n, a, R, seed, i, B=200)
        assert np.array_equal(hist[i].astype(np.int64), h_o), f"set {i}"
        assert c32[i] == np.float32(c_o), f"set {i}"
    _load(m, d, obs)


def test_launch_shapes_and_autotune_do_not_change_results(gpu_mc, data):
    d, obs = data
    m = gpu_mc
    _load(m, d, obs)
    sets = _population(7, 10)
    m.set_population([n for n, _ in sets], [a for _, a in sets])
    m.run_population(R=10, seed=21, graph=False)
    want = m.get_population_result()
    try:
        for shape in ((1, 0, 0), (2, 1, 1), (16, 3, 8), (128, 0, 2), (0, 2, 4)):
            m.set_launch_shape(*shape)
            m.run_population(R=10, seed=21)
            got = m.get_population_result()
            for g, w in zip(got, want):
                assert np.array_equal(g, w), f"shape {shape}"
        m.set_launch_shape(0, 0, 0)
        bt, wpc, q, ms = m.autotune(population=True, R=10, reps=1)
        assert ms > 0 and m.get_launch_shape() == (bt, wpc, q)
        m.run_population(R=10, seed=21)
        got = m.get_population_result()
        for g, w in zip(got, want):
            assert np.array_equal(g, w), "after autotune"
    finally:
        m.set_launch_shape(0, 0, 0)


def test_population_errors(gpu_mc, data):
    d, obs = data
    m = gpu_mc
    _load(m, d, obs)
    with pytest.raises(pda.PdaError, match="set 1"):
        m.set_population([2, 9], [synth.make_args([0.3, 0.7], [1e3, 1e3]), np.zeros(90)])
    with pytest.raises(pda.PdaError, match="set 0"):
        m.set_population([2], [np.array([0.3, 1.5, 1e3, 1e3, 0, 0], np.float32)])
    with pytest.raises(pda.PdaError, match="need"):
        m._L.pda_set_population  # noqa: B018 (the symbol exists)
        nst = np.array([2, 3], np.int32)
        a = np.zeros(7, np.float32)
        pda._check(m._L.pda_set_population(m._h, 0, 2, nst.ctypes.data_as(pda._i32p),
                                           a.ctypes.data_as(pda._f32p), 7))


def test_full_population_workload_against_oracle(gpu_mc, oracle):
    """The batched-population workload at full size (256 sets of 2-4 states x
    50,000 bursts x 32 resamples in one launch sequence): four sets spread
    over the population equal the oracle bit for bit."""
    cfg = synth.config(2)
    d, R, pop = cfg["data"], cfg["R"], cfg["population"]
    assert len(pop) == 256 and d.nbursts == 50_000
    m = gpu_mc
    m.load(d, SgDivSr=np.zeros(d.nbursts, np.float32), nbins=50, left_pad=2)
    m.set_photophysics(0, *cfg["photophysics"])
    m.set_burst_range(0, -1, 0)
    m.set_population([n for n, _ in pop], [a for _, a in pop])
    seed = cfg["seeds"]["mc"]
    m.run_population(R=R, seed=seed, set_base=0, set_stride=1)
    hist, _, _, _ = m.get_population_result()
    od = oracle.Data.from_bursts(d)
    for i in (0, 85, 170, 255):
        n, a = pop[i]
        p = oracle.make_params(n, a, *cfg["photophysics"])
        want = oracle.run_threaded(od, p, R=R, seed=seed, set_id=i, nthreads=16)
        assert np.array_equal(hist[i].astype(np.int64), want), f"set {i} (n={n})"
