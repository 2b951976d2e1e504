"""GPU: the Rao-Blackwellised (RB) forward, the adjoint kernel and the torch
autograd Function, through the C ABI.

* Forward: the 2^-32 fixed-point expected histogram is bit-exact against the
  oracle's float path (2-4 states, background on and off, burst ranges), and
  it is the conditional expectation of the plain simulator's histogram given
  the same walks (the plain run with the same keys differs from it by
  binomial noise only; a 5 % efficiency error is rejected).
* Adjoint: equal to the oracle's FP64 gradient for the same cotangent within
  1e-3 relative; the efficiency gradient equals central finite differences
  of the GPU forward (h = 1e-4 relative) within 1e-3; the rate gradient
  agrees with central differences over independent seeds.
* Jacobian: g @ J equals the adjoint for random cotangents (units whose
  pmf comes from the per-count-vector memo, and units with acceptor totals
  of 48 or more, built one wave per unit), and its rows equal the oracle's
  FP64 gradient for indicator cotangents within 1e-3.
* torch: the autograd Function's chi2 and gradient match the library's.
"""
import numpy as np
import pytest

from gsmfretda_amd import pda, synth

pytestmark = pytest.mark.gpu

PH = synth.PHOTOPHYSICS
ARGS3 = synth.make_args([0.25, 0.5, 0.8], [800, 2000, 600, 2500, 1200, 4000])


def _data(n_bursts, seed, bg=True):
    args = synth.make_args([0.2, 0.55, 0.85], [900, 2500, 700, 3000, 1500, 5000])
    d = synth.generate(n_bursts, args, 3, seed=seed)
    if not bg:
        d.bg_ad = d.bg_dd = 0.0
    return d


def _load(m, d, obs=None, B=50):
    if obs is None:
        obs = np.random.default_rng(1).uniform(0, 1, d.nbursts).astype(np.float32)
    m.load(d, SgDivSr=obs, nbins=B, left_pad=2)
    m.set_photophysics(0, *PH)
    m.set_burst_range(0, -1, 0)
    return obs


def _obs_hist(obs, B=50):
    """The library's binning of SgDivSr: floor(double(x) * B) on [0, 1)."""
    x = obs[(obs >= 0) & (obs < 1)].astype(np.float64)
    return np.bincount(np.floor(x * B).astype(np.int64), minlength=B)[:B].astype(np.float64)


def _rb(m, n, args, R, seed, set_id=0, scores=True):
    assert m.set_params(n, 0, np.asarray(args, np.float32))
    m.rb_forward(R=R, seed=seed, set_id=set_id, scores=scores)
    return m.rb_result()


@pytest.mark.parametrize("n,bg,rng_", [(2, True, (0, -1)), (3, True, (5, 170)), (3, False, (0, -1)),
                                       (4, True, (0, -1)), (3, True, (10, 13))])
def test_rb_forward_bit_exact(gpu_mc, oracle, n, bg, rng_):
    d = _data(200, 31 + n, bg)
    _load(gpu_mc, d)
    rng = np.random.default_rng(n)
    if n == 3:
        args = ARGS3
    else:
        E = np.sort(rng.uniform(0.05, 0.95, n))
        args = synth.make_args(E, synth.loguniform_rates(rng, n * (n - 1), 3e2, 8e3))
    gpu_mc.set_burst_range(rng_[0], rng_[1], 0)
    M, fixed, c2, eff = _rb(gpu_mc, n, args, 12, 77, set_id=3)
    od = oracle.Data.from_bursts(d)
    p = oracle.make_params(n, np.asarray(args, np.float32), *PH)
    b1 = d.nbursts if rng_[1] < 0 else rng_[1]
    want = oracle.rb_run(od, p, 12, 77, 3, b0=rng_[0], b1=b1)
    assert np.array_equal(fixed, want), np.nonzero(fixed != want)
    gpu_mc.set_burst_range(0, -1, 0)


@pytest.mark.parametrize("base,geom", [(200, 60), (400, 150)])
def test_rb_forward_bit_exact_long_bursts(gpu_mc, oracle, base, geom):
    """Long bursts: a window counts its first colour run only, so even bursts
    of 400 photons keep most acceptor totals below 48; each of these data sets
    lists one unit for rb_big_kernel (test_rb_units_of_every_size covers all
    three kernels' units in numbers)."""
    d = synth.generate(120, synth.make_args([0.2, 0.55, 0.85], [900, 2500, 700, 3000, 1500, 5000]),
                       3, base_photons=base, mean_geom=geom, seed=61)
    _load(gpu_mc, d)
    M, fixed, c2, eff = _rb(gpu_mc, 3, ARGS3, 6, 91, set_id=1)
    od = oracle.Data.from_bursts(d)
    p = oracle.make_params(3, np.asarray(ARGS3, np.float32), *PH)
    want = oracle.rb_run(od, p, 6, 91, 1)
    assert np.array_equal(fixed, want), np.nonzero(fixed != want)


def test_rb_forward_full_workload_bit_exact(gpu_mc, oracle):
    """The Rao-Blackwellised workload at full size (3 states, 100,000 bursts
    x 64 resamples, with scores): the fixed-point histogram equals the
    oracle's."""
    cfg = synth.config(3)
    d = cfg["data"]
    gpu_mc.load(d, SgDivSr=np.zeros(d.nbursts, np.float32), nbins=50, left_pad=2)
    gpu_mc.set_photophysics(0, *cfg["photophysics"])
    gpu_mc.set_burst_range(0, -1, 0)
    M, fixed, c2, eff = _rb(gpu_mc, cfg["n"], cfg["args"], cfg["R"], 3003)
    od = oracle.Data.from_bursts(d)
    p = oracle.make_params(cfg["n"], np.asarray(cfg["args"], np.float32), *cfg["photophysics"])
    want = oracle.rb_run(od, p, cfg["R"], 3003, 0)
    assert np.array_equal(fixed, want), np.nonzero(fixed != want)


def test_rb_forward_does_not_depend_on_the_launch_shape(gpu_mc):
    d = _data(300, 37)
    _load(gpu_mc, d)
    saved = gpu_mc.get_launch_shape()
    try:
        got = []
        for tiles, q in ((0, 0), (1, 64), (3, 5), (16, 1)):
            gpu_mc.set_launch_shape(batch_tiles=tiles, queues=q)
            got.append(_rb(gpu_mc, 3, ARGS3, 10, 91, set_id=1)[1])
        for g in got[1:]:
            assert np.array_equal(g, got[0])
    finally:
        gpu_mc.set_launch_shape(*saved)


def test_rb_is_the_expectation_of_the_plain_histogram(gpu_mc):
    d = _data(3000, 41)
    _load(gpu_mc, d)
    R, seed = 64, 5
    M, _, _, _ = _rb(gpu_mc, 3, ARGS3, R, seed, set_id=0)
    H, _, _, _ = pda.evaluate(gpu_mc, 3, ARGS3, R, seed, set_id=0)
    keep = M > 5
    stat = np.sum((H[keep] - M[keep]) ** 2 / M[keep])
    # given the walks the plain histogram is a sum of independent per-unit
    # outcomes with mean M: E[stat] <= bins kept
    assert stat < keep.sum() + 6 * np.sqrt(2 * keep.sum()), stat
    bad = np.array(ARGS3, np.float32)
    bad[:3] *= 1.05
    Mb, _, _, _ = _rb(gpu_mc, 3, bad, R, seed, set_id=0)
    stat_b = np.sum((H[keep] - Mb[keep]) ** 2 / np.maximum(Mb[keep], 1e-9))
    assert stat_b > 5 * (keep.sum() + 6 * np.sqrt(2 * keep.sum())), stat_b


def _chi2_cotangent(M, obs_hist, R, n, left_pad=2):
    """d chi2 / d M_b of the reduced chi2 on expected counts."""
    m = M / R
    keep = (np.arange(len(M)) >= left_pad) & (M > 0)
    dof = keep.sum() - n * (n + 1)
    g = np.zeros_like(M)
    g[keep] = (1.0 - obs_hist[keep] ** 2 / m[keep] ** 2) / (R * dof)
    return g


def test_rb_adjoint_matches_fp64_oracle(gpu_mc, oracle):
    d = _data(400, 51)
    obs = _load(gpu_mc, d)
    R, seed = 16, 8
    M, _, _, _ = _rb(gpu_mc, 3, ARGS3, R, seed, set_id=1)
    oh = _obs_hist(obs)
    g = _chi2_cotangent(M, oh, R, 3)
    grad = gpu_mc.rb_adjoint(g, 3)
    od = oracle.Data.from_bursts(d)
    Mo, go = oracle.rb_grad(od, oracle.make_params(3, ARGS3, *PH), R, seed, g, set_id=1)
    assert np.allclose(M, Mo, rtol=1e-5, atol=1e-6)
    scale = np.abs(go).max()
    assert np.all(np.abs(grad[:9] - go) <= 1e-3 * np.maximum(np.abs(go), 1e-3 * scale)), (grad, go)
    assert np.all(grad[9:] == 0.0)


def test_rb_units_beyond_the_fast_walk(gpu_mc, oracle):
    """A 30 kHz donor / 15 kHz acceptor background: windows whose background
    means exceed 64 leave the RB walk's fast form, and the general-form replay
    pass walks those units again.  The forward equals the oracle's bit for
    bit, the adjoint its FP64 gradient, and the plain walk on the same units
    abandons at least 1 % of them (the same windows)."""
    d = _data(400, 57)
    d.bg_dd, d.bg_ad = 3e4, 1.5e4
    obs = _load(gpu_mc, d)
    R, seed = 16, 12
    M, fixed, _, _ = _rb(gpu_mc, 3, ARGS3, R, seed, set_id=2)
    od = oracle.Data.from_bursts(d)
    p = oracle.make_params(3, np.asarray(ARGS3, np.float32), *PH)
    want = oracle.rb_run(od, p, R, seed, 2)
    assert np.array_equal(fixed, want), np.nonzero(fixed != want)
    oh = _obs_hist(obs)
    g = _chi2_cotangent(M, oh, R, 3)
    grad = gpu_mc.rb_adjoint(g, 3)
    Mo, go = oracle.rb_grad(od, p, R, seed, g, set_id=2)
    # the rate components carry the replayed units' scores: per component;
    # the efficiency components come from float leave-one-out pmf
    # differences, and under this chi2's cotangent (~1e9) the middle state's
    # cancels to 5 % of the largest: against the largest efficiency component
    rate_scale = np.abs(go[3:9]).max()
    assert np.all(np.abs(grad[3:9] - go[3:9]) <= 1e-3 * np.maximum(np.abs(go[3:9]), 1e-3 * rate_scale)), (grad, go)
    assert np.all(np.abs(grad[:3] - go[:3]) <= 1e-3 * np.abs(go[:3]).max()), (grad, go)
    gpu_mc.reSampleTimes = R
    gpu_mc.run_kernel(sid=0, seed=seed, set_id=2)
    gpu_mc.synchronize(0)
    units, replays = gpu_mc.run_stats(0)
    assert units == d.nbursts * R and replays >= units // 100, (units, replays)


def test_rb_at_200_bins(gpu_mc, oracle):
    """The reference's default of 200 E bins: forward bit-exact, adjoint
    within 1e-3 of the FP64 oracle, Jacobian equal to the adjoint."""
    d = _data(300, 71)
    obs = _load(gpu_mc, d, B=200)
    R, seed = 8, 12
    M, fixed, _, _ = _rb(gpu_mc, 3, ARGS3, R, seed, set_id=4)
    assert len(M) == 200
    od = oracle.Data.from_bursts(d)
    p = oracle.make_params(3, np.asarray(ARGS3, np.float32), *PH)
    assert np.array_equal(fixed, oracle.rb_run(od, p, R, seed, 4, B=200))
    g = _chi2_cotangent(M, _obs_hist(obs, 200), R, 3)
    grad = gpu_mc.rb_adjoint(g, 3)
    _, go = oracle.rb_grad(od, p, R, seed, g, set_id=4)
    scale = np.abs(go).max()
    assert np.all(np.abs(grad[:9] - go) <= 1e-3 * np.maximum(np.abs(go), 1e-3 * scale)), (grad, go)
    J = gpu_mc.rb_jacobian(3, 200)
    assert J.shape == (200, 12)
    assert np.allclose(g @ J[:, :9], grad[:9], rtol=1e-5, atol=1e-5 * np.abs(grad).max())
    _load(gpu_mc, d)


@pytest.mark.parametrize("long_bursts", [False, True])
def test_rb_jacobian_is_the_adjoint_of_every_cotangent(gpu_mc, oracle, long_bursts):
    # long bursts: many units with acceptor totals >= 48 (one wave per unit)
    if long_bursts:
        d = synth.generate(60, synth.make_args([0.2, 0.55, 0.85], [900, 2500, 700, 3000, 1500, 5000]),
                           3, base_photons=400, mean_geom=150, seed=57)
    else:
        d = _data(400, 53)
    _load(gpu_mc, d)
    R, seed = 8, 21
    M, _, _, _ = _rb(gpu_mc, 3, ARGS3, R, seed, set_id=2)
    J = gpu_mc.rb_jacobian(3, len(M))
    assert J.shape == (50, 12) and np.all(J[:, 9:] == 0.0)
    rng = np.random.default_rng(5)
    for _ in range(3):
        g = rng.normal(size=50)
        adj = gpu_mc.rb_adjoint(g, 3)
        scale = np.abs(adj).max()
        # the adjoint rounds cotangent differences to float
        assert np.all(np.abs(g @ J - adj) <= 1e-5 * np.maximum(np.abs(adj), 1e-2 * scale)), (g @ J, adj)
    # rows against the oracle's FP64 forward-mode derivative (indicator cotangents)
    od = oracle.Data.from_bursts(d)
    p = oracle.make_params(3, ARGS3, *PH)
    for b in np.argsort(-M)[:4]:
        e = np.zeros(50)
        e[b] = 1.0
        _, go = oracle.rb_grad(od, p, R, seed, e, set_id=2)
        sc = np.abs(go).max()
        assert np.all(np.abs(J[b, :9] - go) <= 1e-3 * np.maximum(np.abs(go), 1e-3 * sc)), (b, J[b], go)


def test_rb_units_of_every_size(gpu_mc, oracle):
    """Efficiencies near 0 and 1 make a window's colour run span most of its
    photons, so acceptor totals reach the burst sizes: units from the memo
    (Mt < 48), from the big list (48 <= Mt < 256, small pmf buffers) and from
    the huge list (Mt >= 256, buffers for the run's largest total) in one run.
    Forward bit-exact against the oracle; adjoint against its FP64 gradient;
    Jacobian rows equal to the adjoint of their cotangents."""
    E = [0.002, 0.5, 0.998]
    args = synth.make_args(E, [300, 800, 400, 900, 500, 700])
    d = synth.generate(60, args, 3, base_photons=200, mean_geom=200, seed=71)
    obs = _load(gpu_mc, d)
    R, seed = 4, 5
    od = oracle.Data.from_bursts(d)
    p = oracle.make_params(3, np.asarray(args, np.float32), *PH)
    mt = np.array([int(np.sum(oracle.rb_unit(od, p, b, r, seed, 0)[0]))
                   for b in range(d.nbursts) for r in range(R)])
    assert (mt < 48).sum() > 20 and ((mt >= 48) & (mt < 256)).sum() > 20 and (mt >= 256).sum() > 20
    M, fixed, _, _ = _rb(gpu_mc, 3, args, R, seed, set_id=0)
    want = oracle.rb_run(od, p, R, seed, 0)
    assert np.array_equal(fixed, want), np.nonzero(fixed != want)
    g = _chi2_cotangent(M, _obs_hist(obs), R, 3)
    grad = gpu_mc.rb_adjoint(g, 3)
    _, go = oracle.rb_grad(od, p, R, seed, g, set_id=0)
    scale = np.abs(go).max()
    assert np.all(np.abs(grad[:9] - go) <= 1e-3 * np.maximum(np.abs(go), 1e-3 * scale)), (grad, go)
    J = gpu_mc.rb_jacobian(3, len(M))
    rng = np.random.default_rng(9)
    for _ in range(2):
        c = rng.normal(size=50)
        adj = gpu_mc.rb_adjoint(c, 3)
        sc = np.abs(adj).max()
        assert np.all(np.abs(c @ J - adj) <= 1e-5 * np.maximum(np.abs(adj), 1e-2 * sc)), (c @ J, adj)


def test_rb_efficiency_gradient_matches_central_differences(gpu_mc, oracle):
    d = _data(20000, 61)
    obs = _load(gpu_mc, d)
    oh = _obs_hist(obs)
    R, seed = 32, 13
    M, _, c2, _ = _rb(gpu_mc, 3, ARGS3, R, seed)
    g = _chi2_cotangent(M, oh, R, 3)
    grad = gpu_mc.rb_adjoint(g, 3)
    a0 = np.asarray(ARGS3, np.float32)
    for s in range(3):
        h = np.float32(1e-4) * a0[s]
        ap, am = a0.copy(), a0.copy()
        ap[s] += h
        am[s] -= h
        _, _, cp, _ = _rb(gpu_mc, 3, ap, R, seed)
        _, _, cm, _ = _rb(gpu_mc, 3, am, R, seed)
        # the library differentiates in the acceptor probability p(E), E after
        # the float distance round trip: divide by the change of p
        pp = oracle.lib().or_state_efficiency(float(ap[s]), PH[3])
        pm = oracle.lib().or_state_efficiency(float(am[s]), PH[3])
        fd = (cp - cm) / (np.float64(np.float32(pp)) - np.float64(np.float32(pm)))
        assert abs(grad[s] - fd) <= 1e-3 * abs(fd), (s, grad[s], fd)


def test_rb_rate_gradient_agrees_with_central_differences(gpu_mc):
    d = _data(4000, 71)
    obs = _load(gpu_mc, d)
    oh = _obs_hist(obs)
    R = 32
    a0 = np.asarray(ARGS3, np.float32)
    M0, _, _, _ = _rb(gpu_mc, 3, a0, R, 1000)
    g = _chi2_cotangent(M0, oh, R, 3)
    lr, fd = [], []
    for seed in range(40):
        M, _, _, _ = _rb(gpu_mc, 3, a0, R, 2000 + seed)
        lr.append(gpu_mc.rb_adjoint(g, 3)[3:9])
        row = []
        for i in range(6):
            h = np.float32(0.05) * a0[3 + i]
            ap, am = a0.copy(), a0.copy()
            ap[3 + i] += h
            am[3 + i] -= h
            Mp, _, _, _ = _rb(gpu_mc, 3, ap, R, 2000 + seed, scores=False)
            Mm, _, _, _ = _rb(gpu_mc, 3, am, R, 2000 + seed, scores=False)
            row.append(g @ (Mp - Mm) / (np.float64(ap[3 + i]) - np.float64(am[3 + i])))
        fd.append(row)
    lr, fd = np.array(lr), np.array(fd)
    diff = lr - fd
    z = diff.mean(0) / (diff.std(0, ddof=1) / np.sqrt(len(diff)))
    assert np.all(np.abs(z) < 4.0), (lr.mean(0), fd.mean(0), z)


def test_torch_autograd_function(gpu_mc):
    torch = pytest.importorskip("torch")
    from gsmfretda_amd import rb_torch
    d = _data(1500, 81)
    obs = _load(gpu_mc, d)
    oh = _obs_hist(obs)
    R, seed = 16, 4
    args = torch.tensor(np.asarray(ARGS3, np.float64), requires_grad=True)
    loss = rb_torch.rb_chi2(args, gpu_mc, 3, R, seed, torch.tensor(oh))
    loss.backward()
    M, _, c2, _ = gpu_mc.rb_result()
    assert abs(loss.item() - c2) <= 1e-12 * abs(c2)
    g = _chi2_cotangent(M, oh, R, 3)
    # the adjoint of this same forward, called directly
    want = gpu_mc.rb_adjoint(g, 3)
    assert np.allclose(args.grad.numpy(), want, rtol=1e-9, atol=0), (args.grad, want)


def test_torch_backward_refuses_after_another_forward_on_its_slot(gpu_mc):
    """The adjoint differentiates the slot's last forward: a direct
    MC.rb_forward on the same slot between forward and backward is refused,
    one on another slot is not."""
    torch = pytest.importorskip("torch")
    from gsmfretda_amd import rb_torch
    d = _data(400, 83)
    obs = _load(gpu_mc, d)
    oh = torch.tensor(_obs_hist(obs))
    R, seed = 8, 5
    args = torch.tensor(np.asarray(ARGS3, np.float64), requires_grad=True)
    loss = rb_torch.rb_chi2(args, gpu_mc, 3, R, seed, oh, sid=0)
    _rb(gpu_mc, 3, ARGS3, R, seed + 1)  # slot 0 again
    with pytest.raises(pda.PdaError, match="slot 0"):
        loss.backward()
    loss = rb_torch.rb_chi2(args, gpu_mc, 3, R, seed, oh, sid=0)
    gpu_mc.set_photophysics(1, *PH)
    gpu_mc.set_burst_range(0, -1, 1)
    assert gpu_mc.set_params(3, 1, np.asarray(ARGS3, np.float32))
    gpu_mc.rb_forward(R=R, seed=seed + 2, sid=1)  # another slot
    gpu_mc.rb_result(1)
    loss.backward()
    assert np.all(np.isfinite(args.grad.numpy()))


def test_rb_errors(gpu_mc):
    d = _data(50, 91)
    _load(gpu_mc, d)
    with pytest.raises(pda.PdaError, match="v = 0"):
        a = np.array(ARGS3, np.float32)
        a[9] = 2.0
        _rb(gpu_mc, 3, a, 4, 1)
    with pytest.raises(pda.PdaError, match="n <= 4"):
        _rb(gpu_mc, 5, synth.make_args(np.linspace(0.1, 0.9, 5), np.full(20, 1e3)), 4, 1)
    _rb(gpu_mc, 3, ARGS3, 4, 1, scores=False)
    with pytest.raises(pda.PdaError, match="scores"):
        gpu_mc.rb_adjoint(np.zeros(50), 3)
