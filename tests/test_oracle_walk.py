"""Pins the oracle's dwell walk, histogram and chi2.

* A second, pure-Python restatement of mc_kernel (src/mc.cu:83-197, window
  scan with sticky flags, cumulative window ends, dwell at the chosen rate,
  background, broadening) over the same primitives, on tiny hand-built and
  random bursts: must agree with the C oracle bit for bit.
* Closed forms: E_state = 1 / 0, single-window bursts whose per-sample E
  follows Binomial(count, 1/2) / count exactly, hand-counted sticky runs.
* Histogram edges and chi2 against exact rational arithmetic.
* Golden fixture regression (tests/golden/).
"""
import math
import os
from fractions import Fraction

import numpy as np
import pytest
from scipy import stats

from gsmfretda_amd import synth
from tests.conftest import ROOT, bits

f32 = np.float32


def py_unit(oracle, d, p, b, r, seed, set_id):
    """Appendix-A restatement of one (burst, resample) unit in Python."""
    L = oracle.lib()
    n = p.n
    E = [f32(x) for x in p.E[:n]]
    v = [f32(x) for x in p.v[:n]]
    K = np.array(p.K[:n * n], np.float32)
    Pi = np.array(p.Pi2j[:n * n], np.float32)
    p0 = np.array(p.p0[:n], np.float32)
    gmx = f32(f32(p.gamma) - f32(p.crosstalk))
    omdd = f32(f32(1.0) - f32(p.dexdir))
    r0 = f32(p.r0)
    clk = f32(d.clk_p)
    P = int(d.phcount[b])
    base = int(d.istart[b])
    t0 = int(d.start[b])
    stop_rel = int(d.stop[b]) - t0
    T = f32(d.T[b])

    def blk(w, sub):
        return [int(x) for x in oracle.philox([w & 0xFFFFFFFF, sub, r, b], [seed, set_id])]

    def categorical(probs, u):
        a = f32(0)
        for i, q in enumerate(probs):
            a = f32(a + q)
            if a > u:
                return i
        return len(probs) - 1

    s = categorical(p0, f32(L.or_u01f(blk(0xFFFFFFFF, 0)[0])))
    tsum, F, acc = f32(0), f32(0), f32(0)
    w0, cursor, w = 0, 0, 0
    while T > tsum and w < (1 << 22):
        A = blk(w, 0)
        col = [Pi[s * n + (ii + 1 if ii >= s else ii)] for ii in range(n - 1)]
        j = categorical(col, f32(L.or_u01f(A[0])))
        if j >= s:
            j += 1
        k = K[j + s * n]
        tsum = f32(tsum + f32(f32(L.or_logf(L.or_u01f(A[1]))) / f32(-k)))
        if not tsum < T:
            w1 = stop_rel
        else:
            w1 = int(f32(f32(w0) + f32(tsum / clk)))
        sd = sa = False
        nd = na = 0
        x = 0
        for i in range(cursor, P):
            tt = int(d.times[base + i]) - t0
            isD, isA = bool(d.mask_dd[base + i]), bool(d.mask_ad[base + i])
            if (isD and tt >= w1) or (isA and tt >= w1) or i == P - 1:
                cursor = i
                f = f32(f32(gmx * f32(nd)) + f32(omdd * f32(na)))
                F = f32(F + f)
                if f32(d.bg_dd) < f32(1e-4):
                    m = math.floor(f32(f + f32(0.5)))
                else:
                    B = blk(w, 1)
                    dt = f32(f32(w1 - w0) * clk)
                    bga = f32(L.or_bg_capped(float(f32(dt * f32(d.bg_ad))), float(f32(omdd * f32(na))),
                                             L.or_u01f(B[0]), L.or_u01f(B[2]), seed, set_id, b, r, w,
                                             0x100000, None))
                    bgd = f32(L.or_bg_capped(float(f32(dt * f32(d.bg_dd))), float(f32(gmx * f32(nd))),
                                             L.or_u01f(B[1]), L.or_u01f(B[3]), seed, set_id, b, r, w,
                                             0x200000, None))
                    m = math.floor(f32(f32(f32(f - bgd) - bga) + f32(0.5)))
                C = blk(w, 2)
                z = f32(L.or_normal(C[0], C[1]))
                xx = f32(f32(1) / E[s] - f32(1)) if E[s] != 0 else f32(np.inf)
                base_r = f32(r0 * f32(L.or_expf(float(f32(f32(L.or_logf(float(xx))) * f32(1 / 6))))))
                rd = f32(f32(z * v[s]) + base_r)
                q = f32(rd / r0)
                q2 = f32(q * q)
                de = f32(f32(1) / f32(f32(1) + f32(f32(q2 * q2) * q2)))
                x = L.or_binomial(int(m), float(de), seed, set_id, b, r, w, A[2], A[3], None)
                break
            if isD:
                sd = True
            if isA:
                sa = True
            if sa and sd:
                continue
            if sa:
                na += 1
            if sd:
                nd += 1
        acc = f32(acc + f32(x))
        w0 = w1
        s = j
        w += 1
    Ev = f32(acc / F) if F > 0 else f32(0)
    return int(acc), F, Ev


def tiny_bursts(rng, nb, p_neither=0.2, p_both=0.05):
    """Hand-sized bursts: 0-25 photons, random channels incl. neither / both."""
    P = rng.integers(0, 26, nb)
    P[:4] = [0, 1, 2, 3]
    dur = rng.integers(20, 4000, nb)
    start = 1000 + np.cumsum(dur + 50) - dur
    stop = start + dur - 1
    istart = np.concatenate([[0], np.cumsum(P)[:-1]])
    times, mad, mdd = [], [], []
    for b in range(nb):
        t = np.sort(rng.integers(0, dur[b], P[b]))
        times.append(start[b] + t)
        u = rng.random(P[b])
        a = (u < 0.5).astype(np.uint8)
        dd = 1 - a
        nei = u > 1 - p_neither
        a[nei] = 0
        dd[nei] = 0
        both = rng.random(P[b]) < p_both
        a[both] = 1
        dd[both] = 1
        mad.append(a)
        mdd.append(dd)
    T = (dur * 12.5e-9).astype(np.float32)
    T[5] = 0.0
    return synth.BurstData(start=start.astype(np.int64), stop=stop.astype(np.int64),
                           istart=istart.astype(np.int64), phcount=P.astype(np.int32),
                           times=np.concatenate(times).astype(np.int64),
                           mask_ad=np.concatenate(mad), mask_dd=np.concatenate(mdd), T=T)


@pytest.mark.parametrize("bg,n,vs", [(0.0, 2, 0.0), (1000.0, 3, 0.0), (5e5, 3, 4.0), (1000.0, 4, 0.0)])
def test_python_restatement_agrees(oracle, bg, n, vs):
    rng = np.random.default_rng(int(bg) % 97 + n)
    d = tiny_bursts(rng, 40)
    d.bg_dd = bg
    d.bg_ad = bg / 2
    E = np.sort(rng.uniform(0, 1, n))
    E[0] = 0.0 if n > 2 else E[0]
    rates = np.exp(rng.uniform(np.log(1e4), np.log(1e6), n * (n - 1)))
    rates[0] = 0.0 if n == 4 else rates[0]   # a zero-rate transition (Q16)
    args = synth.make_args(E, rates, v=np.full(n, vs))
    p = oracle.make_params(n, args)
    od = oracle.Data.from_bursts(d)
    R = 6
    got = oracle.run(od, p, R=R, seed=77, set_id=3)
    for b in range(d.nbursts):
        for r in range(R):
            a, F, Ev = py_unit(oracle, d, p, b, r, 77, 3)
            u = b * R + r
            assert got["acc"][u] == a, (b, r)
            assert bits(np.array([got["F"][u]]))[0] == bits(np.array([F]))[0], (b, r)
            assert bits(np.array([got["E"][u]]))[0] == bits(np.array([Ev]))[0], (b, r)


def test_closed_forms_extreme_efficiencies(oracle):
    d = synth.generate(300, synth.make_args([0.5, 0.5], [3e3, 3e3]), 2, seed=8)
    od = oracle.Data.from_bursts(d)
    # E = 1 for every state, gamma = 1, no crosstalk / direct excitation, no background:
    # every corrected count is an integer and every photon is drawn as acceptor -> E = 1.
    d.bg_dd = 0.0
    od = oracle.Data.from_bursts(d)
    p = oracle.make_params(2, synth.make_args([1.0, 1.0], [3e3, 3e3]), 1.0, 0.0, 0.0, 52.0)
    out = oracle.run(od, p, R=8, seed=1)
    pos = out["F"] > 0
    assert np.all(out["E"][pos] == 1.0) and np.all(out["E"][~pos] == 0.0)
    assert out["hist"].sum() == (~pos).sum()          # E = 1.0 is outside [0,1) (Q13)
    assert np.all(out["acc"][pos] == out["F"][pos])
    # E = 0: no acceptor draws at all
    p = oracle.make_params(2, synth.make_args([0.0, 0.0], [3e3, 3e3]))
    out = oracle.run(od, p, R=8, seed=2)
    assert np.all(out["acc"] == 0) and np.all(out["E"] == 0)


def _single_window_data(pattern_list):
    """One burst per channel pattern ('D','A','N','B'), photons 10 ticks apart."""
    start, stop, istart, P, times, mad, mdd = [], [], [], [], [], [], []
    off, t = 0, 1000
    for pat in pattern_list:
        start.append(t)
        istart.append(off)
        P.append(len(pat))
        for i, ch in enumerate(pat):
            times.append(t + 10 * i)
            mad.append(1 if ch in "AB" else 0)
            mdd.append(1 if ch in "DB" else 0)
        stop.append(t + 10 * max(len(pat) - 1, 0))
        off += len(pat)
        t += 10 * len(pat) + 1000
    n = len(pattern_list)
    return synth.BurstData(start=np.array(start, np.int64), stop=np.array(stop, np.int64),
                           istart=np.array(istart, np.int64), phcount=np.array(P, np.int32),
                           times=np.array(times, np.int64), mask_ad=np.array(mad, np.uint8),
                           mask_dd=np.array(mdd, np.uint8),
                           T=np.full(n, 1e-3, np.float32), bg_dd=0.0, bg_ad=0.0)


def test_sticky_runs_by_hand(oracle):
    """Single window (rates 1e-6 /s, so the first dwell outlasts T): the window
    closes at the last photon, and the count is the first same-colour run,
    neither-channel photons included once a flag is set (Q1, Q5, Q19)."""
    patterns = ["DDDADD", "AAD", "NNDNDAD", "NNNA", "BDDA", "DNNNNNA", "ADDDD", "N", "DDDDDD"]
    expect_counts = [3, 2, 3, 0, 0, 6, 1, 0, 5]   # last photon never counted
    d = _single_window_data(patterns)
    od = oracle.Data.from_bursts(d)
    p = oracle.make_params(2, synth.make_args([1.0, 1.0], [1e-6, 1e-6]), 1.0, 0.0, 0.0, 52.0)
    out = oracle.run(od, p, R=4, seed=3)
    F = out["F"].reshape(len(patterns), 4)
    assert np.array_equal(F[:, 0], np.array(expect_counts, np.float32))
    assert np.all(F == F[:, :1])


def test_single_window_binomial_law(oracle):
    """E_state = 1/2 with one window of c counted photons: acc ~ Binomial(c, 1/2) exactly."""
    d = _single_window_data(["DDDDDDDDDDDDA"] * 50)   # c = 12
    od = oracle.Data.from_bursts(d)
    p = oracle.make_params(2, synth.make_args([0.5, 0.5], [1e-6, 1e-6]), 1.0, 0.0, 0.0, 52.0)
    out = oracle.run(od, p, R=400, seed=4)
    assert np.all(out["F"] == 12.0)
    cnt = np.bincount(out["acc"], minlength=13)
    pv = stats.chisquare(cnt[2:11], stats.binom.pmf(np.arange(2, 11), 12, 0.5) * cnt.sum()
                         * (cnt[2:11].sum() / (stats.binom.pmf(np.arange(2, 11), 12, 0.5).sum()
                                               * cnt.sum()))).pvalue
    assert pv > 1e-3
    assert abs(out["acc"].mean() - 6.0) < 0.1


WINDOW_LAM, WINDOW_DELTA, WINDOW_P = 2e4, 400, 101


def window_end_data(nb):
    """Identical single-colour bursts (donor photons WINDOW_DELTA ticks apart,
    WINDOW_P of them), T = the burst span, no background."""
    P, delta = WINDOW_P, WINDOW_DELTA
    ticks = np.arange(P, dtype=np.int64) * delta
    start = 1000 + np.arange(nb, dtype=np.int64) * (P * delta + 1000)
    return synth.BurstData(start=start, stop=start + P * delta,
                           istart=np.arange(nb, dtype=np.int64) * P,
                           phcount=np.full(nb, P, np.int32),
                           times=(start[:, None] + ticks[None, :]).ravel(),
                           mask_ad=np.zeros(nb * P, np.uint8), mask_dd=np.ones(nb * P, np.uint8),
                           T=np.full(nb, P * delta * synth.CLK_P, np.float32), bg_ad=0.0, bg_dd=0.0)


WINDOW_ARGS = synth.make_args([1.0, 0.0], [WINDOW_LAM, WINDOW_LAM])
WINDOW_PH = (1.0, 0.0, 0.0, 52.0)


def window_end_prediction(cumulative=True, kmax=23):
    """Expected acceptor count per unit on window_end_data under WINDOW_ARGS:
    2 states, E = (1, 0), both rates lam, initial state 0 (its inflow lam > 1
    exceeds every uniform), gamma = 1 and no crosstalk, so a window's trial
    count is its photon count and state-0 windows (k even) give every photon
    to the acceptor, state-1 windows none.

    Window k ends at W_{k+1} (seconds after the burst start) and photon j
    (tick j delta) is counted in a window <= k iff its tick lies below the
    truncated end tick, i.e. W_{k+1} >= (j delta + 1) clk; the last photon is
    never counted (src/mc.cu:140-148).  A window whose dwell sum reaches T
    ends at the stop, after every photon -- and with T the burst span this
    changes no assignment, since then W >= S >= T.  So

        E[acc] = sum_{j < P-1} sum_{k even} P(W_{k+1} >= y_j) - P(W_k >= y_j).

    src/mc.cu:132 and :184 set W_{k+1} = W_k + S_{k+1} with S the cumulative
    dwell sum (mcSpendTime), so W_k = sum_{m=1..k} m tau_m for iid
    tau ~ Exp(lam): a hypoexponential law of rates lam / m, whose tail is
    sum_m C_m exp(-lam y / m) with C_m = prod_{l != m} m / (m - l) (exact
    rationals, evaluated at 50 digits).  The misreading W_k = S_k (window
    ends at burst start + dwell sum) makes W_k Erlang(k, lam)."""
    import mpmath as mp
    mp.mp.dps = 50
    lam = mp.mpf(WINDOW_LAM)
    clk = mp.mpf(synth.CLK_P)

    if cumulative:
        coef = {}
        for k in range(1, kmax + 2):
            cs = []
            for m in range(1, k + 1):
                c = Fraction(1)
                for l in range(1, k + 1):
                    if l != m:
                        c *= Fraction(m, m - l)
                cs.append(mp.mpf(c.numerator) / c.denominator)
            coef[k] = cs

        def sf(y, k):
            if k == 0:
                return mp.mpf(0)
            return mp.fsum(c * mp.exp(-lam / m * y) for m, c in enumerate(coef[k], start=1))
    else:
        def sf(y, k):
            if k == 0:
                return mp.mpf(0)
            return mp.gammainc(k, lam * y, mp.inf, regularized=True)

    tot = mp.mpf(0)
    for j in range(WINDOW_P - 1):
        y = mp.mpf(j * WINDOW_DELTA + 1) * clk
        for k in range(0, kmax, 2):
            tot += sf(y, k + 1) - sf(y, k)
    return float(tot)


def test_cumulative_window_ends_closed_form(oracle):
    """Q2 (src/mc.cu:132, 184): window ends add the cumulative dwell sum to
    the previous end.  The oracle's mean acceptor count on evenly spaced,
    single-colour bursts equals the closed form of that reading and rejects
    the closed form of window ends at burst start + dwell sum."""
    d = window_end_data(2000)
    p = oracle.make_params(2, WINDOW_ARGS, *WINDOW_PH)
    out = oracle.run(oracle.Data.from_bursts(d), p, R=100, seed=11)
    assert np.all(out["F"] == WINDOW_P - 1)
    acc = out["acc"].astype(np.float64)
    se = acc.std(ddof=1) / np.sqrt(acc.size)
    want = window_end_prediction(cumulative=True)
    assert abs(want - window_end_prediction(cumulative=True, kmax=19)) < 1e-9  # tail converged
    wrong = window_end_prediction(cumulative=False, kmax=61)
    assert abs(acc.mean() - want) < 4.5 * se, (acc.mean(), want, se)
    assert abs(acc.mean() - wrong) > 20 * se, (acc.mean(), wrong, se)


def test_histogram_edges(oracle):
    x = np.array([0.0, 0.02, np.nextafter(np.float32(0.02), np.float32(1)), 0.5, 0.999999, 1.0,
                  -0.1, np.nan, 0.98], np.float32)
    h = oracle.hist(x, B=50)
    want = np.zeros(50, np.int64)
    for k in (0, 0, 1, 25, 49, 49):
        want[k] += 1
    assert np.array_equal(h, want)


def test_chi2_exact(oracle):
    rng = np.random.default_rng(9)
    B, R, n, pad = 50, 7, 3, 2
    O = rng.integers(0, 40, B).astype(np.float32)
    M = rng.integers(0, 300, B)
    M[[3, 10, 11]] = 0
    c32, c64, eff = oracle.chi2(O, M, R, n, pad)
    keep = [b for b in range(pad, B) if M[b] > 0]
    assert eff == len(keep)
    exact = sum((Fraction(float(O[b])) - Fraction(int(M[b]), R)) ** 2 / Fraction(int(M[b]), R)
                for b in keep) / (len(keep) - n * (n + 1))
    assert abs(c64 - float(exact)) <= 1e-12 * abs(float(exact))
    # the reference's float accumulation (src/streamWorker.cpp:197-206)
    acc = f32(0)
    for b in keep:
        dlt = f32(O[b] - f32(f32(M[b]) / f32(R)))
        acc = f32(np.float64(acc) + np.float64(dlt) * np.float64(dlt) / np.float64(f32(f32(M[b]) / f32(R))))
    assert f32(c32) == f32(acc / f32(len(keep) - n * (n + 1)))
    assert abs(c32 - float(exact)) <= 1e-5 * abs(float(exact))


def test_golden_fixture(oracle):
    path = os.path.join(ROOT, "tests", "golden", "walk_small.npz")
    g = np.load(path, allow_pickle=False)
    d = synth.BurstData(start=g["start"], stop=g["stop"], istart=g["istart"], phcount=g["phcount"],
                        times=g["times"], mask_ad=g["mask_ad"], mask_dd=g["mask_dd"], T=g["T"],
                        clk_p=float(g["clk_p"]), bg_ad=float(g["bg_ad"]), bg_dd=float(g["bg_dd"]))
    od = oracle.Data.from_bursts(d)
    for k in range(int(g["nsets"])):
        n = int(g[f"n{k}"])
        p = oracle.make_params(n, g[f"args{k}"], *g[f"ph{k}"])
        out = oracle.run(od, p, R=int(g["R"]), seed=int(g["seed"]), set_id=k)
        assert np.array_equal(out["acc"], g[f"acc{k}"])
        assert np.array_equal(bits(out["F"]), bits(g[f"F{k}"]))
        assert np.array_equal(bits(out["E"]), bits(g[f"E{k}"]))
        assert np.array_equal(out["hist"], g[f"hist{k}"])
