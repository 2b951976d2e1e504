"""An independent reference-semantics Monte Carlo of mc_kernel in numpy.

Written from the reference's algorithm (src/mc.cu:83-197, src/gen_rand.cuh:
35-107, src/binom.cuh) with numpy's own random generator in place of curand:
next state by the categorical over P_i2j's column, dwell Exp(K(j,s)) at the
chosen transition's rate, cumulative float window ends, the sticky-flag count
of the first same-colour run, capped Poisson background, normal distance
broadening and numpy's binomial.  It shares no code and no random stream with
the oracle or the HIP kernels, so agreement with them is distributional only
(two-sample chi-square on the E histograms).

Test infrastructure; resamples of one burst are vectorised.
"""
import numpy as np

from gsmfretda_amd import synth


def _burst_tables(times_rel, ch):
    """For a burst: channel-photon ticks, and for each photon i the first
    channel photon >= i and the end of its same-colour run."""
    P = len(ch)
    first = np.full(P + 1, P, np.int64)
    run_end = np.full(P + 1, P, np.int64)
    nxt_d = nxt_a = P
    for i in range(P - 1, -1, -1):
        c = ch[i]
        if c:
            first[i] = i
            run_end[i] = {1: nxt_a, 2: nxt_d, 3: i}[int(c)]
        else:
            first[i] = first[i + 1]
        if c & 1:
            nxt_d = i
        if c & 2:
            nxt_a = i
    return first, run_end


def simulate(data, n, args, R, rng, gamma=0.34, dexdir=0.08, crosstalk=0.07, r0=52.0):
    """Per-sample E of R resamples of every burst, reference semantics."""
    args = np.asarray(args, np.float64)
    E = args[:n].astype(np.float32)
    v = args[n * n:].astype(np.float32)
    K = synth.rate_matrix(n, args).astype(np.float32)     # K[i, j] = rate j -> i
    colsum = np.array([np.float32(sum(np.float32(K[ii, j]) for ii in range(n) if ii != j))
                       for j in range(n)], np.float32)
    p0 = np.array([np.float32(np.float32(K[i, :].sum(dtype=np.float32)) - np.float32(0))
                   for i in range(n)], np.float32)
    gmx = np.float32(np.float32(gamma) - np.float32(crosstalk))
    omdd = np.float32(np.float32(1.0) - np.float32(dexdir))
    clk = np.float32(data.clk_p)
    out = np.zeros(data.nbursts * R, np.float32)
    for b in range(data.nbursts):
        P = int(data.phcount[b])
        if P <= 1:
            continue
        base = int(data.istart[b])
        t0 = int(data.start[b])
        tt = data.times[base:base + P] - t0
        ch = (data.mask_dd[base:base + P].astype(np.int64) |
              (data.mask_ad[base:base + P].astype(np.int64) << 1))
        first, run_end = _burst_tables(tt, ch)
        chan_idx = np.nonzero(ch)[0]
        chan_t = tt[chan_idx]
        T = np.float32(data.T[b])
        stop_rel = int(data.stop[b]) - t0
        # initial state: first i with cumulative inflow > u (src/gen_rand.cuh:35-50)
        u = rng.uniform(size=R)
        cum = np.cumsum(p0)
        s = np.minimum((u[:, None] >= cum[None, :]).sum(axis=1), n - 1)
        tsum = np.zeros(R, np.float32)
        w0 = np.zeros(R, np.int64)
        cur = np.zeros(R, np.int64)
        acc = np.zeros(R, np.float32)
        F = np.zeros(R, np.float32)
        act = np.full(R, T > 0)
        while act.any():
            idx = np.nonzero(act)[0]
            si = s[idx]
            # next state j != s with probability K(j, s) / sum
            probs = K[:, si].T / colsum[si][:, None]
            probs[np.arange(len(idx)), si] = 0.0
            cj = np.cumsum(probs, axis=1)
            uj = rng.uniform(size=len(idx))
            j = np.minimum((uj[:, None] >= cj).sum(axis=1), n - 1)
            j = np.where(j == si, (si + 1) % n, j)
            k = K[j, si]
            with np.errstate(divide="ignore"):
                tau = rng.exponential(1.0, size=len(idx)).astype(np.float32) / k
            ts = (tsum[idx] + tau).astype(np.float32)
            wf = (w0[idx].astype(np.float32) + (ts / clk).astype(np.float32)).astype(np.float32)
            w1 = np.where(ts >= T, stop_rel, np.floor(np.minimum(wf, 2.0**62)).astype(np.int64))
            # close: first channel photon at or after w1 (from the cursor), else the last
            pos = np.searchsorted(chan_t, w1, side="left")
            cand = np.where(pos < len(chan_idx), chan_idx[np.minimum(pos, len(chan_idx) - 1)], P)
            cand = np.where(cand < cur[idx], np.maximum(first[cur[idx]], cur[idx]), cand)
            close = np.minimum(cand, P - 1)
            f = first[cur[idx]]
            cnt = np.where(f < close, np.minimum(close, run_end[f]) - f, 0)
            colour = np.where(f < P, ch[np.minimum(f, P - 1)], 0)
            nd = np.where(colour == 1, cnt, 0).astype(np.float32)
            na = np.where(colour == 2, cnt, 0).astype(np.float32)
            fv = (gmx * nd + omdd * na).astype(np.float32)
            if data.bg_dd < 1e-4:
                m = np.floor(fv + np.float32(0.5))
            else:
                dt = ((w1 - w0[idx]).astype(np.float32) * clk).astype(np.float32)
                def bg(rate, cap):
                    kk = rng.poisson(np.maximum(dt * np.float32(rate), 0.0))
                    alt = np.floor(cap * rng.uniform(size=len(idx)).astype(np.float32))
                    return np.where(kk <= cap, kk, alt).astype(np.float32)
                bga = bg(data.bg_ad, (omdd * na).astype(np.float32))
                bgd = bg(data.bg_dd, (gmx * nd).astype(np.float32))
                m = np.floor(((fv - bgd) - bga) + np.float32(0.5))
            base_r = np.where(E[si] > 0, r0 * np.power(np.maximum(1 / E[si] - 1, 0), 1 / 6), np.inf)
            rd = rng.normal(size=len(idx)).astype(np.float32) * v[si] + base_r
            de = 1.0 / (1.0 + (rd / r0) ** 6)
            x = rng.binomial(np.maximum(m, 0).astype(np.int64), np.clip(de, 0, 1))
            F[idx] += fv
            acc[idx] += x.astype(np.float32)
            tsum[idx] = ts
            w0[idx] = w1
            s[idx] = j
            cur[idx] = close
            done = (close == P - 1) | ~(ts < T)
            act[idx[done]] = False
        Eb = np.where(F > 0, acc / np.where(F > 0, F, 1), 0).astype(np.float32)
        out[b * R:(b + 1) * R] = Eb
    return out
