"""Deterministic synthetic smFRET burst data in the reference's HDF5 layout.

The reference reads /sub_bursts_l/{start,stop,istart,istop}, times, mask_ad,
mask_dd, T_burst_duration, SgDivSr, clk_p, bg_ad_rate and bg_dd_rate
(src/loadHdf5.cpp:37-74); this module produces the same arrays from a known
kinetic model, so the whole path can be exercised without the (unavailable)
measured data:

* photons per burst = base + Geometric(mean m) on {0, 1, ...};
* burst duration ~ LogNormal(ln median, sigma), in 12.5 ns ticks;
* photon ticks are sorted uniform draws in the burst window; bursts are
  placed one after another on an absolute macrotime axis;
* each photon is an acceptor photon with probability E of the state the
  true continuous-time Markov trajectory is in at that tick, else a donor
  photon (optionally a fraction of photons in neither channel, as the
  acceptor-excitation photons of ALEX data).

Parameter vectors use mc::set_params' layout (src/mc.cu:790-800):
args = [E(n), K off-diagonals (n(n-1)), v(n)], the off-diagonal entry for
(i, j), i != j, being the rate of j -> i, listed row-major over i.
"""
from dataclasses import dataclass, field

import numpy as np

CLK_P = 12.5e-9


@dataclass
class BurstData:
    start: np.ndarray
    stop: np.ndarray
    istart: np.ndarray
    phcount: np.ndarray
    times: np.ndarray
    mask_ad: np.ndarray
    mask_dd: np.ndarray
    T: np.ndarray
    clk_p: float = CLK_P
    bg_ad: float = 500.0
    bg_dd: float = 1000.0
    truth: dict = field(default_factory=dict)

    @property
    def nbursts(self):
        return len(self.start)

    @property
    def nphotons(self):
        return len(self.times)

    @property
    def istop(self):
        return self.istart + self.phcount.astype(np.int64) - 1

    def burst_ad(self):
        """The reference's "time if acceptor else 0" array (src/loadHdf5.cpp:120)."""
        return np.where(self.mask_ad != 0, self.times, 0).astype(np.int64)

    def burst_dd(self):
        return np.where(self.mask_dd != 0, self.times, 0).astype(np.int64)


def rate_matrix(n, args):
    """K(i, j) = rate j -> i (genMatK's column-major K without the diagonal)."""
    kargs = np.asarray(args, np.float64)[n:n + n * (n - 1)]
    K = np.zeros((n, n))
    for i in range(n):
        for j in range(n):
            if i < j:
                K[i, j] = kargs[i * (n - 1) + j - 1]
            elif i > j:
                K[i, j] = kargs[i * (n - 1) + j]
    return K


def make_args(E, K_offdiag, v=None):
    n = len(E)
    if v is None:
        v = np.zeros(n)
    return np.concatenate([np.asarray(E, np.float64), np.asarray(K_offdiag, np.float64),
                           np.asarray(v, np.float64)]).astype(np.float32)


def stationary(K):
    n = K.shape[0]
    Q = K.copy()
    Q[np.diag_indices(n)] = -K.sum(axis=0)
    A = np.vstack([Q, np.ones(n)])
    b = np.zeros(n + 1)
    b[-1] = 1.0
    pi = np.linalg.lstsq(A, b, rcond=None)[0]
    pi = np.clip(pi, 0, None)
    return pi / pi.sum()


def generate(n_bursts, args, n_states, base_photons=50, mean_geom=40.0, dur_median=1e-3,
             dur_sigma=0.5, clk_p=CLK_P, bg_ad=500.0, bg_dd=1000.0, seed=0, p_neither=0.0,
             gap_ticks=200_000, first_tick=1_000_000):
    """Synthetic bursts from the true kinetic model `args` (mc::set_params layout)."""
    rng = np.random.default_rng(seed)
    n = n_states
    args = np.asarray(args, np.float64)
    E = args[:n]
    K = rate_matrix(n, args)
    exit_rate = K.sum(axis=0)

    P = base_photons + rng.geometric(1.0 / (mean_geom + 1.0), size=n_bursts) - 1
    P = P.astype(np.int64)
    dur_s = rng.lognormal(np.log(dur_median), dur_sigma, size=n_bursts)
    dur = np.maximum(np.rint(dur_s / clk_p).astype(np.int64), 1)
    dur = np.minimum(dur, (1 << 31) - 2)
    T = (dur * clk_p).astype(np.float32)

    istart = np.zeros(n_bursts, np.int64)
    if n_bursts > 1:
        istart[1:] = np.cumsum(P)[:-1]
    total = int(P.sum())
    owner = np.repeat(np.arange(n_bursts, dtype=np.int64), P)
    ticks = rng.integers(0, dur[owner]) if total else np.zeros(0, np.int64)
    key = np.sort((owner << 32) | ticks)
    ticks = key & 0xFFFFFFFF

    start = np.zeros(n_bursts, np.int64)
    start[0] = first_tick
    if n_bursts > 1:
        start[1:] = first_tick + np.cumsum(dur + gap_ticks)[:-1]
    stop = start + dur - 1
    times = start[owner] + ticks

    # true trajectory: continuous-time Markov chain, stationary start
    pi = stationary(K)
    cum_jump = np.zeros((n, n))
    for s in range(n):
        p = K[:, s] / exit_rate[s] if exit_rate[s] > 0 else np.eye(n)[s]
        cum_jump[s] = np.cumsum(p)
    state = rng.choice(n, size=n_bursts, p=pi)
    t = np.zeros(n_bursts)
    active = np.arange(n_bursts)
    seg_b, seg_end, seg_s = [], [], []
    while active.size:
        s = state[active]
        with np.errstate(divide="ignore"):
            dwell = rng.exponential(1.0, size=active.size) / exit_rate[s] / clk_p
        t_end = t[active] + dwell
        seg_b.append(active)
        seg_end.append(np.minimum(np.ceil(t_end), dur[active]).astype(np.int64))
        seg_s.append(s)
        t[active] = t_end
        u = rng.random(active.size)
        state[active] = np.minimum((u[:, None] > cum_jump[s]).sum(axis=1), n - 1)
        active = active[t_end < dur[active]]
    seg_b = np.concatenate(seg_b)
    seg_key = (seg_b << 32) | np.concatenate(seg_end)
    seg_s = np.concatenate(seg_s)
    order = np.argsort(seg_key, kind="stable")
    seg_key, seg_s = seg_key[order], seg_s[order]
    seg = np.searchsorted(seg_key, key, side="right")
    ph_state = seg_s[np.minimum(seg, len(seg_s) - 1)]

    is_acc = rng.random(total) < E[ph_state]
    mask_ad = is_acc.astype(np.uint8)
    mask_dd = (~is_acc).astype(np.uint8)
    if p_neither > 0:
        nei = rng.random(total) < p_neither
        mask_ad[nei] = 0
        mask_dd[nei] = 0
    return BurstData(start=start, stop=stop, istart=istart, phcount=P.astype(np.int32),
                     times=times.astype(np.int64), mask_ad=mask_ad, mask_dd=mask_dd, T=T,
                     clk_p=clk_p, bg_ad=bg_ad, bg_dd=bg_dd,
                     truth=dict(n=n, args=np.asarray(args, np.float32)))


def loguniform_rates(rng, count, lo, hi):
    return np.exp(rng.uniform(np.log(lo), np.log(hi), size=count))


# Photophysics of every synthetic configuration: gamma, DexDirAem, Dch2Ach, r0.
PHOTOPHYSICS = (1.0, 0.0, 0.05, 52.0)


def config(c, n_bursts=None):
    """Workload c (0..4): dict(data, n, args, R, sets, seeds)."""
    seeds = dict(data=1000 + c, observed=2000 + c, mc=3000 + c, params=4000 + c)
    prng = np.random.default_rng(seeds["params"])
    if c == 0:
        n, args = 2, make_args([0.30, 0.70], [1e3, 1e3])
        N, R, base, mg, med, sig = 5_000, 20, 50, 40.0, 1e-3, 0.5
    elif c in (1, 2, 3):
        prng1 = np.random.default_rng(4001)
        n, args = 3, make_args([0.15, 0.50, 0.85], loguniform_rates(prng1, 6, 1e2, 1e4))
        N = {1: 100_000, 2: 50_000, 3: 100_000}[c]
        R = {1: 100, 2: 32, 3: 64}[c]
        base, mg, med, sig = 50, 40.0, 1e-3, 0.5
    elif c == 4:
        n, args = 4, make_args([0.10, 0.35, 0.65, 0.90], loguniform_rates(prng, 12, 1e3, 1e5))
        N, R, base, mg, med, sig = 1_000_000, 32, 30, 80.0, 2e-3, 0.6
    else:
        raise ValueError(c)
    if n_bursts is not None:
        N = n_bursts
    data = generate(N, args, n, base_photons=base, mean_geom=mg, dur_median=med,
                    dur_sigma=sig, seed=seeds["data"])
    out = dict(data=data, n=n, args=args, R=R, seeds=seeds, photophysics=PHOTOPHYSICS)
    if c == 2:
        out["population"] = population(256, prng)
    if c == 4:
        out["population"] = population(64, prng, lo=1e3, hi=1e5, states=(4,))
    return out


def population(count, rng, lo=1e2, hi=1e4, states=(2, 3, 4)):
    """Random parameter sets: n ~ U(states), sorted E ~ U[0.05, 0.95], log-uniform rates."""
    out = []
    for _ in range(count):
        n = int(rng.choice(states))
        E = np.sort(rng.uniform(0.05, 0.95, size=n))
        out.append((n, make_args(E, loguniform_rates(rng, n * (n - 1), lo, hi))))
    return out
