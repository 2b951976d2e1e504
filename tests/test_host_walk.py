"""The product's walk algorithm compiled for the host (tests/host/host_walk.cpp:
pda_walk.h + pda_layout.cpp) against the oracle's literal scan, bit for bit.

This checks, without a GPU, the parts of the HIP path that differ from the
reference's algorithm: the burst-relative re-encoding (re-timed neither
photons, channel-run links), the binary-searched window ends, the skipped
draws (n == 2 next state, v == 0 broadening) and the early exit once the
cursor reaches the last photon.
"""
import ctypes

import numpy as np
import pytest

from gsmfretda_amd import synth
from tests.conftest import bits, build_hostwalk
from tests.test_oracle_walk import tiny_bursts

_P = ctypes.POINTER


@pytest.fixture(scope="module")
def hostwalk():
    L = ctypes.CDLL(build_hostwalk())
    L.hw_error.restype = ctypes.c_char_p
    L.hw_replays.restype = ctypes.c_int64
    return L


def _ptr(a, t):
    return a.ctypes.data_as(_P(t))


def run_host(L, d, n, args, ph, R, seed, set_id=0, dump=False):
    nu = d.nbursts * R
    acc = np.zeros(nu, np.int32)
    F = np.zeros(nu, np.float32)
    E = np.zeros(nu, np.float32)
    a = np.ascontiguousarray(args, np.float32)
    phv = np.asarray(ph, np.float32)
    arrs = [np.ascontiguousarray(x, t) for x, t in (
        (d.start, np.int64), (d.stop, np.int64), (d.istart, np.int64), (d.phcount, np.int32),
        (d.times, np.int64), (d.mask_ad, np.uint8), (d.mask_dd, np.uint8), (d.T, np.float32))]
    link = np.zeros(3 * max(int(d.phcount.sum()), 1), np.uint32) if dump else None
    rc = L.hw_run(_ptr(arrs[0], ctypes.c_int64), _ptr(arrs[1], ctypes.c_int64),
                  _ptr(arrs[2], ctypes.c_int64), _ptr(arrs[3], ctypes.c_int32),
                  ctypes.c_int64(d.nbursts), _ptr(arrs[4], ctypes.c_int64),
                  _ptr(arrs[5], ctypes.c_uint8), _ptr(arrs[6], ctypes.c_uint8),
                  ctypes.c_int64(len(arrs[4])), _ptr(arrs[7], ctypes.c_float),
                  ctypes.c_float(d.clk_p), ctypes.c_float(d.bg_ad), ctypes.c_float(d.bg_dd),
                  ctypes.c_int(n), _ptr(a, ctypes.c_float), _ptr(phv, ctypes.c_float),
                  ctypes.c_int64(0), ctypes.c_int64(d.nbursts), ctypes.c_int(R),
                  ctypes.c_uint32(seed), ctypes.c_uint32(set_id), _ptr(acc, ctypes.c_int32),
                  _ptr(F, ctypes.c_float), _ptr(E, ctypes.c_float),
                  _ptr(link, ctypes.c_uint32) if dump else None)
    if rc != 0:
        raise ValueError(L.hw_error().decode())
    return (acc, F, E, link) if dump else (acc, F, E)


def compare(oracle, L, d, n, args, ph, R, seed, set_id=0):
    acc, F, E = run_host(L, d, n, args, ph, R, seed, set_id)
    o = oracle.run(oracle.Data.from_bursts(d), oracle.make_params(n, args, *ph), R=R, seed=seed,
                   set_id=set_id)
    assert np.array_equal(acc, o["acc"]), f"{(acc != o['acc']).sum()} units differ"
    assert np.array_equal(bits(F), bits(o["F"]))
    assert np.array_equal(bits(E), bits(o["E"]))


@pytest.mark.parametrize("c", [0, 1, 4])
def test_config_subsets(oracle, hostwalk, c):
    cfg = synth.config(c, n_bursts=600 if c != 4 else 300)
    compare(oracle, hostwalk, cfg["data"], cfg["n"], cfg["args"], cfg["photophysics"],
            min(cfg["R"], 16), cfg["seeds"]["mc"])


@pytest.mark.parametrize("seed", range(4))
def test_tiny_bursts_all_branches(oracle, hostwalk, seed):
    rng = np.random.default_rng(50 + seed)
    d = tiny_bursts(rng, 60)
    d.bg_dd = [0.0, 1e3, 1e6, 2e5][seed]
    d.bg_ad = [0.0, 5e2, 3e6, 1e5][seed]
    n = 2 + seed
    E = np.sort(rng.uniform(0, 1, n))
    rates = np.exp(rng.uniform(np.log(1e3), np.log(1e6), n * (n - 1)))
    if seed == 3:
        rates[2] = 0.0
    v = np.where(rng.random(n) < 0.5, rng.uniform(0, 8, n), 0.0)
    before = hostwalk.hw_replays()
    compare(oracle, hostwalk, d, n, synth.make_args(E, rates, v), (0.34, 0.08, 0.07, 52.0), 9,
            1000 + seed, set_id=seed)
    if seed == 2 and not v.any():
        assert hostwalk.hw_replays() > before   # background means > 64 leave the fast form


def test_eight_states_and_long_runs(oracle, hostwalk):
    # long acceptor runs in a state with E near 1 reach the BTPE regime
    args = synth.make_args([0.05, 0.2, 0.35, 0.5, 0.65, 0.8, 0.9, 0.97],
                           np.full(56, 150.0), np.zeros(8))
    d = synth.generate(200, synth.make_args([0.97, 0.97], [10.0, 10.0]), 2, base_photons=300,
                       mean_geom=200, seed=3)
    before = hostwalk.hw_replays()
    compare(oracle, hostwalk, d, 8, args, (1.0, 0.0, 0.0, 52.0), 12, 5)
    assert hostwalk.hw_replays() > before           # BTPE draws are replayed


def test_background_replays(oracle, hostwalk):
    rng = np.random.default_rng(52)
    d = tiny_bursts(rng, 60)
    d.bg_dd, d.bg_ad = 1e6, 3e6
    before = hostwalk.hw_replays()
    compare(oracle, hostwalk, d, 4, synth.make_args([0.1, 0.3, 0.6, 0.9], np.full(12, 1e4)),
            (0.34, 0.08, 0.07, 52.0), 9, 1002)
    assert hostwalk.hw_replays() > before


@pytest.mark.parametrize("seed", range(3))
def test_windows_past_the_stop(oracle, hostwalk, seed):
    # The last window of a unit ends at the burst's stop tick.  Stops before
    # the last photons and durations T above (stop - start) clk_p let earlier
    # windows end past the stop, so the last window's search starts beyond
    # the stop's lower bound (8-36 such final windows per case here).
    rng = np.random.default_rng(70 + seed)
    d = tiny_bursts(rng, 80)
    dur = d.stop - d.start + 1
    d.stop = d.start + (dur * rng.uniform(0.1, 1.0, d.nbursts)).astype(np.int64)
    d.T = (d.T * np.float32([1.0, 1.2, 1.0][seed])).astype(np.float32)
    d.bg_dd, d.bg_ad = [0.0, 1e4, 1e3][seed], [0.0, 5e3, 1e3][seed]
    n = 3
    args = synth.make_args([0.2, 0.5, 0.8], np.full(6, 2e5 * (seed + 1)))
    compare(oracle, hostwalk, d, n, args, (0.34, 0.08, 0.07, 52.0), 12, 300 + seed)


def test_link_encoding_brute_force(hostwalk):
    rng = np.random.default_rng(8)
    d = tiny_bursts(rng, 40)
    args = synth.make_args([0.3, 0.7], [1e3, 1e3])
    *_, link = run_host(hostwalk, d, 2, args, (1.0, 0.0, 0.0, 52.0), 1, 1, dump=True)
    link = link.reshape(-1, 3)
    off = 0
    for b in range(d.nbursts):
        P = int(d.phcount[b])
        base = int(d.istart[b])
        col = [int(d.mask_dd[base + i]) | (int(d.mask_ad[base + i]) << 1) for i in range(P)]
        t = [int(d.times[base + i] - d.start[b]) for i in range(P)]
        for i in range(P):
            f = next((j for j in range(i, P) if col[j]), P)
            tick, lx, ly = (int(x) for x in link[off + i])
            assert (lx & 0x3FFFFFFF) == f
            if f < P:
                assert lx >> 30 == col[f]
                other = {1: 2, 2: 1, 3: 3}[col[f]]
                run_end = f if col[f] == 3 else next(
                    (j for j in range(f + 1, P) if col[j] & other), P)
                assert ly == run_end
                assert tick == t[f]              # neither photons carry the next channel tick
            else:
                assert ly == P and tick == 0x80000000
        off += P


def test_rejects_bad_data(hostwalk):
    d = synth.generate(10, synth.make_args([0.3, 0.7], [1e3, 1e3]), 2, seed=1)
    bad = synth.BurstData(**{**d.__dict__, "times": d.times.copy()})
    bad.times[int(d.istart[3]) + 1] = bad.times[int(d.istart[3])] - 5   # out of time order
    with pytest.raises(ValueError, match="out of time order"):
        run_host(hostwalk, bad, 2, synth.make_args([0.3, 0.7], [1e3, 1e3]), (1, 0, 0, 52), 1, 1)
    bad = synth.BurstData(**{**d.__dict__, "times": d.times.copy()})
    bad.times[int(d.istart[2])] = d.start[2] - 1                        # before the burst start
    with pytest.raises(ValueError, match="outside"):
        run_host(hostwalk, bad, 2, synth.make_args([0.3, 0.7], [1e3, 1e3]), (1, 0, 0, 52), 1, 1)
    with pytest.raises(ValueError, match="rate"):
        run_host(hostwalk, d, 2, synth.make_args([0.3, 0.7], [-1.0, 1e3]), (1, 0, 0, 52), 1, 1)
    with pytest.raises(ValueError, match="outside"):
        run_host(hostwalk, d, 2, synth.make_args([0.3, 1.7], [1.0, 1e3]), (1, 0, 0, 52), 1, 1)
