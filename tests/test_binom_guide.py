"""The fast walk's guided binomial draw (pda_walk.h binomial_guided /
build_guide_row) against the inversion loop it replaces (binomial_table, the
loop of src/binom.cuh:135-169), compiled for the host.

The guided draw must give the loop's X for every uniform, or -1 (the unit
replays in the general form, which runs the loop) only inside the rounding
band around a threshold.  Checked at every threshold pair +- 3 grid steps and
at random grid points, for every state and photon count of several models;
the loop's exact crossing points, found by bisection on the loop itself, must
lie inside the bands (a band that missed its crossing, or an off-by-one in
up/lo, fails here).
"""
import ctypes

import numpy as np
import pytest

from gsmfretda_amd import synth
from tests.conftest import build_hostwalk
from tests.test_host_walk import run_host

_P = ctypes.POINTER


@pytest.fixture(scope="module")
def hw():
    L = ctypes.CDLL(build_hostwalk())
    L.hw_error.restype = ctypes.c_char_p
    L.hw_replays.restype = ctypes.c_int64
    return L


def _check(L, n, args, ph, n_rand=1000, seed=7):
    a = np.ascontiguousarray(args, np.float32)
    p = np.asarray(ph, np.float32)
    out = np.zeros(8, np.int64)
    rc = L.hw_guide_check(n, a.ctypes.data_as(_P(ctypes.c_float)), p.ctypes.data_as(_P(ctypes.c_float)),
                          n_rand, ctypes.c_uint64(seed), out.ctypes.data_as(_P(ctypes.c_int64)))
    assert rc == 0, L.hw_error().decode()
    return out


MODELS = {
    "3-state headline": (3, synth.config(1, n_bursts=10)["args"]),
    "4-state fast dynamics": (4, synth.config(4, n_bursts=10)["args"]),
    "E near 0 and 1": (2, synth.make_args([0.02, 0.98], [1e3, 1e3])),
    "E = 1/2": (2, synth.make_args([0.5, 0.5], [1e3, 2e3])),
    "E near 1/2 from both sides": (3, synth.make_args([0.49, 0.51, 0.3], [1e3] * 6)),
}


@pytest.mark.parametrize("name", list(MODELS))
def test_guided_draw_is_the_inversion_loop(hw, name):
    n, args = MODELS[name]
    out = _check(hw, n, args, synth.PHOTOPHYSICS)
    compared, differ, band, loop_replay_only, crossing_outside, longest, guide_bad, rnd_band = out
    assert compared > 100_000
    assert differ == 0            # both give an X: the same X
    assert loop_replay_only == 0  # the loop's X > bound is always a replay of the guided draw
    assert crossing_outside == 0  # every exact crossing lies in (lo_k, up_k]
    assert guide_bad == 0
    assert rnd_band == 0          # a random uniform falls in a band with probability ~1e-14
    assert 0 < band < compared    # the probes inside the bands are replays
    assert longest <= 128


def test_host_walk_guided_equals_loop(hw):
    """Whole units: the fast walk with guided draws and with the loop give the
    same (acc, F) for every unit, and the same replays."""
    c = synth.config(1, n_bursts=400)
    d = c["data"]
    res = []
    for guided in (1, 0):
        hw.hw_set_guided(guided)
        try:
            r0 = hw.hw_replays()
            acc, F, E = run_host(hw, d, c["n"], c["args"], c["photophysics"], 16, seed=11)
            res.append((acc, F, E, hw.hw_replays() - r0))
        finally:
            hw.hw_set_guided(1)
    (a1, f1, e1, r1), (a0, f0, e0, r0) = res
    assert np.array_equal(a1, a0)
    assert np.array_equal(f1.view(np.uint32), f0.view(np.uint32))
    assert np.array_equal(e1.view(np.uint32), e0.view(np.uint32))
    assert r1 == r0


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(MODELS))
def test_device_rows_equal_host_rows(hw, name):
    """The rows guide_kernel builds on the GPU equal build_guide_row_host's
    bit for bit (same operations, same order), so the equivalence checks
    above hold for the rows the GPU walk reads."""
    from gsmfretda_amd import pda
    n, args = MODELS[name]
    a = np.ascontiguousarray(args, np.float32)
    ph = np.asarray(synth.PHOTOPHYSICS, np.float32)
    h_thr = np.zeros((8, 8384, 2), np.int64)
    h_guide = np.zeros((8, 128, 64), np.uint8)
    assert hw.hw_build_guide(n, a.ctypes.data_as(_P(ctypes.c_float)), ph.ctypes.data_as(_P(ctypes.c_float)),
                             h_thr.ctypes.data_as(_P(ctypes.c_int64)),
                             h_guide.ctypes.data_as(_P(ctypes.c_uint8))) == 0
    m = pda.MC(0, 1)
    try:
        m.set_photophysics(0, *synth.PHOTOPHYSICS)
        assert m.set_params(n, 0, a)
        d_thr, d_guide = m.debug_guide(0)
    finally:
        m.close()
    rows = 0
    for s in range(n):
        for mm in range(128):
            off = mm * (mm + 3) // 2
            h_row, d_row = h_thr[s, off:off + mm + 2], d_thr[s, off:off + mm + 2]
            end = int(np.argmax(h_row[:, 0] == np.iinfo(np.int64).max))
            assert end > 0
            assert np.array_equal(d_row[:end + 1], h_row[:end + 1]), (s, mm)
            assert np.array_equal(d_guide[s, mm], h_guide[s, mm]), (s, mm)
            rows += 1
    assert rows == n * 128


def test_poisson_peeled_steps_equal_the_loop(hw):
    """background_single (first two inversion steps without a branch) equals
    the plain inversion loop bitwise: at random uniforms, at every cumulative
    sum of the first steps and its float neighbours, and on long loops (past
    k = 255 and into subnormal pk: uniforms next to 1, caps of 300 and
    1000)."""
    out = np.zeros(4, np.int64)
    assert hw.hw_poisson_check(200_000, 2_000, ctypes.c_uint64(3), out.ctypes.data_as(_P(ctypes.c_int64))) == 0
    compared, differ, past_255, subnormal = out
    assert compared > 3_000_000
    assert differ == 0
    assert past_255 > 100
    assert subnormal > 100


def test_unit_division_by_multiply_and_shift(hw):
    """unit_div (a multiply by ceil(2^(31+l) / R) and a shift, pda_walk.h),
    which the walks use for the burst of a unit index, equals u / R for every
    R up to 4096 and random R up to 2^32 - 1, at the edges (0, R - 1, R, the
    multiples of R next to 2^31, 2^31 - 1) and at random u below 2^31."""
    out = np.zeros(2, np.int64)
    assert hw.hw_unit_div_check(20_000, 200, ctypes.c_uint64(9), out.ctypes.data_as(_P(ctypes.c_int64))) == 0
    assert out[0] > 4_000_000
    assert out[1] == 0
