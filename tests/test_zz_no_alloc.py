"""GPU: pda_run_noalloc, the evaluation without device allocation, equals
pda_run through the C ABI.

Two calls in a row on one slot, the second with a smaller input (fewer
bursts and resamples): each call's histogram, both chi2, bins kept, units
and replayed units equal those of pda_run with the same keys.  The first
workload abandons units (a 30 kHz background), so the replay list in the
fixed scratch array is exercised; the result goes to a device buffer the
caller owns (allocated here with the HIP runtime the library itself uses).
A call beyond PDA_NOALLOC_MAX_UNITS is refused before anything is launched.
"""
import ctypes

import numpy as np
import pytest

from gsmfretda_amd import pda, synth

pytestmark = pytest.mark.gpu

HDR = np.dtype([("chi2_64", "<f8"), ("chi2_ref", "<f4"), ("effN", "<i4"), ("units", "<i8"),
                ("nbins", "<i4"), ("replays", "<i4")])
B = 50


class DeviceBuffer:
    """Device memory of the caller, from the HIP runtime the library loaded."""

    def __init__(self, nbytes):
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                       ctypes.c_int]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.nbytes = nbytes
        self.ptr = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(self.ptr), nbytes) == 0
        self.write(np.full(nbytes, 0xAB, np.uint8))

    def write(self, a):
        assert self.hip.hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, 1) == 0  # host to device

    def read(self):
        out = np.zeros(self.nbytes, np.uint8)
        assert self.hip.hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, 2) == 0  # to host
        return out

    def free(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()


def _noalloc(m, buf, R, seed, set_id):
    m.run_noalloc(buf.ptr.value, buf.nbytes, R=R, seed=seed, set_id=set_id, sid=0)
    m.synchronize(0)
    raw = buf.read()
    return raw[:HDR.itemsize].view(HDR)[0], raw[HDR.itemsize:HDR.itemsize + 4 * B].view(np.int32)


def _run(m, R, seed, set_id):
    m.reSampleTimes = R
    m.run_kernel(sid=0, seed=seed, set_id=set_id)
    m.synchronize(0)
    hist, c64, c32, eff = m.get_hist_chi2(0)
    units, replays = m.run_stats(0)
    return hist, c64, c32, eff, units, replays


def _same(hdr, hist, want):
    h, c64, c32, eff, units, replays = want
    assert np.array_equal(hist, h)
    assert hdr["chi2_64"] == c64
    assert np.float32(hdr["chi2_ref"]).view(np.uint32) == np.float32(c32).view(np.uint32)
    assert hdr["effN"] == eff and hdr["units"] == units and hdr["nbins"] == B
    assert hdr["replays"] == replays


def test_noalloc_equals_run_twice_then_smaller(gpu_mc):
    m = gpu_mc
    cfg = synth.config(1, n_bursts=4000)
    d = cfg["data"]
    d.bg_dd, d.bg_ad = 3e4, 1.5e4  # a few per cent of units are replayed
    obs = np.random.default_rng(5).uniform(0, 1, d.nbursts).astype(np.float32)
    m.load(d, SgDivSr=obs, nbins=B, left_pad=2)
    m.set_photophysics(0, *cfg["photophysics"])
    assert m.set_params(cfg["n"], 0, cfg["args"])
    buf = DeviceBuffer(HDR.itemsize + 4 * B)

    m.set_burst_range(0, -1, 0)
    want = _run(m, 100, 71, 4)
    assert want[5] > 0, "the workload should abandon units"
    hdr, hist = _noalloc(m, buf, 100, 71, 4)
    _same(hdr, hist, want)

    m.set_burst_range(500, 1700, 0)
    want = _run(m, 37, 72, 9)
    hdr, hist = _noalloc(m, buf, 37, 72, 9)
    _same(hdr, hist, want)
    assert hdr["units"] == 1200 * 37

    # too many units for the fixed scratch: refused, nothing launched
    m.set_burst_range(0, -1, 0)
    with pytest.raises(pda.PdaError, match="at most"):
        m.run_noalloc(buf.ptr.value, buf.nbytes, R=pda_max_r(d.nbursts), seed=1, sid=0)
    with pytest.raises(pda.PdaError, match="result buffer"):
        m.run_noalloc(buf.ptr.value, 16, R=4, seed=1, sid=0)
    m.set_burst_range(0, -1, 0)
    buf.free()


def pda_max_r(nbursts):
    """Smallest R whose run exceeds the no-alloc limit of 16,000,000 units."""
    return 16_000_000 // nbursts + 1
