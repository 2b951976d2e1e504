"""ctypes wrapper of the CPU oracle (oracle/pda_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or the timed CPU baseline.  The
product package (gsmfretda_amd) never imports it.

The oracle restates gSMFRETda's mc_kernel (src/mc.cu:83-197) one resample at
a time on the keyed Philox4x32-10 streams that the HIP kernels also consume.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libpda_oracle.so")
MAXN = 8

_i64p = ctypes.POINTER(ctypes.c_int64)
_i32p = ctypes.POINTER(ctypes.c_int32)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)
_u32p = ctypes.POINTER(ctypes.c_uint32)


class OrData(ctypes.Structure):
    _fields_ = [
        ("nbursts", ctypes.c_int64),
        ("start", _i64p), ("stop", _i64p), ("istart", _i64p),
        ("phcount", _i32p),
        ("times", _i64p),
        ("mask_ad", _u8p), ("mask_dd", _u8p),
        ("T", _f32p),
        ("clk_p", ctypes.c_float), ("bg_ad", ctypes.c_float), ("bg_dd", ctypes.c_float),
    ]


class OrParams(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int),
        ("E", ctypes.c_float * MAXN), ("v", ctypes.c_float * MAXN), ("p0", ctypes.c_float * MAXN),
        ("K", ctypes.c_float * (MAXN * MAXN)), ("Pi2j", ctypes.c_float * (MAXN * MAXN)),
        ("gamma", ctypes.c_float), ("dexdir", ctypes.c_float),
        ("crosstalk", ctypes.c_float), ("r0", ctypes.c_float),
    ]


class OrStats(ctypes.Structure):
    _fields_ = [(name, ctypes.c_int64) for name in (
        "windows", "photons_visited", "philox_blocks",
        "binom_inv_iters", "binom_btpe_iters", "poisson_iters",
        "units", "live_windows", "live_philox_blocks", "search_steps", "binom_draws",
        "live_binom_inv_iters", "live_btpe_iters", "bg_draws", "live_poisson_iters")]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built: {LIB_PATH} (run oracle/Makefile)")
        L = ctypes.CDLL(LIB_PATH)
        L.or_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
        L.or_u01f.argtypes = [ctypes.c_uint32]; L.or_u01f.restype = ctypes.c_float
        L.or_u53.argtypes = [ctypes.c_uint32, ctypes.c_uint32]; L.or_u53.restype = ctypes.c_double
        for name, t in (("or_logf", ctypes.c_float), ("or_expf", ctypes.c_float),
                        ("or_log", ctypes.c_double), ("or_exp", ctypes.c_double),
                        ("or_cos2pif", ctypes.c_float)):
            getattr(L, name).argtypes = [t]
            getattr(L, name).restype = t
        L.or_normal.argtypes = [ctypes.c_uint32, ctypes.c_uint32]; L.or_normal.restype = ctypes.c_float
        L.or_categorical.argtypes = [ctypes.c_int, _f32p, ctypes.c_float]
        L.or_categorical.restype = ctypes.c_int
        L.or_binomial.argtypes = [ctypes.c_int, ctypes.c_double] + [ctypes.c_uint32] * 7 + [
            ctypes.POINTER(OrStats)]
        L.or_binomial.restype = ctypes.c_int64
        L.or_bg_capped.argtypes = [ctypes.c_float] * 4 + [ctypes.c_uint32] * 6 + [
            ctypes.POINTER(OrStats)]
        L.or_bg_capped.restype = ctypes.c_float
        L.or_set_params.argtypes = [ctypes.POINTER(OrParams), ctypes.c_int, _f32p, ctypes.c_int,
                                    ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float]
        L.or_set_params.restype = ctypes.c_int
        L.or_unit.argtypes = [ctypes.POINTER(OrData), ctypes.POINTER(OrParams), ctypes.c_int64,
                              ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                              _i32p, _f32p, _f32p, ctypes.POINTER(OrStats)]
        L.or_run.argtypes = [ctypes.POINTER(OrData), ctypes.POINTER(OrParams), ctypes.c_int64,
                             ctypes.c_int64, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32,
                             _i32p, _f32p, _f32p, ctypes.c_int, _i64p, ctypes.POINTER(OrStats)]
        L.or_run_threaded.argtypes = [ctypes.POINTER(OrData), ctypes.POINTER(OrParams),
                                      ctypes.c_int64, ctypes.c_int64, ctypes.c_int,
                                      ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, _i64p,
                                      ctypes.c_int]
        L.or_hist.argtypes = [_f32p, ctypes.c_int64, ctypes.c_int, _i64p]
        L.or_chi2.argtypes = [_f32p, _i64p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                              ctypes.c_int, _f32p, _f64p, ctypes.POINTER(ctypes.c_int)]
        L.or_rb_unit.argtypes = [ctypes.POINTER(OrData), ctypes.POINTER(OrParams), ctypes.c_int64,
                                 ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _i32p, _f32p,
                                 _f64p, _i32p]
        L.or_rb_run.argtypes = [ctypes.POINTER(OrData), ctypes.POINTER(OrParams), ctypes.c_int64,
                                ctypes.c_int64, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32,
                                ctypes.c_int, ctypes.POINTER(ctypes.c_uint64)]
        L.or_rb_grad.argtypes = [ctypes.POINTER(OrData), ctypes.POINTER(OrParams), ctypes.c_int64,
                                 ctypes.c_int64, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32,
                                 ctypes.c_int, _f64p, _f64p, _f64p]
        L.or_log_factorials.restype = _f64p
        L.or_state_efficiency.argtypes = [ctypes.c_float, ctypes.c_float]
        L.or_state_efficiency.restype = ctypes.c_float
        _lib = L
    return _lib


def _ptr(a, t):
    return a.ctypes.data_as(t)


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, np.uint32)
    lib().or_philox4x32_10(_ptr(c, _u32p), _ptr(k, _u32p), _ptr(out, _u32p))
    return out


class Data:
    """Holds numpy arrays alive and exposes them as an or_data struct."""

    def __init__(self, start, stop, istart, phcount, times, mask_ad, mask_dd, T,
                 clk_p, bg_ad, bg_dd):
        self.arrays = dict(
            start=np.ascontiguousarray(start, np.int64), stop=np.ascontiguousarray(stop, np.int64),
            istart=np.ascontiguousarray(istart, np.int64),
            phcount=np.ascontiguousarray(phcount, np.int32),
            times=np.ascontiguousarray(times, np.int64),
            mask_ad=np.ascontiguousarray(mask_ad, np.uint8),
            mask_dd=np.ascontiguousarray(mask_dd, np.uint8),
            T=np.ascontiguousarray(T, np.float32))
        a = self.arrays
        self.s = OrData(len(a["start"]), _ptr(a["start"], _i64p), _ptr(a["stop"], _i64p),
                        _ptr(a["istart"], _i64p), _ptr(a["phcount"], _i32p),
                        _ptr(a["times"], _i64p), _ptr(a["mask_ad"], _u8p),
                        _ptr(a["mask_dd"], _u8p), _ptr(a["T"], _f32p),
                        float(clk_p), float(bg_ad), float(bg_dd))

    @classmethod
    def from_bursts(cls, b):
        """From a gsmfretda_amd.synth.BurstData-like object (HDF5 field names)."""
        return cls(b.start, b.stop, b.istart, b.phcount, b.times, b.mask_ad, b.mask_dd,
                   b.T, b.clk_p, b.bg_ad, b.bg_dd)

    @property
    def nbursts(self):
        return int(self.s.nbursts)


def make_params(n, args, gamma=0.34, dexdir=0.08, crosstalk=0.07, r0=52.0):
    a = np.ascontiguousarray(args, np.float32)
    p = OrParams()
    rc = lib().or_set_params(ctypes.byref(p), int(n), _ptr(a, _f32p), len(a), gamma, dexdir,
                             crosstalk, r0)
    if rc != 0:
        raise ValueError(f"bad parameter set: n={n}, {len(a)} args")
    return p


def run(data, params, R, seed, set_id=0, b0=0, b1=None, B=50, want_units=True, stats=False):
    """Oracle over bursts [b0, b1) x R.  Returns dict(acc, F, E, hist[, stats])."""
    if b1 is None:
        b1 = data.nbursts
    nu = (b1 - b0) * R
    acc = np.zeros(nu if want_units else 0, np.int32)
    F = np.zeros_like(acc, dtype=np.float32)
    E = np.zeros_like(acc, dtype=np.float32)
    hist = np.zeros(B, np.int64)
    st = OrStats()
    lib().or_run(ctypes.byref(data.s), ctypes.byref(params), b0, b1, R, seed, set_id,
                 _ptr(acc, _i32p) if want_units else None,
                 _ptr(F, _f32p) if want_units else None,
                 _ptr(E, _f32p) if want_units else None, B, _ptr(hist, _i64p),
                 ctypes.byref(st) if stats else None)
    out = dict(hist=hist)
    if want_units:
        out.update(acc=acc, F=F, E=E)
    if stats:
        out["stats"] = st.as_dict()
    return out


def run_threaded(data, params, R, seed, set_id=0, b0=0, b1=None, B=50, nthreads=None):
    if b1 is None:
        b1 = data.nbursts
    if nthreads is None:
        nthreads = os.cpu_count() or 1
    hist = np.zeros(B, np.int64)
    lib().or_run_threaded(ctypes.byref(data.s), ctypes.byref(params), b0, b1, R, seed, set_id,
                          B, _ptr(hist, _i64p), int(nthreads))
    return hist


def hist(x, B=50):
    x = np.ascontiguousarray(x, np.float32)
    h = np.zeros(B, np.int64)
    lib().or_hist(_ptr(x, _f32p), len(x), B, _ptr(h, _i64p))
    return h


def chi2(O, M, R, nstates, left_pad=2):
    O = np.ascontiguousarray(O, np.float32)
    M = np.ascontiguousarray(M, np.int64)
    c32 = ctypes.c_float()
    c64 = ctypes.c_double()
    eff = ctypes.c_int()
    lib().or_chi2(_ptr(O, _f32p), _ptr(M, _i64p), len(O), left_pad, R, nstates,
                  ctypes.byref(c32), ctypes.byref(c64), ctypes.byref(eff))
    return c32.value, c64.value, eff.value


def rb_unit(data, params, burst, r, seed, set_id=0, transitions=False):
    """Walk summary of one unit for the Rao-Blackwellised form:
    (photon counts per state, F, scores of the rates in args order
    [, windows per transition in args order])."""
    n = params.n
    cnt = np.zeros(n, np.int32)
    F = ctypes.c_float()
    S = np.zeros(max(n * (n - 1), 1), np.float64)
    N = np.zeros(max(n * (n - 1), 1), np.int32)
    lib().or_rb_unit(ctypes.byref(data.s), ctypes.byref(params), burst, r, seed, set_id,
                     _ptr(cnt, _i32p), ctypes.byref(F), _ptr(S, _f64p), _ptr(N, _i32p))
    if transitions:
        return cnt, F.value, S[:n * (n - 1)], N[:n * (n - 1)]
    return cnt, F.value, S[:n * (n - 1)]


def rb_run(data, params, R, seed, set_id=0, b0=0, b1=None, B=50):
    """Expected histogram of the RB form in 2^-32 units (float path)."""
    if b1 is None:
        b1 = data.nbursts
    M = np.zeros(B, np.uint64)
    lib().or_rb_run(ctypes.byref(data.s), ctypes.byref(params), b0, b1, R, seed, set_id, B,
                    M.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
    return M


def rb_grad(data, params, R, seed, g, set_id=0, b0=0, b1=None):
    """FP64 expected histogram M and the gradient of sum_b g_b M_b in the
    args (E then the off-diagonal rates in args order)."""
    if b1 is None:
        b1 = data.nbursts
    g = np.ascontiguousarray(g, np.float64)
    n = params.n
    M = np.zeros(len(g), np.float64)
    grad = np.zeros(n * n, np.float64)
    lib().or_rb_grad(ctypes.byref(data.s), ctypes.byref(params), b0, b1, R, seed, set_id,
                     len(g), _ptr(g, _f64p), _ptr(M, _f64p), _ptr(grad, _f64p))
    return M, grad


def log_factorials(count):
    p = lib().or_log_factorials()
    return np.array([p[i] for i in range(count + 1)])
