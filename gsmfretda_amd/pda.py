"""Python host interface of the MI355X Monte-Carlo PDA path (ctypes over the C ABI).

`MC` mirrors gSMFRETda's `class mc` (src/mc.hpp:32-180) method for method, so
tests and callers read like the reference's gpuWorker loop
(src/gpuWorker.cpp:69-83):

    m = MC(0, 1)
    m.init_data_gpu(...)             # bursts resident in HBM, once
    m.set_params(n, sid, args)       # E, K off-diagonals, v
    m.setBurstBd(begin, end, sid)
    m.run_kernel(N, sid)             # async on the GPU
    while not m.streamQuery(sid): ...
    m.get_hist_chi2(sid)             # int histogram + chi2, computed on device

The shared library must exist (gsmfretda_amd/libgsmfretda_pda.so, built by
gsmfretda_amd.build); there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgsmfretda_pda.so")

PDA_RUN_SAMPLES = 1
PDA_RUN_NO_GRAPH = 2
PDA_RB_SCORES = 1
MAX_STATES = 8
MAX_BINS = 1024

# Symbols declared in include/gsmfretda_pda.h.
EXPORTS = (
    "pda_last_error", "pda_version", "pda_create", "pda_destroy", "pda_num_streams",
    "pda_stream", "pda_load_bursts", "pda_load_bursts_tagged", "pda_num_bursts",
    "pda_set_observed", "pda_set_observed_hist", "pda_set_params", "pda_set_photophysics",
    "pda_set_burst_range", "pda_run", "pda_query", "pda_synchronize", "pda_get_result",
    "pda_run_stats", "pda_get_samples", "pda_samples_host", "pda_last_run_ms", "pda_last_pass_ms",
    "pda_set_launch_shape", "pda_get_launch_shape", "pda_autotune",
    "pda_set_population", "pda_run_population", "pda_get_population_result",
    "pda_population_stats", "pda_rb_forward", "pda_rb_get_result", "pda_rb_adjoint",
    "pda_rb_jacobian", "pda_rb_stats", "pda_rb_forward_population",
    "pda_rb_get_population_result", "pda_rb_adjoint_population", "pda_rb_jacobian_population",
    "pda_run_noalloc", "pda_debug_primitive",
    "pda_debug_microbench", "pda_debug_guide",
)


class PdaError(RuntimeError):
    pass


_i64p = ctypes.POINTER(ctypes.c_int64)
_i32p = ctypes.POINTER(ctypes.c_int32)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p

_lib = None


def load_library(path=LIB_PATH):
    """Loads the HIP library; raises if it was not built (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise PdaError(f"{path} is missing: build it with gsmfretda_amd.build.build()")
    L = ctypes.CDLL(path)
    L.pda_last_error.restype = ctypes.c_char_p
    L.pda_version.restype = ctypes.c_char_p
    L.pda_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)]
    L.pda_destroy.argtypes = [_vp]
    L.pda_num_streams.argtypes = [_vp]
    L.pda_stream.argtypes = [_vp, ctypes.c_int]
    L.pda_stream.restype = _vp
    L.pda_load_bursts.argtypes = [_vp, _i64p, _i64p, _i64p, _i32p, ctypes.c_int64, _i64p, _u8p,
                                  _u8p, ctypes.c_int64, _f32p, ctypes.c_float, ctypes.c_float,
                                  ctypes.c_float]
    L.pda_load_bursts_tagged.argtypes = [_vp, _i64p, _i64p, _i64p, _i32p, ctypes.c_int64, _i64p,
                                         _i64p, ctypes.c_int64, _f32p, ctypes.c_float,
                                         ctypes.c_float, ctypes.c_float]
    L.pda_num_bursts.argtypes = [_vp]
    L.pda_num_bursts.restype = ctypes.c_int64
    L.pda_set_observed.argtypes = [_vp, _f32p, ctypes.c_int64, ctypes.c_int, ctypes.c_int]
    L.pda_set_observed_hist.argtypes = [_vp, _f32p, ctypes.c_int, ctypes.c_int]
    L.pda_set_params.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _f32p, ctypes.c_int]
    L.pda_set_photophysics.argtypes = [_vp, ctypes.c_int] + [ctypes.c_float] * 4
    L.pda_set_burst_range.argtypes = [_vp, ctypes.c_int, ctypes.c_int64, ctypes.c_int64, _i64p]
    L.pda_run.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32,
                          ctypes.c_int]
    L.pda_query.argtypes = [_vp, ctypes.c_int]
    L.pda_synchronize.argtypes = [_vp, ctypes.c_int]
    L.pda_get_result.argtypes = [_vp, ctypes.c_int, _i32p, ctypes.c_int, _f64p, _f32p,
                                 ctypes.POINTER(ctypes.c_int)]
    L.pda_get_samples.argtypes = [_vp, ctypes.c_int, _f32p, _i32p, _f32p, ctypes.c_int64]
    L.pda_run_stats.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int64),
                                ctypes.POINTER(ctypes.c_int32)]
    L.pda_samples_host.argtypes = [_vp, ctypes.c_int]
    L.pda_samples_host.restype = _f32p
    L.pda_last_run_ms.argtypes = [_vp, ctypes.c_int, _f32p]
    L.pda_last_pass_ms.argtypes = [_vp, ctypes.c_int, _f32p, _f32p]
    L.pda_debug_primitive.argtypes = [_vp, ctypes.c_int, _vp, _vp, ctypes.c_int64]
    L.pda_rb_forward_population.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_uint32,
                                            ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int]
    L.pda_rb_get_population_result.argtypes = [_vp, ctypes.c_int,
                                               ctypes.POINTER(ctypes.c_uint64), _f64p,
                                               ctypes.c_int, _f64p, _i32p]
    L.pda_rb_adjoint_population.argtypes = [_vp, ctypes.c_int, _f64p, ctypes.c_int, _f64p,
                                            ctypes.c_int]
    L.pda_rb_jacobian_population.argtypes = [_vp, ctypes.c_int, _i32p, _f64p, ctypes.c_int,
                                             ctypes.c_int]
    L.pda_run_noalloc.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_uint32,
                                  ctypes.c_uint32, ctypes.c_void_p, ctypes.c_int64]
    L.pda_debug_microbench.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                       ctypes.POINTER(ctypes.c_float),
                                       ctypes.POINTER(ctypes.c_int64),
                                       ctypes.POINTER(ctypes.c_int64)]
    L.pda_debug_guide.argtypes = [_vp, ctypes.c_int, _vp, ctypes.c_int64, _vp, ctypes.c_int64]
    _ip = ctypes.POINTER(ctypes.c_int)
    L.pda_set_launch_shape.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.pda_get_launch_shape.argtypes = [_vp, _ip, _ip, _ip]
    L.pda_autotune.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _ip,
                               _ip, _ip, _f32p]
    L.pda_set_population.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _i32p, _f32p,
                                     ctypes.c_int64]
    L.pda_run_population.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_uint32,
                                     ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int]
    L.pda_get_population_result.argtypes = [_vp, ctypes.c_int, _i32p, ctypes.c_int, _f64p,
                                            _f32p, _i32p]
    L.pda_population_stats.argtypes = [_vp, ctypes.c_int, _i64p, _i32p, _i32p, _f32p]
    _u64p = ctypes.POINTER(ctypes.c_uint64)
    L.pda_rb_forward.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_uint32,
                                 ctypes.c_uint32, ctypes.c_int]
    L.pda_rb_get_result.argtypes = [_vp, ctypes.c_int, _u64p, _f64p, ctypes.c_int, _f64p, _i32p]
    L.pda_rb_adjoint.argtypes = [_vp, ctypes.c_int, _f64p, ctypes.c_int, _f64p, ctypes.c_int]
    L.pda_rb_jacobian.argtypes = [_vp, ctypes.c_int, _f64p, ctypes.c_int, ctypes.c_int]
    L.pda_rb_stats.argtypes = [_vp, ctypes.c_int, _f32p, _f32p, _f32p]
    _lib = L
    return L


def _check(rc):
    if rc < 0:
        raise PdaError(f"pda error {rc}: {_lib.pda_last_error().decode()}")
    return rc


def _p(a, t):
    return a.ctypes.data_as(t)


class MC:
    """The reference's `class mc` (src/mc.hpp:32-180) over libgsmfretda_pda.so."""

    def __init__(self, devid=0, streamNum=1):
        L = load_library()
        h = _vp()
        _check(L.pda_create(devid, streamNum, ctypes.byref(h)))
        self._h = h
        self._L = L
        self.devid = devid
        self.streamNum = L.pda_num_streams(h)
        self.reSampleTimes = 4          # src/mc.cu:308
        self.s_n = [0] * self.streamNum
        self.sz_burst = 0
        self.seed = 1234
        self._N = [0] * self.streamNum
        self._hist_bins = 200           # fret_hist_num default (cmdline.ggo)
        self._left_pad = 2               # src/main.cpp:155
        self._rb_runs = [0] * self.streamNum  # RB forwards started per slot

    # -- lifetime ------------------------------------------------------------
    def close(self):
        if self._h:
            self._L.pda_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    def getStreams(self):
        return [self._L.pda_stream(self._h, s) for s in range(self.streamNum)]

    def set_gpuid(self):
        pass

    # -- data ------------------------------------------------------------------
    def init_data_gpu(self, istart, start, stop, phCount, times, mask_ad, mask_dd, T, SgDivSr,
                      clk_p, bg_ad, bg_dd, nbins=None, left_pad=None):
        """mc::init_data_gpu (src/mc.cu:456-521) from the HDF5 datasets
        (times + masks instead of the burst_ad/burst_dd encoding)."""
        a = dict(istart=np.ascontiguousarray(istart, np.int64),
                 start=np.ascontiguousarray(start, np.int64),
                 stop=np.ascontiguousarray(stop, np.int64),
                 ph=np.ascontiguousarray(phCount, np.int32),
                 times=np.ascontiguousarray(times, np.int64),
                 mad=np.ascontiguousarray(mask_ad, np.uint8),
                 mdd=np.ascontiguousarray(mask_dd, np.uint8),
                 T=np.ascontiguousarray(T, np.float32))
        _check(self._L.pda_load_bursts(
            self._h, _p(a["start"], _i64p), _p(a["stop"], _i64p), _p(a["istart"], _i64p),
            _p(a["ph"], _i32p), len(a["start"]), _p(a["times"], _i64p), _p(a["mad"], _u8p),
            _p(a["mdd"], _u8p), len(a["times"]), _p(a["T"], _f32p), clk_p, bg_ad, bg_dd))
        self.sz_burst = len(a["start"])
        if SgDivSr is not None:
            self.set_observed(SgDivSr, nbins or self._hist_bins,
                              self._left_pad if left_pad is None else left_pad)

    def init_data_gpu_tagged(self, istart, start, stop, phCount, burst_ad, burst_dd, T,
                             SgDivSr, clk_p, bg_ad, bg_dd):
        """mc::init_data_gpu with the reference's exact photon encoding."""
        a = [np.ascontiguousarray(x, t) for x, t in (
            (istart, np.int64), (start, np.int64), (stop, np.int64), (phCount, np.int32),
            (burst_ad, np.int64), (burst_dd, np.int64), (T, np.float32))]
        _check(self._L.pda_load_bursts_tagged(
            self._h, _p(a[0], _i64p), _p(a[1], _i64p), _p(a[2], _i64p), _p(a[3], _i32p),
            len(a[1]), _p(a[4], _i64p), _p(a[5], _i64p), len(a[4]), _p(a[6], _f32p),
            clk_p, bg_ad, bg_dd))
        self.sz_burst = len(a[1])
        if SgDivSr is not None:
            self.set_observed(SgDivSr, self._hist_bins, self._left_pad)

    def load(self, bursts, SgDivSr=None, nbins=50, left_pad=2):
        """Convenience: load a gsmfretda_amd.synth.BurstData."""
        self.init_data_gpu(bursts.istart, bursts.start, bursts.stop, bursts.phcount,
                           bursts.times, bursts.mask_ad, bursts.mask_dd, bursts.T, SgDivSr,
                           bursts.clk_p, bursts.bg_ad, bursts.bg_dd, nbins, left_pad)

    def load_hdf5(self, path, nbins=None, left_pad=2):
        """Bursts and SgDivSr from a file in the reference's HDF5 schema
        (src/loadHdf5.cpp:37-74), loaded as init_data_gpu does; returns the
        BurstData."""
        from .hdf5 import read_bursts
        d, sg = read_bursts(path)
        self.load(d, SgDivSr=sg, nbins=nbins or self._hist_bins, left_pad=left_pad)
        return d

    def set_observed(self, SgDivSr, nbins=50, left_pad=2):
        x = np.ascontiguousarray(SgDivSr, np.float32)
        _check(self._L.pda_set_observed(self._h, _p(x, _f32p), len(x), nbins, left_pad))
        self._hist_bins = nbins
        self._left_pad = left_pad

    def set_observed_hist(self, ohist, left_pad=2):
        h = np.ascontiguousarray(ohist, np.float32)
        _check(self._L.pda_set_observed_hist(self._h, _p(h, _f32p), len(h), left_pad))
        self._hist_bins = len(h)
        self._left_pad = left_pad

    def get_max_conversion_capacity(self, max_stateNum=8):
        return self.sz_burst > 0

    # -- per parameter set -------------------------------------------------------
    def set_nstates(self, n, sid):
        prev = self.s_n[sid]
        self.s_n[sid] = n
        return prev if prev != n else n

    def set_params(self, n, sid, args):
        """mc::set_params: args = [E(n), K off-diagonals (n(n-1)), v(n)].
        Returns False on a bad argument count, like the reference."""
        a = np.ascontiguousarray(args, np.float32)
        rc = self._L.pda_set_params(self._h, sid, n, _p(a, _f32p), len(a))
        if rc == -1:
            return False
        _check(rc)
        self.s_n[sid] = n
        return True

    def set_photophysics(self, sid, gamma=0.34, dexdir=0.08, crosstalk=0.07, r0=52.0):
        _check(self._L.pda_set_photophysics(self._h, sid, gamma, dexdir, crosstalk, r0))

    def setBurstBd(self, cstart, cstop, sid):
        """mc::setBurstBd (src/mc.cu:523-535), including its range rules."""
        rcstart, rcstop = cstart, cstop
        if cstop >= self.sz_burst or cstop < 0:
            rcstop = self.sz_burst
        if cstart >= rcstop:
            rcstart = 0
        N = rcstop - rcstart
        if cstart == -1:
            return -1
        _check(self._L.pda_set_burst_range(self._h, sid, rcstart, rcstop, None))
        self._N[sid] = N
        return N

    def set_burst_range(self, begin, end, sid):
        n = ctypes.c_int64()
        _check(self._L.pda_set_burst_range(self._h, sid, begin, end, ctypes.byref(n)))
        self._N[sid] = n.value
        return n.value

    def init_randstate(self, N, sid):
        pass  # stateless Philox streams: nothing to initialise

    def set_params_buff(self, oldS_n, N, sid):
        pass

    def run_kernel(self, N=None, sid=0, seed=None, set_id=None, samples=False):
        """mc::run_kernel: async launch of reSampleTimes resamples per burst."""
        flags = PDA_RUN_SAMPLES if samples else 0
        s = self.seed if seed is None else seed
        k = sid if set_id is None else set_id
        _check(self._L.pda_run(self._h, sid, self.reSampleTimes, s & 0xFFFFFFFF,
                               k & 0xFFFFFFFF, flags))

    def streamQuery(self, sid):
        return _check(self._L.pda_query(self._h, sid)) == 1

    def synchronize(self, sid):
        _check(self._L.pda_synchronize(self._h, sid))

    # -- results ---------------------------------------------------------------
    def get_hist_chi2(self, sid, nbins=None):
        nb = nbins or self._hist_bins
        h = np.zeros(nb, np.int32)
        c64 = ctypes.c_double()
        c32 = ctypes.c_float()
        eff = ctypes.c_int()
        _check(self._L.pda_get_result(self._h, sid, _p(h, _i32p), nb, ctypes.byref(c64),
                                      ctypes.byref(c32), ctypes.byref(eff)))
        return h, c64.value, c32.value, eff.value

    def run_noalloc(self, d_result, result_bytes, R=None, seed=None, set_id=0, sid=0):
        """pda_run without device allocation: the result (pda_result header,
        then int32 hist[nbins]) goes to the device buffer at address
        d_result; asynchronous on the slot's stream."""
        R = self.reSampleTimes if R is None else R
        seed = self.seed if seed is None else seed
        _check(self._L.pda_run_noalloc(self._h, sid, R, seed, set_id, d_result, result_bytes))

    def run_stats(self, sid=0):
        """(units, units replayed in the general form) of the slot's last run."""
        u, r = ctypes.c_int64(), ctypes.c_int32()
        _check(self._L.pda_run_stats(self._h, sid, ctypes.byref(u), ctypes.byref(r)))
        return u.value, r.value

    def get_samples(self, sid, n):
        E = np.zeros(n, np.float32)
        acc = np.zeros(n, np.int32)
        F = np.zeros(n, np.float32)
        _check(self._L.pda_get_samples(self._h, sid, _p(E, _f32p), _p(acc, _i32p),
                                       _p(F, _f32p), n))
        return E, acc, F

    def last_run_ms(self, sid):
        ms = ctypes.c_float()
        _check(self._L.pda_last_run_ms(self._h, sid, ctypes.byref(ms)))
        return ms.value

    def last_pass_ms(self, sid):
        a, b = ctypes.c_float(), ctypes.c_float()
        _check(self._L.pda_last_pass_ms(self._h, sid, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    # -- launch shape and autotuner ---------------------------------------------
    def set_launch_shape(self, batch_tiles=0, waves_per_cu=0, queues=0):
        _check(self._L.pda_set_launch_shape(self._h, batch_tiles, waves_per_cu, queues))

    def get_launch_shape(self):
        v = [ctypes.c_int() for _ in range(3)]
        _check(self._L.pda_get_launch_shape(self._h, *[ctypes.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def autotune(self, sid=0, population=False, R=None, reps=3):
        """Picks the fastest launch shape for the slot's current work; returns
        (batch_tiles, waves_per_cu, queues, best_ms)."""
        v = [ctypes.c_int() for _ in range(3)]
        ms = ctypes.c_float()
        R = self.reSampleTimes if R is None else R
        _check(self._L.pda_autotune(self._h, sid, int(bool(population)), R, reps,
                                    *[ctypes.byref(x) for x in v], ctypes.byref(ms)))
        return tuple(x.value for x in v) + (ms.value,)

    # -- population runs -----------------------------------------------------------
    def set_population(self, nstates, args_list, sid=0):
        """A generation of parameter sets: nstates[i] states, args_list[i] the
        mc::set_params vector [E(n), K off-diagonals n(n-1), v(n)] of set i."""
        nstates = np.ascontiguousarray(nstates, np.int32)
        packed = np.ascontiguousarray(np.concatenate([np.asarray(a, np.float32).ravel()
                                                      for a in args_list]), np.float32)
        _check(self._L.pda_set_population(self._h, sid, len(nstates), _p(nstates, _i32p),
                                          _p(packed, _f32p), packed.size))
        self._pop_sets = len(nstates)

    def run_population(self, R=None, seed=None, set_base=0, set_stride=1, sid=0, graph=True):
        """Asynchronous evaluation of the slot's population; set i uses the
        random-stream key (seed, set_base + i * set_stride)."""
        R = self.reSampleTimes if R is None else R
        seed = self.seed if seed is None else seed
        flags = 0 if graph else PDA_RUN_NO_GRAPH
        _check(self._L.pda_run_population(self._h, sid, R, seed, set_base, set_stride, flags))

    def get_population_result(self, sid=0, nbins=None):
        """(hist [nsets, nbins] int32, chi2_64 [nsets], chi2_ref [nsets], effN [nsets])."""
        B = self._hist_bins if nbins is None else nbins
        S = self._pop_sets
        hist = np.zeros((S, B), np.int32)
        c64 = np.zeros(S, np.float64)
        c32 = np.zeros(S, np.float32)
        eff = np.zeros(S, np.int32)
        _check(self._L.pda_get_population_result(self._h, sid, _p(hist, _i32p), B,
                                                 _p(c64, _f64p), _p(c32, _f32p),
                                                 _p(eff, _i32p)))
        return hist, c64, c32, eff

    def population_stats(self, sid=0):
        """(units per set, replayed units, graph captures so far, GPU ms of the last run)."""
        u = ctypes.c_int64()
        r, c = ctypes.c_int32(), ctypes.c_int32()
        ms = ctypes.c_float()
        _check(self._L.pda_population_stats(self._h, sid, ctypes.byref(u), ctypes.byref(r),
                                            ctypes.byref(c), ctypes.byref(ms)))
        return u.value, r.value, c.value, ms.value

    # -- Rao-Blackwellised form ---------------------------------------------------
    def rb_forward(self, R=None, seed=None, set_id=0, sid=0, scores=True):
        """Asynchronous RB evaluation of the slot's parameter set (v = 0, n <= 4)."""
        R = self.reSampleTimes if R is None else R
        seed = self.seed if seed is None else seed
        _check(self._L.pda_rb_forward(self._h, sid, R, seed, set_id,
                                      PDA_RB_SCORES if scores else 0))
        self._rb_runs[sid] += 1

    def rb_run_id(self, sid=0):
        """Number of RB forwards started on the slot: the adjoint and the
        Jacobian differentiate the slot's last one."""
        return self._rb_runs[sid]

    def rb_result(self, sid=0, nbins=None):
        """(expected counts M_b as float64, the same in 2^-32 units, chi2, effN)."""
        B = self._hist_bins if nbins is None else nbins
        fixed = np.zeros(B, np.uint64)
        M = np.zeros(B, np.float64)
        c = ctypes.c_double()
        e = ctypes.c_int32()
        _check(self._L.pda_rb_get_result(self._h, sid, fixed.ctypes.data_as(
            ctypes.POINTER(ctypes.c_uint64)), _p(M, _f64p), B, ctypes.byref(c), ctypes.byref(e)))
        return M, fixed, c.value, e.value

    def rb_adjoint(self, g, n, sid=0):
        """d chi2 / d args (mc::set_params layout) for the cotangent g = d chi2 / d M."""
        g = np.ascontiguousarray(g, np.float64)
        out = np.zeros(n * (n + 1), np.float64)
        _check(self._L.pda_rb_adjoint(self._h, sid, _p(g, _f64p), len(g), _p(out, _f64p),
                                      len(out)))
        return out

    def rb_jacobian(self, n, nbins, sid=0):
        """d M_b / d args of the last RB forward: array (nbins, n (n + 1)),
        mc::set_params layout; g @ J equals rb_adjoint(g)."""
        J = np.zeros((nbins, n * (n + 1)), np.float64)
        _check(self._L.pda_rb_jacobian(self._h, sid, _p(J, _f64p), nbins, J.shape[1]))
        return J

    def rb_forward_population(self, R=None, seed=None, set_base=0, set_stride=1, sid=0,
                              scores=True):
        """Asynchronous RB forward of every set of the slot's population;
        set k uses the key (seed, set_base + k * set_stride)."""
        R = self.reSampleTimes if R is None else R
        seed = self.seed if seed is None else seed
        _check(self._L.pda_rb_forward_population(self._h, sid, R, seed, set_base, set_stride,
                                                 PDA_RB_SCORES if scores else 0))
        self._rb_runs[sid] += 1

    def rb_population_result(self, sid=0, nbins=None):
        """(M [nsets, B] float64, the same in 2^-32 units, chi2 [nsets], effN [nsets])."""
        B = self._hist_bins if nbins is None else nbins
        S = self._pop_sets
        fixed = np.zeros((S, B), np.uint64)
        M = np.zeros((S, B), np.float64)
        c2 = np.zeros(S, np.float64)
        eff = np.zeros(S, np.int32)
        _check(self._L.pda_rb_get_population_result(
            self._h, sid, fixed.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), _p(M, _f64p), B,
            _p(c2, _f64p), _p(eff, _i32p)))
        return M, fixed, c2, eff

    def rb_adjoint_population(self, g, n, sid=0):
        """d chi2 / d args of every set for cotangents g [nsets, B]: [nsets, n (n + 1)]."""
        g = np.ascontiguousarray(g, np.float64)
        out = np.zeros((g.shape[0], n * (n + 1)), np.float64)
        _check(self._L.pda_rb_adjoint_population(self._h, sid, _p(g, _f64p), g.shape[1],
                                                 _p(out, _f64p), n * (n + 1)))
        return out

    def rb_jacobian_population(self, n, nbins, which=None, sid=0):
        """d M_b / d args of the sets of the last RB population forward:
        [nsets, nbins, n (n + 1)] (rows of sets with which[k] == 0 are zero)."""
        S = self._pop_sets
        J = np.zeros((S, nbins, n * (n + 1)), np.float64)
        w = None if which is None else np.ascontiguousarray(which, np.int32)
        _check(self._L.pda_rb_jacobian_population(self._h, sid, None if w is None else _p(w, _i32p),
                                                  _p(J, _f64p), nbins, n * (n + 1)))
        return J

    def rb_stats(self, sid=0):
        """GPU ms of the last RB walk, whole forward, last adjoint or Jacobian."""
        v = [ctypes.c_float() for _ in range(3)]
        _check(self._L.pda_rb_stats(self._h, sid, *[ctypes.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def debug_primitive(self, op, inp, out_dtype, n):
        inp = np.ascontiguousarray(inp)
        out = np.zeros(n * {0: 4}.get(op, 1), out_dtype)
        _check(self._L.pda_debug_primitive(self._h, op, inp.ctypes.data, out.ctypes.data, n))
        return out

    def debug_guide(self, sid=0):
        """The slot's binomial threshold rows as the GPU built them: (thr int64
        [8, 8384, 2] of (up, lo), row m of a state at m (m + 3) / 2, guide uint8
        [8, 128, 64])."""
        thr = np.zeros((8, 8384, 2), np.int64)
        guide = np.zeros((8, 128, 64), np.uint8)
        _check(self._L.pda_debug_guide(self._h, sid, thr.ctypes.data, thr.nbytes,
                                       guide.ctypes.data, guide.nbytes))
        return thr, guide

    def debug_microbench(self, op, iters, param=0.0):
        """(GPU ms, events, inner iterations) of one event microbenchmark launch."""
        ms, ev, inner = ctypes.c_float(), ctypes.c_int64(), ctypes.c_int64()
        _check(self._L.pda_debug_microbench(self._h, op, iters, param, ctypes.byref(ms),
                                            ctypes.byref(ev), ctypes.byref(inner)))
        return ms.value, ev.value, inner.value


def evaluate(m, n, args, R, seed, set_id=0, sid=0, begin=0, end=-1, samples=False,
             photophysics=None):
    """One parameter-set evaluation, blocking: returns (hist, chi2_64, chi2_ref, effN)."""
    if photophysics is not None:
        m.set_photophysics(sid, *photophysics)
    if not m.set_params(n, sid, args):
        raise PdaError("bad parameter set")
    m.set_burst_range(begin, end, sid)
    m.reSampleTimes = R
    m.run_kernel(sid=sid, seed=seed, set_id=set_id, samples=samples)
    m.synchronize(sid)
    return m.get_hist_chi2(sid)
