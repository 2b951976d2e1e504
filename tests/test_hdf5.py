"""HDF5 ingest (gsmfretda_amd.hdf5) without libhdf5.

The reader is checked against files written by the HDF5 library itself (the
PyTables test files shipped in this image, whose contents PyTables' own tests
state): 6x5 arrays i + j in int32/int64/float64 of both byte orders,
float16/32/64 arrays, a chunked extendible big-endian array with unwritten
chunks, and shuffle + deflate chunked index arrays whose int32 and float64
copies must agree; unsupported classes and filters must raise H5Error.  The
burst schema of src/loadHdf5.cpp:37-74 round-trips through the writer, and
bursts read back evaluate bit for bit as the originals (oracle on the CPU,
the HIP path on the GPU).
"""
import glob
import os

import numpy as np
import pytest

from gsmfretda_amd import hdf5, synth

PT = "/opt/conda/lib/python3.9/site-packages/tables/tests/"
needs_pt = pytest.mark.skipif(not os.path.isdir(PT), reason="HDF5 library test files absent")


@needs_pt
@pytest.mark.parametrize("fn,dt", [("smpl_f64be.h5", "float64"), ("smpl_f64le.h5", "float64"),
                                   ("smpl_i64be.h5", "int64"), ("smpl_i64le.h5", "int64"),
                                   ("smpl_i32be.h5", "int32"), ("smpl_i32le.h5", "int32")])
def test_known_numeric_arrays(fn, dt):
    a = hdf5.H5File(PT + fn).read("/TestArray")
    want = np.add.outer(np.arange(6), np.arange(5)).astype(dt)
    assert a.dtype == np.dtype(dt) and np.array_equal(a, want)


@needs_pt
def test_float_widths():
    f = hdf5.H5File(PT + "float.h5")
    want = np.arange(6) + np.arange(5).reshape(5, 1)
    for k in ("float16", "float32", "float64"):
        a = f.read("/" + k)
        assert a.dtype == np.dtype(k) and np.array_equal(a, want.astype(k))
    with pytest.raises(hdf5.H5Error, match="float size 16"):
        f.read("/longdouble")


@needs_pt
def test_chunked_extendible_array():
    a = hdf5.H5File(PT + "smpl_SDSextendible.h5").read("/ExtendibleArray")
    want = np.zeros((10, 5), np.int32)
    want[:2, :3] = 1
    want[2, :3] = 1
    want[:2, 3:] = 3
    want[3:, 0] = 2
    assert a.dtype == np.int32 and np.array_equal(a, want)


@needs_pt
def test_shuffle_deflate_chunks():
    f = hdf5.H5File(PT + "indexes_2_1.h5")
    g = "/_i_table1/"
    s3, s4 = f.read(g + "var3/sorted"), f.read(g + "var4/sorted")
    assert s3.dtype == np.int32 and s4.dtype == np.float64
    assert np.array_equal(s3, s4) and np.array_equal(s3.ravel(), np.arange(16))
    for k in ("ranges", "bounds", "abounds", "zbounds", "mbounds", "sortedLR"):
        assert np.array_equal(f.read(g + "var3/" + k), f.read(g + "var4/" + k)), k
    assert np.array_equal(f.read(g + "var4/indices").ravel(), np.arange(16))


@needs_pt
def test_every_library_file_reads_or_refuses_cleanly():
    ok = 0
    for fn in sorted(glob.glob(PT + "*.h5")):
        f = hdf5.H5File(fn)

        def walk(g, depth=0):
            nonlocal ok
            for k in f.keys(g):
                path = g.rstrip("/") + "/" + k
                try:
                    f.read(path)
                    ok += 1
                except hdf5.H5Error as e:
                    if "not a dataset" in str(e) and depth < 3:
                        try:
                            walk(path, depth + 1)
                        except hdf5.H5Error:
                            pass
        walk("/")
    assert ok >= 80


def _bursts(n=300, seed=3):
    args = synth.make_args([0.3, 0.7], [1e3, 1e3])
    d = synth.generate(n, args, 2, seed=seed, p_neither=0.05)
    sg = np.random.default_rng(seed).uniform(0, 1, n).astype(np.float32)
    return d, sg


def test_schema_round_trip(tmp_path):
    d, sg = _bursts()
    p = str(tmp_path / "bursts.hdf5")
    hdf5.write_bursts(p, d, sg)
    f = hdf5.H5File(p)
    assert f.keys("/") == sorted(["sub_bursts_l", "times", "mask_ad", "mask_dd", "T_burst_duration",
                                  "SgDivSr", "clk_p", "bg_ad_rate", "bg_dd_rate"])
    assert f.keys("/sub_bursts_l") == ["istart", "istop", "start", "stop"]
    assert f.read("clk_p").shape == ()
    e, sg2 = hdf5.read_bursts(p)
    for k in ("start", "stop", "istart", "phcount", "times", "mask_ad", "mask_dd", "T"):
        assert np.array_equal(getattr(e, k), getattr(d, k)), k
    assert np.array_equal(sg2, sg)
    assert (e.clk_p, e.bg_ad, e.bg_dd) == (np.float32(d.clk_p), np.float32(d.bg_ad),
                                           np.float32(d.bg_dd))


def test_schema_errors(tmp_path):
    d, sg = _bursts(50)
    p = str(tmp_path / "bad.hdf5")
    hdf5.write(p, {"/sub_bursts_l/start": d.start, "/sub_bursts_l/stop": d.stop[:-1],
                   "/sub_bursts_l/istart": d.istart, "/sub_bursts_l/istop": d.istop,
                   "times": d.times, "mask_ad": d.mask_ad, "mask_dd": d.mask_dd,
                   "T_burst_duration": d.T, "SgDivSr": sg, "clk_p": np.float32(1.25e-8),
                   "bg_ad_rate": np.float32(0), "bg_dd_rate": np.float32(0)})
    with pytest.raises(hdf5.H5Error, match="differ in length"):
        hdf5.read_bursts(p)
    with pytest.raises(KeyError):
        hdf5.H5File(p).read("/nope")
    q = tmp_path / "not.h5"
    q.write_bytes(b"plain text" * 100)
    with pytest.raises(hdf5.H5Error, match="not an HDF5 file"):
        hdf5.H5File(str(q))


def test_read_bursts_evaluate_like_the_originals(oracle, tmp_path):
    d, sg = _bursts(200, seed=9)
    p = str(tmp_path / "b.hdf5")
    hdf5.write_bursts(p, d, sg)
    e, _ = hdf5.read_bursts(p)
    prm = oracle.make_params(2, synth.make_args([0.3, 0.7], [1e3, 1e3]), *synth.PHOTOPHYSICS)
    a = oracle.run(oracle.Data.from_bursts(d), prm, R=8, seed=11)
    b = oracle.run(oracle.Data.from_bursts(e), prm, R=8, seed=11)
    assert np.array_equal(a["hist"], b["hist"]) and np.array_equal(a["acc"], b["acc"])


@pytest.mark.gpu
def test_hdf5_bursts_on_the_gpu_equal_the_oracle_on_the_originals(gpu_mc, oracle, tmp_path):
    """Bursts written in the reference's schema, read back and loaded into
    the GPU: every unit's acceptor count, corrected count and E, the
    histogram and chi2 equal the oracle's run on the original in-memory
    arrays (src/loadHdf5.cpp:37-74, 101-128)."""
    from tests.conftest import bits
    d, sg = _bursts(500, seed=21)
    p = str(tmp_path / "f2.hdf5")
    hdf5.write_bursts(p, d, sg)
    n, args, ph, R, seed = 3, synth.make_args([0.2, 0.5, 0.8], [700, 2500, 900, 3000, 1200, 4000]), \
        synth.PHOTOPHYSICS, 24, 606
    gpu_mc.load_hdf5(p, nbins=50, left_pad=2)
    gpu_mc.set_photophysics(0, *ph)
    assert gpu_mc.set_params(n, 0, args)
    gpu_mc.set_burst_range(0, -1, 0)
    gpu_mc.reSampleTimes = R
    gpu_mc.run_kernel(sid=0, seed=seed, set_id=2, samples=True)
    gpu_mc.synchronize(0)
    hist, c64, c32, _ = gpu_mc.get_hist_chi2(0)
    E, acc, F = gpu_mc.get_samples(0, d.nbursts * R)
    o = oracle.run(oracle.Data.from_bursts(d), oracle.make_params(n, args, *ph), R=R, seed=seed,
                   set_id=2)
    assert np.array_equal(acc, o["acc"])
    assert np.array_equal(bits(F), bits(o["F"])) and np.array_equal(bits(E), bits(o["E"]))
    assert np.array_equal(hist.astype(np.int64), o["hist"])
    w32, w64, _ = oracle.chi2(oracle.hist(sg).astype(np.float32), o["hist"], R, n)
    assert np.float32(c32).view(np.uint32) == np.float32(w32).view(np.uint32)
    assert abs(c64 - w64) <= 1e-12 * abs(w64)


@pytest.mark.gpu
def test_hdf5_into_the_gpu(gpu_mc, tmp_path):
    from gsmfretda_amd import pda
    d, sg = _bursts(400, seed=12)
    p = str(tmp_path / "g.hdf5")
    hdf5.write_bursts(p, d, sg)
    args = synth.make_args([0.3, 0.7], [1e3, 1e3])
    gpu_mc.load(d, SgDivSr=sg, nbins=50, left_pad=2)
    want = pda.evaluate(gpu_mc, 2, args, 16, 5, photophysics=synth.PHOTOPHYSICS)
    gpu_mc.load_hdf5(p, nbins=50, left_pad=2)
    got = pda.evaluate(gpu_mc, 2, args, 16, 5, photophysics=synth.PHOTOPHYSICS)
    assert np.array_equal(got[0], want[0]) and got[1] == want[1]
