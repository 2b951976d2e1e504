"""The C-ABI library loads, exports every entry point include/*.h declares,
keeps the mc shim's method set, and fails loudly (no CPU fallback)."""
import os
import re
import subprocess

import pytest

from gsmfretda_amd import pda
from tests.conftest import ROOT

HEADER = os.path.join(ROOT, "include", "gsmfretda_pda.h")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(pda_\w+)\s*\(", text, re.M)))


def dynsyms(flags):
    out = subprocess.run(["nm", "-D", "--defined-only", *flags, pda.LIB_PATH], check=True,
                         capture_output=True, text=True).stdout
    return out


def test_library_loads():
    L = pda.load_library()
    assert L.pda_version().decode().startswith("gsmfretda_amd")


def test_exports_every_declared_symbol():
    names = declared()
    assert len(names) >= 20
    syms = {line.split()[-1] for line in dynsyms([]).splitlines() if " T " in line}
    missing = [n for n in names if n not in syms]
    assert not missing, missing
    assert sorted(pda.EXPORTS) == names


def test_mc_shim_methods_exported():
    """The reference's class mc methods (src/mc.hpp:79-179) are all present."""
    text = dynsyms(["-C"])
    for m in ("mc::mc(int, int, unsigned char, unsigned long, bool)", "mc::~mc()",
              "mc::getStreams(int*)", "mc::streamQuery(int)", "mc::set_nstates(int, int)",
              "mc::set_gpuid()", "mc::setBurstBd(int, int, int)", "mc::set_reSampleTimes(int)",
              "mc::free_data_gpu()", "mc::init_randstate(int, int)", "mc::run_kernel(int, int)",
              "mc::get_res(int, int)", "mc::set_params_buff(int, int, int)",
              "mc::get_max_conversion_capacity(int)", "mc::init_data_gpu(",
              "mc::set_params(int, int, std::vector<float, std::allocator<float> >&)"):
        assert m in text, m


def test_no_gpu_fails_loudly():
    """Without a GPU the context cannot be created and the error is reported."""
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    with pytest.raises(pda.PdaError):
        pda.MC(0, 1)
    assert pda.load_library().pda_last_error().decode() != ""


def test_missing_library_raises(tmp_path):
    saved = pda._lib
    pda._lib = None
    try:
        with pytest.raises(pda.PdaError, match="missing"):
            pda.load_library(str(tmp_path / "nope.so"))
    finally:
        pda._lib = saved


def test_bench_refuses_more_than_one_gpu():
    """bench.py runs on one GPU: --gpus above 1 exits with a message before
    touching a device."""
    r = subprocess.run(["python", os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
