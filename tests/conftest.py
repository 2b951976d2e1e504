import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HOSTWALK = os.path.join(ROOT, "tests", "_build", "libhostwalk.so")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C ABI)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    if not os.path.exists(O.LIB_PATH):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True,
                       capture_output=True)
    O.lib()
    return O


def build_hostwalk():
    """Host build of the product walk (tests/host/host_walk.cpp)."""
    csrc = os.path.join(ROOT, "gsmfretda_amd", "csrc")
    srcs = [os.path.join(ROOT, "tests", "host", "host_walk.cpp"), os.path.join(csrc, "pda_layout.cpp")]
    newest = max(os.path.getmtime(p) for p in srcs + [
        os.path.join(csrc, f) for f in ("pda_walk.h", "pda_math.h", "pda_types.h", "pda_layout.h",
                                     "pda_rb.h")])
    if os.path.exists(HOSTWALK) and os.path.getmtime(HOSTWALK) >= newest:
        return HOSTWALK
    os.makedirs(os.path.dirname(HOSTWALK), exist_ok=True)
    cmd = ["hipcc", "-x", "hip", "--offload-host-only", "-O2", "-fPIC", "-std=c++17",
           "-ffp-contract=off", "-fno-fast-math", "-shared", "-I", csrc, "-I",
           os.path.join(ROOT, "include"), *srcs, "-o", HOSTWALK + ".tmp"]
    subprocess.run(cmd, check=True, capture_output=True)
    os.replace(HOSTWALK + ".tmp", HOSTWALK)
    return HOSTWALK


HOST_PROGRAMS = ("mc_dropin", "ps_loop")


def host_program(name):
    """tests/host/<name>.cpp built against the HIP library into tests/_build/
    (built by __graft_entry__.build(); rebuilt here only when stale)."""
    src = os.path.join(ROOT, "tests", "host", name + ".cpp")
    lib = os.path.join(ROOT, "gsmfretda_amd", "libgsmfretda_pda.so")
    exe = os.path.join(ROOT, "tests", "_build", name)
    deps = [src, os.path.join(ROOT, "include", "gsmfretda_mc.hpp")]
    if os.path.exists(exe) and os.path.getmtime(exe) >= max(os.path.getmtime(p) for p in deps):
        return exe
    os.makedirs(os.path.dirname(exe), exist_ok=True)
    cmd = ["hipcc", "-O2", "-std=c++17", "-pthread", "-I", os.path.join(ROOT, "include"), src,
           "-L", os.path.dirname(lib), "-lgsmfretda_pda",
           "-Wl,-rpath,$ORIGIN/../../gsmfretda_amd", "-o", exe + ".tmp"]
    subprocess.run(cmd, check=True, capture_output=True, timeout=600)
    os.replace(exe + ".tmp", exe)
    return exe


@pytest.fixture(scope="session")
def gpu_mc():
    """One MC context on cuda:0 for the whole GPU session."""
    from gsmfretda_amd import pda
    m = pda.MC(0, 2)
    yield m
    m.close()


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint32 if a.dtype.itemsize == 4 else np.uint64)
