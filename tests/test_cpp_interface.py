"""The C++17 host interface (include/apo/gpu/algorithm.hpp) over the C ABI.

CPU: the header compiles on its own.  GPU: the C++ caller built against it
(apo_amd/examples/diagram.cpp) reproduces the oracle's vertex set.
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import oracle
from apo_amd import build as hip_build
from apo_amd import datasets

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_header_compiles_standalone():
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "use.cpp")
        with open(src, "w") as f:
            f.write('#include "apo/gpu/algorithm.hpp"\n'
                    "int main() { apo_config c; apo_default_config(&c); return c.k == 16 ? 0 : 1; }\n")
        res = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only",
                              "-I", os.path.join(ROOT, "include"), src], capture_output=True, text=True)
        assert res.returncode == 0, res.stderr


@pytest.mark.gpu
def test_cpp_caller_matches_oracle(tmp_path):
    if not os.path.exists(hip_build.EXAMPLE):
        pytest.fail("apo_diagram not built (run __graft_entry__.build())")
    sites = datasets.uniform_sites(3000, seed=4)
    sites.tofile(tmp_path / "sites.bin")
    print("[case] C++ caller, 3000 uniform sites", flush=True)
    res = subprocess.run([hip_build.EXAMPLE, str(tmp_path / "sites.bin"), str(tmp_path / "out.bin")],
                         capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    raw = (tmp_path / "out.bin").read_bytes()
    v = int(np.frombuffer(raw[:8], np.uint64)[0])
    xyzw = np.frombuffer(raw[8:8 + 32 * v], np.float64).reshape(v, 4)
    ids = np.frombuffer(raw[8 + 32 * v:], np.uint32).reshape(v, 4)
    ref_x, ref_ids = oracle.Diagram(sites).vertices()
    assert np.array_equal(ids, ref_ids)
    assert np.array_equal(xyzw.view(np.uint64), ref_x.view(np.uint64))
