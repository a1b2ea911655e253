"""The C++17 host interface (include/apo/gpu/algorithm.hpp) over the C ABI.

CPU: the header compiles on its own.  GPU: the C++ caller built against it
(apo_amd/examples/diagram.cpp) reproduces the oracle's vertices and closed
edges (computeFullApolloniusDiagram) and writes the topology document of the
reference's topology app with the oracle's sets.
"""
import json
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
    # log-uniform radii: this set has 22 closed edges
    sites = datasets.loguniform_sites(20_000, seed=3)
    sites.tofile(tmp_path / "sites.bin")
    print("[case] C++ caller, 20000 log-uniform sites", flush=True)
    res = subprocess.run([hip_build.EXAMPLE, str(tmp_path / "sites.bin"), str(tmp_path / "out.bin"),
                          str(tmp_path / "topology.json")], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    raw = (tmp_path / "out.bin").read_bytes()
    v = int(np.frombuffer(raw[:8], np.uint64)[0])
    o = 8
    xyzw = np.frombuffer(raw[o:o + 32 * v], np.float64).reshape(v, 4)
    o += 32 * v
    ids = np.frombuffer(raw[o:o + 16 * v], np.uint32).reshape(v, 4)
    o += 16 * v
    e = int(np.frombuffer(raw[o:o + 8], np.uint64)[0])
    o += 8
    exyzw = np.frombuffer(raw[o:o + 32 * e], np.float64).reshape(e, 4)
    o += 32 * e
    eids = np.frombuffer(raw[o:o + 16 * e], np.uint32).reshape(e, 4)
    d = oracle.Diagram(sites)
    ref_x, ref_ids = d.vertices()
    assert np.array_equal(ids, ref_ids)
    assert np.array_equal(xyzw.view(np.uint64), ref_x.view(np.uint64))
    ref_ex, ref_eids = d.closed_edges()
    assert e > 0
    assert np.array_equal(eids, ref_eids)
    assert np.array_equal(exyzw.view(np.uint64), ref_ex.view(np.uint64))
    doc = json.loads((tmp_path / "topology.json").read_text())
    topo = d.topology()
    for key in ("bisectors", "trisectors", "quadrisectors"):
        assert np.array_equal(np.asarray(doc[key], np.uint32).reshape(topo[key].shape), topo[key])
