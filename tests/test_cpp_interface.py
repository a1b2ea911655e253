"""The C++17 host interface (include/apo/gpu/algorithm.hpp) over the C ABI.

CPU: the header compiles on its own, a caller written as the reference's
apps are (Real sites, AlgorithmGPU<K>{sites, knnValidationStep}, the
Algorithm scene members, apo::joggle on a Span) compiles against it, and so
does the reference-side shim (apo_amd/examples/apo_gpu_shim.cpp) against its
declarations-only form.  GPU: the C++ caller built against it
(apo_amd/examples/diagram.cpp) reproduces the oracle's vertices and closed
edges (computeFullApolloniusDiagram) and writes the topology document of the
reference's topology app with the oracle's sets; the reference-style caller
reproduces the oracle through the shim and with K = 1 over 4 kNN passes.
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


def _compile_only(srcs, *defines):
    res = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", *defines,
                          "-I", os.path.join(ROOT, "include"), *srcs], capture_output=True, text=True)
    assert res.returncode == 0, res.stderr


def test_reference_style_caller_compiles(tmp_path):
    src = tmp_path / "caller.cpp"
    src.write_text(
        '#include "apo/gpu/algorithm.hpp"\n'
        "int main() {\n"
        "    std::vector<apo::Real> sites(400, 0.5);\n"
        "    apo::joggle(sites, 1e-3);\n"                                   # Span<Real> from a vector
        "    apo::gpu::AlgorithmGPU<1> algorithm{sites, 4};\n"              # benchmark_knn.cu:40
        "    apo::gpu::Algorithm& base = algorithm;\n"
        "    base.build();\n"
        "    apo::gpu::VertexDiagram v = base.toHost(true);\n"
        "    apo::gpu::FullDiagram f = apo::gpu::computeFullApolloniusDiagram(sites);\n"
        "    apo::gpu::Topology t = algorithm.toTopology();\n"
        "    double box = base.m_maxX - base.m_minX + base.m_maxRadius - base.m_minRadius;\n"
        "    return int(v.vertices.size() + f.closedEdgesMin.size() + t.bisectors.size() + box\n"
        "               + algorithm.sample.knnSearch + algorithm.knnValidationStep) & 1;\n"
        "}\n")
    _compile_only([str(src)])


def test_reference_shim_compiles():
    ex = os.path.join(ROOT, "apo_amd", "examples")
    _compile_only([os.path.join(ex, "apo_gpu_shim.cpp"), os.path.join(ex, "reference_caller.cpp")],
                  "-DAPO_GPU_DECLARATIONS_ONLY")


def _read_vertices(raw, o):
    v = int(np.frombuffer(raw[o:o + 8], np.uint64)[0])
    o += 8
    xyzw = np.frombuffer(raw[o:o + 32 * v], np.float64).reshape(v, 4)
    o += 32 * v
    ids = np.frombuffer(raw[o:o + 16 * v], np.uint32).reshape(v, 4)
    return xyzw, ids, o + 16 * v


@pytest.mark.gpu
@pytest.mark.parametrize("k,passes", [(1, 4), (16, 2)])
def test_reference_caller_through_shim(tmp_path, k, passes):
    if not os.path.exists(hip_build.SHIM_CALLER):
        pytest.fail("apo_reference_caller not built (run __graft_entry__.build())")
    sites = datasets.uniform_sites(10_000, seed=1)
    sites.tofile(tmp_path / "sites.bin")
    print(f"[case] reference-style caller, 10000 uniform sites, K={k}, {passes} passes", flush=True)
    res = subprocess.run([hip_build.SHIM_CALLER, str(tmp_path / "sites.bin"), str(tmp_path / "out.bin"), str(k),
                          str(passes)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    raw = (tmp_path / "out.bin").read_bytes()
    xyzw, ids, o = _read_vertices(raw, 0)
    ex, eids, o = _read_vertices(raw, o)
    kx, kids, o = _read_vertices(raw, o)
    assert o == len(raw)
    d = oracle.Diagram(sites)
    ref_x, ref_ids = d.vertices()
    assert np.array_equal(ids, ref_ids)
    assert np.array_equal(xyzw.view(np.uint64), ref_x.view(np.uint64))
    ref_ex, ref_eids = d.closed_edges()
    assert np.array_equal(eids, ref_eids)
    assert np.array_equal(ex.view(np.uint64), ref_ex.view(np.uint64))
    dk = oracle.Diagram(sites, k=k, knn_passes=passes)
    rx, ri = dk.vertices()
    assert np.array_equal(kids, ri)
    assert np.array_equal(kx.view(np.uint64), rx.view(np.uint64))


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
