"""The C-ABI library loads and exports every entry point include/apo_c.h declares
(CPU only: no compute call needs a GPU here)."""
import ctypes
import os
import re

import numpy as np
import pytest

from apo_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "apo_c.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(apo_[a-z_0-9]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert declared_symbols() == sorted(_lib.SIGNATURES)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.library_path()):
        pytest.skip("libapo_hip.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(_lib.library_path())
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_default_config_matches_reference_defaults():
    if not os.path.exists(_lib.library_path()):
        pytest.skip("libapo_hip.so not built")
    cfg = _lib.Config()
    _lib.lib().apo_default_config(ctypes.byref(cfg))
    # AlgorithmGPU<16>(sites, knnValidationStep = 1), 10 validation rounds, <100 restarts
    assert (cfg.k, cfg.knn_passes, cfg.max_iterations, cfg.max_restarts, cfg.stage_mask) == (16, 1, 10, 100, 7)


def test_config_struct_mirrors_header():
    text = open(os.path.join(ROOT, "include", "apo_c.h")).read()
    body = re.search(r"typedef struct apo_config \{(.*?)\} apo_config;", text, flags=re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    fields = re.findall(r"(uint32_t|float)\s+([a-z_0-9]+);", body)
    want = [(name, {"uint32_t": "c_uint", "float": "c_float"}[ty]) for ty, name in fields]
    got = [(name, t.__name__) for name, t in _lib.Config._fields_]
    assert want == got


def test_joggle_matches_the_python_restatement():
    # apo::joggle (utils.cpp:14-25) in the C ABI, FP64 and FP32 sites, against
    # apo_amd.molecules.joggle (itself pinned to the oracle's pcg)
    if not os.path.exists(_lib.library_path()):
        pytest.skip("libapo_hip.so not built")
    from apo_amd import datasets, molecules
    s = datasets.uniform_sites(500, seed=4)
    f = s.copy()
    _lib.lib().apo_joggle_f32(_lib.fptr(f), f.size, 1e-3)
    assert np.array_equal(f, molecules.joggle(s, 1e-3))
    d = s.astype(np.float64)
    _lib.lib().apo_joggle(_lib.dptr(d), d.size, 1e-3)
    v = np.arange(s.size, dtype=np.uint32)
    noise = np.ldexp(molecules.pcg_word(v).astype(np.float64), -32) * 2e-3 - 1e-3
    assert np.array_equal(d.ravel(), s.astype(np.float64).ravel() + noise)


def test_stats_struct_mirrors_header():
    text = open(os.path.join(ROOT, "include", "apo_c.h")).read()
    body = re.search(r"typedef struct apo_stats \{(.*?)\} apo_stats;", text, flags=re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    fields = re.findall(r"(double|uint64_t)\s+([a-z_0-9]+)(\[\d+\])?;", body)
    want = [(name, {"double": "c_double", "uint64_t": "c_ulong"}[ty], bool(arr)) for ty, name, arr in fields]
    got = [(name, (t._type_ if hasattr(t, "_length_") else t).__name__, hasattr(t, "_length_"))
           for name, t in _lib.Stats._fields_]
    assert [w[0] for w in want] == [g[0] for g in got]
    words = sum(int(arr[1:-1]) if arr else 1 for _, _, arr in fields)
    assert ctypes.sizeof(_lib.Stats) == 8 * words
