"""Persisted proposal weights (apo_amd/models/): the committed parameters
load (raw float32, checksum in the JSON header checked), and save/load
round-trips bit for bit.  CPU only."""
import json
import os

import numpy as np
import pytest

from apo_amd import proposal


def test_committed_weights_load():
    p = proposal.load()
    assert p.dtype == np.float32 and p.size == proposal.PARAMS
    assert np.all(np.isfinite(p))
    with open(os.path.splitext(proposal.DEFAULT_WEIGHTS)[0] + ".json") as f:
        header = json.load(f)
    assert header["params"] == proposal.PARAMS
    assert list(header["class_k"]) == list(proposal.CLASS_K)


def test_save_load_round_trip(tmp_path):
    p = np.random.default_rng(1).normal(size=proposal.PARAMS).astype(np.float32)
    path = str(tmp_path / "w.f32")
    proposal.save(p, path)
    assert np.array_equal(proposal.load(path).view(np.uint32), p.view(np.uint32))


def test_checksum_mismatch_is_refused(tmp_path):
    p = proposal.load()
    path = str(tmp_path / "proposal_c5.f32")
    q = p.copy()
    q[0] += 1.0
    proposal.save(q, path)
    with open(os.path.splitext(proposal.DEFAULT_WEIGHTS)[0] + ".json") as f:
        header = f.read()
    (tmp_path / "proposal_c5.json").write_text(header)
    with pytest.raises(ValueError):
        proposal.load(path)
