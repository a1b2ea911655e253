"""Real-molecule inputs (apo_amd/molecules.py).

* MMTF: the reference's sample proteins load with exactly the atom counts
  its benchmark and precision apps list for them
  (app/benchmarks/benchmark_stages.cu:38-42), with Bondi radii for C, N, O
  and S.  The five smallest samples are kept as fixtures under
  tests/golden/mmtf/.
* The MMTF codecs decode hand-made payloads to the values the MMTF 1.0
  specification defines (run-length, delta, recursive indexing, integer
  division), computed here by hand.
* The ball-dataset parser and apo::joggle (pcg noise, checked against the
  oracle's own pcg).
"""
import os
import struct

import numpy as np
import pytest

import oracle
from apo_amd import molecules

HERE = os.path.dirname(os.path.abspath(__file__))
MMTF = os.path.join(HERE, "golden", "mmtf")
# atom counts listed beside these files in the reference's apps
COUNTS = {"5ZCK": 31, "1AGA": 126, "3DIK": 219, "101M": 1413, "1A3F": 2784}


@pytest.mark.parametrize("name", sorted(COUNTS))
def test_mmtf_atom_counts(name):
    s = molecules.load_mmtf(os.path.join(MMTF, f"{name}.mmtf"))
    assert s.shape == (COUNTS[name], 4) and s.dtype == np.float32
    assert np.all(np.isfinite(s))
    assert set(np.unique(s[:, 3]).tolist()) <= {np.float32(v) for v in (1.0, 1.47, 1.52, 1.55, 1.7, 1.73, 1.8, 2.27)}


def test_bondi_radii_of_protein_elements():
    assert [molecules.vdw_radius(e) for e in ("C", "N", "O", "S", "FE", "Zn", "X")] == \
        [1.70, 1.55, 1.52, 1.80, 1.0, 1.39, 1.0]


def _field(codec, length, param, payload):
    return struct.pack(">iii", codec, length, param) + payload


def test_codecs_follow_the_specification():
    # 4: big-endian int32
    assert molecules.decode(_field(4, 3, 0, struct.pack(">iii", 7, -2, 40000))).tolist() == [7, -2, 40000]
    # 5: fixed-length strings
    assert molecules.decode(_field(5, 2, 4, b"A\x00\x00\x00BC\x00\x00")) == ["A", "BC"]
    # 6: run-length encoded characters (value, count)
    assert molecules.decode(_field(6, 3, 0, struct.pack(">iiii", 65, 2, 0, 1))) == ["A", "A", ""]
    # 7: run-length: (1, 4) -> 1 1 1 1
    assert molecules.decode(_field(7, 4, 0, struct.pack(">ii", 1, 4))).tolist() == [1, 1, 1, 1]
    # 8: run-length then delta: (1, 4) -> 1 2 3 4
    assert molecules.decode(_field(8, 4, 0, struct.pack(">ii", 1, 4))).tolist() == [1, 2, 3, 4]
    # 9: run-length then / param: (250, 2) / 100 -> 2.5 2.5
    assert molecules.decode(_field(9, 2, 100, struct.pack(">ii", 250, 2))).tolist() == [2.5, 2.5]
    # 10: int16 recursive index, delta, / 1000: 1000 | 32767 1000 -> 1000, 33767 -> 1.0, 34.767
    got = molecules.decode(_field(10, 2, 1000, struct.pack(">hhh", 1000, 32767, 1000)))
    assert np.allclose(got, [1.0, 34.767], atol=1e-4)
    # 12: int16 recursive index, / param: -32768 -10 -> -32778 / 10
    assert np.allclose(molecules.decode(_field(12, 1, 10, struct.pack(">hh", -32768, -10))), [-3277.8])
    # 14 / 15: recursive index without division
    assert molecules.decode(_field(14, 2, 0, struct.pack(">hhh", 32767, 1, 5))).tolist() == [32768, 5]
    assert molecules.decode(_field(15, 2, 0, struct.pack(">bbb", 127, 127, 3))).tolist() == [257]
    with pytest.raises(ValueError):
        molecules.decode(_field(99, 0, 0, b""))


def test_ball_dataset_parser(tmp_path):
    p = tmp_path / "balls.txt"
    p.write_text("3\n1 0.1 0.2 0.3 0.05\n2\t1.5  2.5 3.5\t0.25\n3 -1e-3 4 5 1.0000001\n")
    s = molecules.parse_ball_dataset(str(p))
    want = np.array([[0.1, 0.2, 0.3, 0.05], [1.5, 2.5, 3.5, 0.25], [-1e-3, 4, 5, 1.0000001]], np.float32)
    assert np.array_equal(s, want)
    p.write_text("4\n1 0 0 0 1\n")
    with pytest.raises(ValueError):
        molecules.parse_ball_dataset(str(p))


def test_joggle_is_the_reference_pcg_noise():
    s = np.zeros((50, 4), np.float32)
    j = molecules.joggle(s, 1e-3)
    want = np.array([oracle.lib().orc_pcg_ldexp(v) * 2e-3 - 1e-3 for v in range(200)], np.float64)
    assert np.array_equal(j.ravel(), want.astype(np.float32))
    assert np.abs(j).max() <= 1e-3
