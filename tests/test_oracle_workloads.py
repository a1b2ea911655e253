"""The parity workloads reach the regimes the kernels treat separately.

The cell kernels take a cell's edges 64 per step (one lane each) and cap
every cell at fixed record capacities; these checks, on the oracle alone,
keep the GPU parity cases covering cells of one, two and four 64-edge
batches and a cell past the edge capacity.
"""
import numpy as np

import oracle
from apo_amd import datasets

K_EDGE_CAP = 264          # per-cell edge capacity of the library
FLAG_EDGE_OVERFLOW = 0x2  # cell flag: edges truncated at the capacity


def _edges(sites):
    c, _, _ = oracle.Diagram(sites).all_cells()
    return c["edge_nb"], c["flags"]


def test_loguniform_parity_case_has_two_batch_cells():
    en, _ = _edges(datasets.loguniform_sites(20_000, seed=3))
    assert np.median(en) < 64
    assert (en > 64).sum() > 100


def test_crowded_shell_has_a_four_batch_cell_within_capacity():
    en, fl = _edges(datasets.crowded_shell_sites(60))
    assert 192 < en.max() <= K_EDGE_CAP
    assert not fl.any()


def test_crowded_shell_overflows_the_edge_capacity():
    en, fl = _edges(datasets.crowded_shell_sites(150))
    assert en.max() == K_EDGE_CAP
    big = int(np.argmax(en))
    assert fl[big] & FLAG_EDGE_OVERFLOW
