"""Writes the golden fixtures of tests/golden/ from the oracle (the CPU
restatement of the reference's method; the reference itself ships no
expected outputs).  Each fixture holds the FP32 input sites and the
oracle's results: LBVH arrays, 16-NN lists, emitted vertices (FP64 centre +
radius, original ids) and the toTopology sets; the batch fixture holds
molecule offsets and the per-molecule vertices with batch ids.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402
from apo_amd import datasets  # noqa: E402


def single(name, sites):
    b = oracle.lbvh(sites)
    d = oracle.Diagram(sites)
    x, i = d.vertices()
    t = d.topology()
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), sites=sites, codes=b["codes"], indices=b["indices"],
                        parent=b["parent"], left=b["left"], right=b["right"], aabb=b["aabb"], sizes=b["sizes"],
                        max_radius=b["max_radius"], knn=d.knns(), xyzw=x, ids=i, bisectors=t["bisectors"],
                        trisectors=t["trisectors"], quadrisectors=t["quadrisectors"])
    print(f"{name}: {len(sites)} sites, {len(x)} vertices")


def batch(name, count, lo, hi, seed):
    sites, offsets = datasets.batched_molecules(count=count, seed=seed, lo=lo, hi=hi)
    xs, ids = [], []
    for m in range(count):
        a, b = int(offsets[m]), int(offsets[m + 1])
        x, i = oracle.Diagram(sites[a:b]).vertices()
        xs.append(x)
        ids.append(i + np.uint32(a))
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), sites=sites, offsets=offsets, xyzw=np.concatenate(xs),
                        ids=np.concatenate(ids))
    print(f"{name}: {count} molecules, {len(sites)} sites, {sum(len(x) for x in xs)} vertices")


if __name__ == "__main__":
    single("uniform_400", datasets.uniform_sites(400, seed=61, r_min=0.01, r_max=0.04))
    single("loguniform_500", datasets.loguniform_sites(500, seed=62))
    single("protein_400", datasets.protein_like_sites(400, seed=63))
    batch("batch_6", 6, 40, 400, 64)
