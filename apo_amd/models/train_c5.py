"""Writes proposal_c5.f32 and proposal_c5.json: the learned proposal
trained on 2M labelled sites (two 1M log-uniform sets, seeds 501 and 502;
class labels from the library's label pass, radius targets from a completed
plain build of each set), 8 epochs of the training kernels (batch 4096,
Adam lr 2e-3, seed 7).  Needs the GPU.

    python -m apo_amd.models.train_c5
"""
from __future__ import annotations

import hashlib
import json
import os
import time

import numpy as np

from apo_amd import datasets, proposal

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "proposal_c5.f32")
HEADER = os.path.join(HERE, "proposal_c5.json")
RECIPE = {"sets": [{"sites": 1_000_000, "generator": "loguniform_sites", "seed": 501},
                   {"sites": 1_000_000, "generator": "loguniform_sites", "seed": 502}],
          "epochs": 8, "batch": 4096, "lr": 2e-3, "seed": 7}


def train():
    t0 = time.perf_counter()
    feats, labs, rads = [], [], []
    for s in RECIPE["sets"]:
        sites = datasets.loguniform_sites(s["sites"], seed=s["seed"])
        f, lab = proposal.labels(sites)
        feats.append(f)
        labs.append(lab)
        rads.append(proposal.radius_targets(sites))
    t_labels = time.perf_counter() - t0
    t0 = time.perf_counter()
    params, loss = proposal.train(np.concatenate(feats), np.concatenate(labs), np.concatenate(rads),
                                  epochs=RECIPE["epochs"], batch=RECIPE["batch"], lr=RECIPE["lr"],
                                  seed=RECIPE["seed"])
    t_train = time.perf_counter() - t0
    return params, loss, t_labels, t_train


def main():
    params, loss, t_labels, t_train = train()
    proposal.save(params, WEIGHTS)
    header = {"layout": "float32 little-endian, W1 [52][64], b1 [64], W2 [64][64], b2 [64], W3 [64][9], b3 [9] "
                        "(weights [in][out]; output 8 is the radius head)",
              "params": int(params.size), "class_k": list(proposal.CLASS_K), "recipe": RECIPE,
              "epoch_loss": [float(x) for x in loss], "label_seconds": t_labels, "train_seconds": t_train,
              "sha256": hashlib.sha256(np.ascontiguousarray(params, "<f4").tobytes()).hexdigest()}
    with open(HEADER, "w") as f:
        json.dump(header, f, indent=1)
    print(json.dumps(header))


if __name__ == "__main__":
    main()
