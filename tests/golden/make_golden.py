"""Writes the golden fixtures under tests/golden/ from the reference itself.

The reference (dnbaker/turbotops v0.1.0) is compiled in place from its own
sources by oracle/Makefile into oracle/_ref/ (scalar kernels, FP contraction
off); this script calls that build through oracle.Reference and stores
inputs and outputs only. Run it where the reference sources exist:

    make -C oracle && python tests/golden/make_golden.py

Inputs are numpy default_rng(seed) draws in float64 cast to float32.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import Reference  # noqa: E402

METRICS = ["l1", "l2", "linf", "lp:1.5", "lp:3", "lp:4", "lp:7.5", "lp:50"]


def clouds():
    rng = np.random.default_rng(20261203)
    out = {
        "gauss37x3": rng.normal(size=(37, 3)),
        "unif50x16": rng.uniform(size=(50, 16)),
        "gauss29x5": rng.normal(size=(29, 5)),
        "grid40x3": rng.integers(0, 4, size=(40, 3)).astype(np.float64),   # exact ties
        "gauss70x17": rng.normal(size=(70, 17)),
    }
    return {k: v.astype(np.float32) for k, v in out.items()}


def main():
    ref = Reference(simd=False)
    rng = np.random.default_rng(7)
    # per-pair metric values, every metric and dims 1..20
    a = rng.normal(scale=3.0, size=(400, 20)).astype(np.float32)
    b = rng.normal(scale=3.0, size=(400, 20)).astype(np.float32)
    dims = rng.integers(1, 21, size=400).astype(np.uint32)
    pairs = {"a": a, "b": b, "dims": dims}
    for m in METRICS:
        pairs["d_" + m] = np.array([ref.metric(m, a[t, :dims[t]], b[t, :dims[t]]) for t in range(400)],
                                   np.float32)
    np.savez_compressed(os.path.join(HERE, "metric_pairs.npz"), **pairs)

    dense, graphs = {}, {}
    for name, x in clouds().items():
        dense["x_" + name] = x
        graphs["x_" + name] = x
        for m in METRICS:
            full, mat = ref.make_dense(x, m)
            dense[f"full_{name}_{m}"] = mat
            dense[f"radius_{name}_{m}"] = np.float32(ref.enclosing_radius(mat, x.shape[0], True))
            _, low = ref.make_dense(x, m, budget=x.shape[0] * (x.shape[0] - 1) * 2)
            dense[f"lower_{name}_{m}"] = low
            rp, col, w = ref.radius_graph(x, m, float(np.float32(0.6) * dense[f"radius_{name}_{m}"]))
            graphs[f"rg_rp_{name}_{m}"], graphs[f"rg_col_{name}_{m}"], graphs[f"rg_w_{name}_{m}"] = rp, col, w
            graphs[f"knn_{name}_{m}"] = ref.knn(x, m, 7)
            rp, col, w = ref.knn_graph(x, m, 7)
            graphs[f"kg_rp_{name}_{m}"], graphs[f"kg_col_{name}_{m}"], graphs[f"kg_w_{name}_{m}"] = rp, col, w
            dims_, births, deaths = ref.mst_h0(x, m)
            graphs[f"mst_{name}_{m}"] = deaths
            idx, d, ap, thr = ref.dense_edges(x, m)
            graphs[f"edges_idx_{name}_{m}"], graphs[f"edges_d_{name}_{m}"] = idx, d
            graphs[f"edges_apparent_{name}_{m}"] = ap
            graphs[f"edges_thr_{name}_{m}"] = np.float32(thr)
    np.savez_compressed(os.path.join(HERE, "dense_small.npz"), **dense)
    np.savez_compressed(os.path.join(HERE, "graphs_small.npz"), **graphs)
    print("wrote", sorted(f for f in os.listdir(HERE) if f.endswith(".npz")))


if __name__ == "__main__":
    main()
