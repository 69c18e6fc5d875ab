#!/usr/bin/env python
"""Train the learned near-field preconditioner on the GPU and evaluate it.

    python treams-rs_amd/train_preconditioner.py [--clusters 4096] [--epochs 2] [--out PATH]

Training set: clusters generated by the method itself, N uniform in {64..1024}, lmax
in {3..6}, k0 U[0.8, 2.0], radii U[0.15, 0.35], eps U[2,6] + i U[0,0.2], fill fraction
U[0.05, 0.25], seeds 1000+; one Adam step per cluster on the dominant error modes of
A M^{-1} - I, the library's forward, backward and Adam kernels.
Evaluation (held-out seeds 9000+ of the same distribution, then the N = 1024, lmax = 6
workload, seed 0): GMRES(50) iterations to a true relative residual of 1e-8 for four
plane waves, in the headline's mixed precision (inner solves on the c64-precision
operator, FP64 refinement) and in FP64, with no preconditioner, block-Jacobi, the
near-field preconditioner without the network (alpha = 1) and with it.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--clusters", type=int, default=4096)
    ap.add_argument("--epochs", type=int, default=2)
    ap.add_argument("--probes", type=int, default=4)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--eval", type=int, default=16)
    ap.add_argument("--out", default=None)
    ap.add_argument("--report", default=None)
    ap.add_argument("--compare", default=None, help="weights to evaluate beside the new ones")
    args = ap.parse_args()

    import torch

    from __graft_entry__ import load_package

    pkg = load_package()
    import importlib

    learned = importlib.import_module("treams_rs_amd.learned")
    out_path = args.out or learned.DEFAULT_WEIGHTS
    t0 = time.perf_counter()
    mlp = learned.train(n_clusters=args.clusters, epochs=args.epochs, probes=args.probes, lr=args.lr, seed=1000,
                        log=lambda m: print(f"[train] {m}", flush=True))
    train_s = time.perf_counter() - t0
    mlp.save(out_path, clusters=args.clusters, epochs=args.epochs, lr=args.lr, train_seconds=train_s)
    print(f"[train] saved {out_path} after {train_s:.1f} s", flush=True)

    nets = {"neumann_alpha1": learned.MLP(seed=1), "learned": mlp}  # MLP(seed): zero last layer, alpha = 1
    if args.compare:
        nets["compare"] = learned.MLP.load(args.compare)

    def iterations(cl, pc, net, mixed):
        if net is not None:
            learned.attach(cl, net)
        a = torch.stack([cl.plane_wave([0, 0, 1], 1), cl.plane_wave([0, 0, 1], 0), cl.plane_wave([0, 0, -1], 1),
                         cl.plane_wave([0, 0, -1], 0)], dim=1)
        _, conv = cl.solve(a, rtol=1e-8, restart=50, max_iterations=2000, preconditioner=pc,
                           mixed_precision="matrix_free" if mixed else False)
        return max(c[0] for c in conv)

    methods = [("none", "none", None), ("block_jacobi", "block_jacobi", None)] + \
        [(name, "learned", net) for name, net in nets.items()]

    def evaluate(lmax, k0, radii, eps, pos):
        row = {}
        cl = pkg.IterativeSphereCluster(lmax, k0, radii, eps, pos)
        for mixed in (True, False):
            for name, pc, net in methods:
                cl.reset()
                row[("mixed_" if mixed else "fp64_") + name] = iterations(cl, pc, net, mixed)
        del cl
        return row

    rows = []
    for s in range(args.eval):
        lmax, k0, radii, eps, pos = learned.training_cluster(9000 + s)
        row = {"seed": 9000 + s, "N": int(radii.size), "lmax": lmax, **evaluate(lmax, k0, radii, eps, pos)}
        rows.append(row)
        print(f"[eval] {row}", flush=True)
    lmax, radii, eps, pos = pkg.workload("n1024_l6", seed=0)
    big = {"workload": "n1024_l6", **evaluate(lmax, 1.3, radii, eps, pos)}
    print(f"[eval] {big}", flush=True)
    keys = [k for k in rows[0] if k.startswith(("mixed_", "fp64_"))]
    report = {"train_clusters": args.clusters, "epochs": args.epochs, "train_seconds": train_s,
              "distribution": "N U{64..1024}, lmax U{3..6}, k0 U[0.8,2], radii U[0.15,0.35], "
                              "eps U[2,6]+iU[0,0.2], fill U[0.05,0.25]; training seeds 1000+, held-out 9000+",
              "heldout": rows, "n1024_l6": big,
              "mean_iterations": {k: float(np.mean([r[k] for r in rows])) for k in keys}}
    print(json.dumps(report), flush=True)
    if args.report:
        with open(args.report, "w") as f:
            json.dump(report, f, indent=1)


if __name__ == "__main__":
    main()
