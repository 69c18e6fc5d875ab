#!/usr/bin/env python
"""Train the learned near-field preconditioner on the GPU and evaluate it.

    python treams-rs_amd/train_preconditioner.py [--clusters 64] [--epochs 20] [--out PATH]

Training clusters: N in {64..256}, lmax in {3..6}, k0 U[0.8, 2.0], radii
U[0.15, 0.35], eps U[2,6] + i U[0,0.2], fill fraction U[0.05, 0.25], seeds 1000+.
Evaluation: held-out seeds 9000+, GMRES iterations to 1e-8 (FP64, restart 50) with
no preconditioner, block-Jacobi, the near-field preconditioner without the network
(alpha = 1) and with it; then the N = 1024, lmax = 6 workload.
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
    ap.add_argument("--clusters", type=int, default=64)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--probes", type=int, default=4)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--eval", type=int, default=8)
    ap.add_argument("--out", default=None)
    ap.add_argument("--report", default=None)
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

    untrained = learned.MLP(seed=1)  # zero last layer: alpha = 1

    def iterations(cl, pc, net):
        if net is not None:
            learned.attach(cl, net)
        a = torch.stack([cl.plane_wave([0, 0, 1], 1), cl.plane_wave([0, 0, -1], 0)], dim=1)
        t = time.perf_counter()
        _, conv = cl.solve(a, rtol=1e-8, restart=50, max_iterations=1000, preconditioner=pc)
        torch.cuda.synchronize()
        return max(c[0] for c in conv), time.perf_counter() - t

    rows = []
    for s in range(args.eval):
        lmax, k0, radii, eps, pos = learned.training_cluster(9000 + s)
        row = {"seed": 9000 + s, "N": int(radii.size), "lmax": lmax}
        for name, pc, net in [("none", "none", None), ("block_jacobi", "block_jacobi", None),
                              ("neumann_alpha1", "learned", untrained), ("learned", "learned", mlp)]:
            cl = pkg.IterativeSphereCluster(lmax, k0, radii, eps, pos)
            row[name] = iterations(cl, pc, net)[0]
            del cl
        rows.append(row)
        print(f"[eval] {row}", flush=True)
    lmax, radii, eps, pos = pkg.workload("n1024_l6", seed=0)
    big = {"workload": "n1024_l6"}
    for name, pc, net in [("none", "none", None), ("neumann_alpha1", "learned", untrained),
                          ("learned", "learned", mlp)]:
        cl = pkg.IterativeSphereCluster(lmax, 1.3, radii, eps, pos)
        its, sec = iterations(cl, pc, net)
        big[name] = its
        big[name + "_seconds"] = sec
        del cl
    print(f"[eval] {big}", flush=True)
    report = {"train_seconds": train_s, "heldout": rows, "n1024_l6": big,
              "mean_iterations": {k: float(np.mean([r[k] for r in rows]))
                                  for k in ("none", "block_jacobi", "neumann_alpha1", "learned")}}
    print(json.dumps(report), flush=True)
    if args.report:
        with open(args.report, "w") as f:
            json.dump(report, f, indent=1)


if __name__ == "__main__":
    main()
