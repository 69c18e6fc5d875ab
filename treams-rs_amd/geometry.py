"""Synthetic sphere clusters: radii first, then centres rejection-sampled uniformly in
a cube to a minimum surface gap (numpy PCG64, fixed seeds)."""
from __future__ import annotations

import numpy as np


def random_cluster(N: int, side: float, gap: float, radius_range=(0.2, 0.3), eps=None, eps_real=(2.0, 6.0),
                   eps_imag=(0.0, 0.2), seed: int = 0, max_tries: int = 2_000_000):
    """Return (radii[N], epsilon[N] complex, positions[N, 3]).

    eps: a fixed complex permittivity, or None for real U[eps_real] + i U[eps_imag].
    """
    rng = np.random.default_rng(seed)
    radii = rng.uniform(radius_range[0], radius_range[1], N)
    if eps is None:
        epsilon = rng.uniform(*eps_real, N) + 1j * rng.uniform(*eps_imag, N)
    else:
        epsilon = np.full(N, complex(eps))
    pos = np.zeros((N, 3))
    # uniform grid of cells for fast rejection
    cell = 2 * radius_range[1] + gap
    ncell = max(1, int(side / cell))
    grid: dict = {}
    tries = 0
    for i in range(N):
        while True:
            tries += 1
            if tries > max_tries:
                raise RuntimeError("could not place spheres: cube too small for the gap")
            c = rng.uniform(-side / 2, side / 2, 3)
            key = tuple(np.minimum(((c + side / 2) / side * ncell).astype(int), ncell - 1))
            ok = True
            for dx in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    for dz in (-1, 0, 1):
                        for j in grid.get((key[0] + dx, key[1] + dy, key[2] + dz), ()):
                            if np.linalg.norm(c - pos[j]) < radii[i] + radii[j] + gap:
                                ok = False
                                break
                        if not ok:
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if ok:
                pos[i] = c
                grid.setdefault(key, []).append(i)
                break
    return radii, epsilon, pos


#: The workloads (N, lmax, cube side, gap, radius range, permittivity) of the cluster path.
WORKLOADS = {
    "n8_l3": dict(N=8, lmax=3, side=2.0, gap=0.1, radius_range=(0.2, 0.3), eps=4 + 0.1j),
    "n512_l4": dict(N=512, lmax=4, side=12.0, gap=0.1, radius_range=(0.2, 0.3), eps=4 + 0.1j),
    "n256_l4_grad": dict(N=256, lmax=4, side=9.0, gap=0.1, radius_range=(0.2, 0.3), eps=None),
    "n1024_l6": dict(N=1024, lmax=6, side=16.0, gap=0.05, radius_range=(0.15, 0.35), eps=None),
}


def workload(name: str, seed: int = 0):
    w = WORKLOADS[name]
    radii, eps, pos = random_cluster(w["N"], w["side"], w["gap"], w["radius_range"], w["eps"], seed=seed)
    return w["lmax"], radii, eps, pos
