"""Differentiable cluster entry points with the shape of the reference's ``treams_rs.diff``
(python/treams_rs/diff.py): each returns the value and a context whose ``pullback``
and ``pushforward`` run on the device.

* ``sphere_cluster(lmax, k0, radii, epsilon, positions)`` -> (X, SphereClusterContext):
  the full interacting T-matrix X = (I - T C)^{-1} T (diff.py:1435-1470,
  cluster/spheres.rs:56-80, its residual cluster/spheres.rs:263-378).
* ``particle_cluster(local, positions, k0, lmax=... | modes=...)`` ->
  (X, ParticleClusterContext): X of particles with arbitrary local T-matrices over the
  concatenated bases (diff.py:1473-1523, cluster/particles.rs:65-250); spherical waves,
  helicity basis, one real wavenumber.
* ``interaction(local, coupling)`` -> (X, InteractionContext): X = (I - T C)^{-1} T for
  dense T and C (diff.py:1638-1655, cluster/interaction.rs:330-440).

The LU factors (this library's HIP kernels) stay with the context, so pullbacks and
pushforwards reuse them.  ``interaction`` forms I - T C and the products of its
derivatives with torch matrix products (library GEMMs); every solve is this library's.
"""
from __future__ import annotations

import numpy as np
import torch

from .cluster import (
    DenseFactor,
    IterativeSphereCluster,
    SphereClusterFactor,
    _device,
    mode_index,
)


def _eye(n: int, dev) -> torch.Tensor:
    return torch.eye(n, dtype=torch.complex128, device=dev)


class SphereClusterContext:
    """Reusable derivative context of ``sphere_cluster`` (cluster/spheres.rs:230-378)."""

    def __init__(self, factor: SphereClusterFactor, x: torch.Tensor):
        self.factor = factor
        self.cluster = factor.cluster
        self._eye = _eye(self.cluster.n, x.device)
        self._x, self._ctx, _ = factor.record(self._eye)

    def pushforward(self, k0=0.0, radii=None, epsilon=None, positions=None) -> torch.Tensor:
        """dX along tangents of k0, the radii, permittivities and positions
        (SphereClusterResidual::pushforward, cluster/spheres.rs:263-300)."""
        dx, _ = self._ctx.pushforward(k0, radii, epsilon, positions)
        return dx

    def pullback(self, cotangent):
        """(k0, radii[N], epsilon[N], positions[N, 3]) gradients of Re<G, X>
        (SphereClusterResidual::pullback, cluster/spheres.rs:321-378)."""
        gk, gr, ge, gp, _, _ = self._ctx.pullback(torch.as_tensor(cotangent).to(self._x.device, torch.complex128))
        return gk, gr, ge, gp


def sphere_cluster(lmax: int, k0: float, radii, epsilon, positions, *, precision: str = "c128",
                   rtol: float = 1e-12):
    """Coupled T-matrix of homogeneous nonmagnetic spheres in vacuum, (n, n) with
    n = spheres * 2 lmax (lmax + 2), sphere by sphere in the reference mode order,
    and its derivative context."""
    cl = IterativeSphereCluster(lmax, k0, radii, epsilon, positions)
    f = SphereClusterFactor(cl, precision, rtol)
    ctx = SphereClusterContext(f, torch.empty(0, device=cl._dev))
    return ctx._x, ctx


class ParticleClusterContext:
    """Reusable derivative context of ``particle_cluster``
    (ParticleClusterResidual, cluster/particles.rs:158-250)."""

    def __init__(self, factor, keep, own, full_shape):
        self.factor = factor
        self.cluster = factor.cluster
        self._keep = keep            # global rows / columns of the particles' own modes
        self._own = own              # per particle: indices of its modes in the full block
        dev = self.cluster._dev
        n = self.cluster.n
        inc = torch.zeros((n, keep.numel()), dtype=torch.complex128, device=dev)
        inc[keep, torch.arange(keep.numel(), device=dev)] = 1.0
        self._inc = inc
        self._xfull, self._ctx, _ = factor.record(inc)
        self.value = self._xfull[keep]
        self._M = full_shape

    def _embed_cols(self, g):
        """A (keep x keep) matrix as (n x keep) rows in the full basis."""
        out = torch.zeros((self.cluster.n, self._keep.numel()), dtype=torch.complex128, device=self._xfull.device)
        out[self._keep] = torch.as_tensor(g).to(out.device, torch.complex128)
        return out

    def pushforward(self, local=None, positions=None, k0: float = 0.0) -> torch.Tensor:
        """dX along tangents of the local T-matrices (one per particle, in its own
        basis), the positions and the wavenumber (cluster/particles.rs:206-220)."""
        full = None
        if local is not None:
            M = self._M
            full = torch.zeros((self.cluster.N, M, M), dtype=torch.complex128, device=self._xfull.device)
            for i, (t, idx) in enumerate(zip(local, self._own)):
                ix = torch.as_tensor(idx, device=full.device)
                full[i][ix[:, None], ix[None, :]] = torch.as_tensor(t).to(full.device, torch.complex128)
        dx, _ = self._ctx.pushforward_local(k0, positions, full, None)
        return dx[self._keep]

    def pullback(self, cotangent):
        """(local gradients [one per particle, own basis], positions[N, 3], k0) of
        Re<G, X> (cluster/particles.rs:231-250)."""
        gk, gp, gl, _, _ = self._ctx.pullback_local(self._embed_cols(cotangent))
        local = []
        for i, idx in enumerate(self._own):
            ix = torch.as_tensor(idx, device=gl.device)
            local.append(gl[i][ix[:, None], ix[None, :]])
        return local, gp, gk


def parity_change(modes) -> np.ndarray:
    """The real orthogonal, symmetric change between helicity and parity coefficients of
    one basis (special/polarization.rs:1-60): helicity waves are (N +- M) / sqrt(2)
    (pol 1 positive, pol 0 negative), parity waves M (pol 0) and N (pol 1), so
    b_N = (a_+ + a_-) / sqrt(2), b_M = (a_+ - a_-) / sqrt(2) for each (l, m).  The same
    matrix maps parity to helicity; every (l, m) needs both polarisations."""
    index = {tuple(int(v) for v in md): k for k, md in enumerate(modes)}
    U = np.zeros((len(modes), len(modes)))
    r = 1.0 / np.sqrt(2.0)
    for (l, m, pol), k in index.items():
        other = index.get((l, m, 1 - pol))
        if other is None:
            raise ValueError(f"the parity basis needs both polarisations of (l, m) = ({l}, {m})")
        # row k: parity pol (1: N, 0: M) from helicity pols (1: +, 0: -)
        plus = k if pol == 1 else other
        minus = other if pol == 1 else k
        U[k, plus] = r
        U[k, minus] = r if pol == 1 else -r
    return U


class _ParityContext:
    """A particle-cluster context in the parity basis: the helicity context's value,
    tangents and cotangents conjugated by the per-particle basis change."""

    def __init__(self, ctx, changes):
        self._ctx = ctx
        self._U = changes
        big = torch.block_diag(*[torch.as_tensor(u, dtype=torch.complex128) for u in changes])
        self._big = big.to(ctx.value.device)
        self.value = self._big @ ctx.value @ self._big

    def pushforward(self, local=None, positions=None, k0: float = 0.0) -> torch.Tensor:
        lh = None if local is None else [u @ np.asarray(t.cpu() if torch.is_tensor(t) else t) @ u
                                         for t, u in zip(local, self._U)]
        return self._big @ self._ctx.pushforward(lh, positions, k0) @ self._big

    def pullback(self, cotangent):
        g = torch.as_tensor(cotangent).to(self._big.device, torch.complex128)
        local, gp, gk = self._ctx.pullback(self._big @ g @ self._big)
        out = []
        for t, u in zip(local, self._U):
            ut = torch.as_tensor(u, dtype=torch.complex128, device=t.device)
            out.append(ut @ t @ ut)
        return out, gp, gk


def particle_cluster(local, positions, k0: float, *, lmax=None, modes=None, poltype: str = "helicity",
                     precision: str = "c128", rtol: float = 1e-12):
    """Coupled T-matrix of particles with arbitrary local T-matrices and its derivative
    context.  ``local``: one (n_i, n_i) T-matrix per particle in its basis, either the
    full basis up to lmax[i] (reference mode order) or the mode list modes[i] of
    (l, m, pol); the particles are embedded in the basis up to the largest degree with
    zero rows and columns (exactly the reduced system).  ``poltype`` "parity" takes and
    returns every matrix in the parity basis (diff.py:1473-1523); the device computes in
    the helicity basis."""
    if poltype not in ("helicity", "parity"):
        raise ValueError("poltype must be 'helicity' or 'parity'")
    if poltype == "parity":
        pos = np.asarray(positions, np.float64).reshape(-1, 3)
        if modes is None:
            lm = np.broadcast_to(np.asarray(lmax, int), (pos.shape[0],))
            modes = [[(l, m, p) for l in range(1, int(li) + 1) for m in range(-l, l + 1) for p in (1, 0)]
                     for li in lm]
        changes = [parity_change(ms) for ms in modes]
        hel = [u @ np.asarray(t.cpu() if torch.is_tensor(t) else t, np.complex128) @ u for t, u in zip(local, changes)]
        _, ctx = particle_cluster(hel, pos, k0, modes=modes, precision=precision, rtol=rtol)
        pctx = _ParityContext(ctx, changes)
        return pctx.value, pctx
    pos = np.asarray(positions, np.float64).reshape(-1, 3)
    N = pos.shape[0]
    if len(local) != N or N == 0:
        raise ValueError("one local matrix and position required per particle")
    if (lmax is None) == (modes is None):
        raise ValueError("give either lmax (one per particle) or modes (one list per particle)")
    if modes is None:
        lm = np.broadcast_to(np.asarray(lmax, int), (N,))
        modes = [[(l, m, p) for l in range(1, int(li) + 1) for m in range(-l, l + 1) for p in (1, 0)] for li in lm]
    L = max(l for ms in modes for (l, _, _) in ms)
    M = 2 * L * (L + 2)
    full = np.zeros((N, M, M), np.complex128)
    keep, own = [], []
    for i, (t, ms) in enumerate(zip(local, modes)):
        t = np.asarray(t.cpu() if torch.is_tensor(t) else t, np.complex128)
        idx = np.array([mode_index(l, m, p) for (l, m, p) in ms], int)
        if t.shape != (len(idx), len(idx)) or len(set(idx.tolist())) != len(idx):
            raise ValueError("local matrix shape must match its basis, modes distinct")
        full[i][np.ix_(idx, idx)] = t
        keep.extend((i * M + idx).tolist())
        own.append(idx)
    cl = IterativeSphereCluster.from_tmatrices(L, k0, pos, full)
    f = SphereClusterFactor(cl, precision, rtol)
    ctx = ParticleClusterContext(f, torch.as_tensor(np.asarray(keep, int), device=cl._dev), own, M)
    return ctx.value, ctx


class InteractionContext:
    """Reusable derivative context of ``interaction`` (InteractionResidual,
    cluster/interaction.rs:330-440)."""

    def __init__(self, local, coupling, precision, rtol):
        self.local = local
        self.coupling = coupling
        n = local.shape[0]
        eye = _eye(n, local.device)
        self.factor = DenseFactor(eye - local @ coupling, precision, rtol)
        self.value, _ = self.factor.solve(local)

    def pushforward(self, local, coupling) -> torch.Tensor:
        """dX = (I - T C)^{-1} (dT (I + C X) + T dC X) (cluster/interaction.rs:359-368)."""
        dt = torch.as_tensor(local).to(self.local.device, torch.complex128)
        dc = torch.as_tensor(coupling).to(self.local.device, torch.complex128)
        x = self.value
        response = torch.eye(x.shape[0], dtype=x.dtype, device=x.device) + self.coupling @ x
        rhs = dt @ response + self.local @ (dc @ x)
        dx, _ = self.factor.solve(rhs)
        return dx

    def pullback(self, cotangent):
        """(T-bar, C-bar) of Re<G, X>: Y = (I - T C)^{-H} G, T-bar = Y (I + C X)^H,
        C-bar = T^H Y X^H (cluster/interaction.rs:379-440)."""
        g = torch.as_tensor(cotangent).to(self.local.device, torch.complex128)
        y, _ = self.factor.solve(g, adjoint=True)
        x = self.value
        response = torch.eye(x.shape[0], dtype=x.dtype, device=x.device) + self.coupling @ x
        tbar = y @ response.conj().T
        cbar = (self.local.conj().T @ y) @ x.conj().T
        return tbar, cbar


def interaction(local, coupling, *, precision: str = "c128", rtol: float = 1e-12):
    """X = (I - T C)^{-1} T for dense T and C (n x n), and its derivative context."""
    t = torch.as_tensor(local).to(_device(), torch.complex128)
    c = torch.as_tensor(coupling).to(_device(), torch.complex128)
    if t.ndim != 2 or t.shape[0] != t.shape[1] or c.shape != t.shape:
        raise ValueError("local and coupling must be square matrices of one shape")
    ctx = InteractionContext(t, c, precision, rtol)
    return ctx.value, ctx
