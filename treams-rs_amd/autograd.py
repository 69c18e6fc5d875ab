"""PyTorch adapter: the cluster solve, the incident plane wave and the cross sections
as ``torch.autograd.Function``s whose forward and backward run in libtreams_gpu.so.

Counterpart of the reference's ``_Execute`` adapter (python/treams_rs/torch.py:102-168),
which records a native computation in the forward and calls its pullback in the
backward; here the native computation is the device solve and the pullback is the
device adjoint (LU factors reused, or adjoint GMRES).  Gradients follow PyTorch's
convention for complex inputs (dL/dRe + i dL/dIm), which is the library's.

Differentiate a scattering cross section::

    cluster = TorchCluster(lmax, k0, radii, epsilon, positions)   # tensors with grad
    a = cluster.plane_wave([0, 0, 1], 1)
    x = cluster.solve(a, method="lu")
    sca, ext = cluster.cross_sections(a, x)
    sca.backward()          # radii.grad, positions.grad, epsilon.grad

The saved record keeps the LU factors (or the GMRES options) until the graph is
freed; a second backward (``retain_graph=True``) reuses them.  Forward mode
(``torch.autograd.forward_ad``) runs the record's pushforward: one tangent solve.
"""
from __future__ import annotations

import numpy as np
import torch

from .cluster import IterativeSphereCluster, SphereClusterFactor


def _np(t, dtype):
    return np.ascontiguousarray(torch.as_tensor(t).detach().cpu().numpy().astype(dtype))


def _meta(t: torch.Tensor):
    """What a gradient must match: (shape, dtype, device) of an input."""
    return (tuple(t.shape), t.dtype, t.device)


def _like(value, meta):
    """Gradient array as a tensor of the input's shape, dtype and device."""
    shape, dtype, device = meta
    t = torch.as_tensor(value)
    if not dtype.is_complex and t.is_complex():
        t = t.real
    return t.to(device=device, dtype=dtype).reshape(shape)


class _Solve(torch.autograd.Function):
    """X = (I - T C)^{-1} T A as a function of radii, positions, epsilon and A."""

    @staticmethod
    def forward(ctx, radii, positions, epsilon, incident, cluster, method, precision, options):
        h = cluster.handle
        if method == "lu":
            factor = SphereClusterFactor(h, precision, options.get("rtol", 1e-12))
            x, record, _ = factor.record(incident.detach())
        elif method == "gmres":
            x, record, _ = h.record(incident.detach(), **options)
        else:
            raise ValueError(f"method must be 'lu' or 'gmres', not {method!r}")
        ctx.record = record
        ctx.metas = tuple(_meta(t) for t in (radii, positions, epsilon, incident))
        return x

    @staticmethod
    def jvp(ctx, tr, tp, te, ti, *_):
        """Forward mode: one tangent solve (the record's pushforward,
        cluster/iterative.rs:362-447 or cluster/spheres.rs:263-300)."""
        def host(t, dtype):
            return None if t is None else _np(t, dtype)

        ti = None if ti is None else ti.detach().to(torch.complex128)
        dx, _ = ctx.record.pushforward(0.0, host(tr, np.float64), host(te, np.complex128),
                                       host(tp, np.float64), ti)
        return dx

    @staticmethod
    def backward(ctx, gx):
        mr, mp, me, mi = ctx.metas
        gk, gr, ge, gp, ginc, _ = ctx.record.pullback(gx.detach().to(torch.complex128))
        return (_like(gr, mr) if ctx.needs_input_grad[0] else None,
                _like(gp, mp) if ctx.needs_input_grad[1] else None,
                _like(ge, me) if ctx.needs_input_grad[2] else None,
                _like(ginc, mi) if ctx.needs_input_grad[3] else None,
                None, None, None, None)


class _PlaneWave(torch.autograd.Function):
    """Regular coefficients of a plane wave at the sphere positions: a_j = a_j(0)
    e^{i k.r_j} (pw/expand.rs:23-66 with the phase of pw/field.rs:20-26)."""

    @staticmethod
    def forward(ctx, positions, cluster, direction, pol):
        a = cluster.handle.plane_wave(direction, pol)
        d = np.asarray(direction, np.float64)
        ctx.kvec = cluster.handle.k0 * d / np.linalg.norm(d)
        ctx.save_for_backward(a)
        ctx.primal = a  # for jvp (forward mode reads no saved tensors)
        ctx.N = cluster.handle.N
        ctx.pos_meta = _meta(positions)
        return a

    @staticmethod
    def jvp(ctx, tp, *_):
        a = ctx.primal
        if tp is None:
            return torch.zeros_like(a)
        # da_j = i (k . dr_j) a_j
        kd = tp.detach().to(torch.float64).reshape(ctx.N, 3) @ torch.as_tensor(ctx.kvec, device=tp.device)
        return (1j * kd.to(a.device)[:, None] * a.reshape(ctx.N, -1)).reshape(a.shape)

    @staticmethod
    def backward(ctx, ga):
        (a,) = ctx.saved_tensors
        if not ctx.needs_input_grad[0]:
            return None, None, None, None
        # dL = Re sum conj(ga) i k.dr_j a_j
        phase = torch.sum(torch.conj(ga.reshape(ctx.N, -1)) * 1j * a.reshape(ctx.N, -1), dim=1).real
        g = phase.detach().cpu().numpy()[:, None] * ctx.kvec[None, :]
        return _like(g, ctx.pos_meta), None, None, None


class _CrossSections(torch.autograd.Function):
    """(sigma_sca, sigma_ext) of one illumination
    (python/treams_rs/_framework_tmatrix.py:231-289): sca = Re(x^H R x) s,
    ext = -Re(a^H x) s with s = 0.5 / (flux k^2) and the regular overlap R."""

    @staticmethod
    def forward(ctx, x, incident, positions, cluster, flux):
        h = cluster.handle
        sca, ext = h.cross_sections(incident.detach(), x.detach(), flux)
        ctx.cluster, ctx.flux = cluster, flux
        ctx.save_for_backward(x.detach(), incident.detach())
        ctx.pos_meta = _meta(positions)
        dev = x.device
        return (torch.tensor(sca, dtype=torch.float64, device=dev), torch.tensor(ext, dtype=torch.float64, device=dev))

    @staticmethod
    def backward(ctx, gsca, gext):
        x, a = ctx.saved_tensors
        h = ctx.cluster.handle
        s = 0.5 / (ctx.flux * h.k0 ** 2)
        gs = float(gsca) if gsca is not None else 0.0
        ge = float(gext) if gext is not None else 0.0
        gx = ga = gp = None
        if ctx.needs_input_grad[0]:
            gx = (h.overlap(x) + h.overlap(x, adjoint=True)) * (s * gs) - a * (s * ge)
        if ctx.needs_input_grad[1]:
            ga = -x * (s * ge)
        if ctx.needs_input_grad[2] and gs != 0.0:
            gp = _like(h.overlap_pullback(x * (s * gs), x), ctx.pos_meta)
        return gx, ga, gp, None, None


class TorchCluster:
    """Sphere cluster with tensor inputs; its solve, plane wave and cross sections are
    differentiable with respect to the radii, positions, permittivities and incident
    fields.  The device handle is built from the tensors' values at construction."""

    def __init__(self, lmax: int, k0: float, radii: torch.Tensor, epsilon, positions: torch.Tensor):
        self.radii = torch.as_tensor(radii)
        self.positions = torch.as_tensor(positions)
        eps = torch.as_tensor(epsilon)
        if eps.ndim == 0:
            eps = eps.expand(self.radii.shape)
        self.epsilon = eps
        self.handle = IterativeSphereCluster(lmax, k0, _np(self.radii, np.float64),
                                             _np(self.epsilon, np.complex128), _np(self.positions, np.float64))

    @property
    def n(self) -> int:
        return self.handle.n

    def plane_wave(self, direction, pol) -> torch.Tensor:
        return _PlaneWave.apply(self.positions, self, direction, pol)

    def solve(self, incident: torch.Tensor, method: str = "lu", precision: str = "c128", **options) -> torch.Tensor:
        """method "lu" (factors reused by the backward) or "gmres" (adjoint GMRES with
        the same options: rtol, restart, mixed_precision, preconditioner, ...)."""
        eps = self.epsilon if self.epsilon.is_complex() else self.epsilon.to(torch.complex128)
        return _Solve.apply(self.radii, self.positions, eps, incident, self, method, precision, options)

    def cross_sections(self, incident: torch.Tensor, x: torch.Tensor, flux: float = 0.5):
        return _CrossSections.apply(x, incident, self.positions, self, float(flux))
