"""Learned near-field preconditioner: network, training on GPU, weights on disk.

The preconditioner (libtreams_gpu.so, tg_learned.hip) is
    M^{-1} x = x + T sum_{(i,j) in k-nearest pairs} D(alpha_ij) C_ij x_j,
alpha_ij[l] = 1 + out[2l] + i out[2l+1], out = MLP(features_ij).  The network is the
library's MFMA MLP (BF16 operands, FP32 accumulation); training minimises
|A M^{-1} V - V|^2 / |V|^2 over random probe blocks V with gradients from
tg_cluster_learned_loss and tg_mlp_backward, and Adam on the device.
"""
from __future__ import annotations

import ctypes
import os
import time

import numpy as np
import torch

from ._lib import check, lib
from .cluster import IterativeSphereCluster, _columns, _device, _stream_ptr
from .geometry import random_cluster

FEATURES = 64
OUTPUTS = 16
WIDTH = 256
HIDDEN_LAYERS = 3
K_NEAREST = 8
HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_WEIGHTS = os.path.join(HERE, "weights", "pc_mlp.npz")


class MLP:
    """FEATURES -> WIDTH (x HIDDEN_LAYERS, ReLU) -> OUTPUTS on the GPU."""

    def __init__(self, n_in=FEATURES, width=WIDTH, n_out=OUTPUTS, hidden_layers=HIDDEN_LAYERS, seed=0):
        _device()
        h = ctypes.c_void_p()
        check(lib().tg_mlp_create(n_in, width, n_out, hidden_layers, seed, ctypes.c_void_p(_stream_ptr()),
                                  ctypes.byref(h)))
        self._h = h
        self._destroy = lib().tg_mlp_destroy  # bound now: module globals are gone at exit
        self.dims = (n_in, width, n_out, hidden_layers)
        self.count = int(lib().tg_mlp_param_count(h))
        self.step = 0

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._destroy(h)
            self._h = None

    def params(self) -> np.ndarray:
        out = np.zeros(self.count, np.float32)
        check(lib().tg_mlp_get_params(self._h, out.ctypes.data))
        return out

    def set_params(self, values) -> None:
        v = np.ascontiguousarray(values, np.float32)
        assert v.size == self.count
        check(lib().tg_mlp_set_params(self._h, v.ctypes.data))

    def grads(self) -> np.ndarray:
        out = np.zeros(self.count, np.float32)
        check(lib().tg_mlp_get_grads(self._h, out.ctypes.data))
        return out

    def zero_grad(self):
        check(lib().tg_mlp_zero_grad(self._h))

    def forward(self, x: torch.Tensor) -> torch.Tensor:
        x = x.to(_device(), torch.float32).contiguous()
        y = torch.empty((x.shape[0], self.dims[2]), dtype=torch.float32, device=x.device)
        check(lib().tg_mlp_forward(self._h, x.data_ptr(), x.shape[0], y.data_ptr()))
        return y

    def backward(self, dy: torch.Tensor) -> None:
        dy = dy.to(_device(), torch.float32).contiguous()
        check(lib().tg_mlp_backward(self._h, dy.data_ptr()))

    def adam(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        self.step += 1
        check(lib().tg_mlp_adam_step(self._h, lr, beta1, beta2, eps, self.step))

    def save(self, path=DEFAULT_WEIGHTS, **meta):
        os.makedirs(os.path.dirname(path), exist_ok=True)
        np.savez(path, params=self.params(), dims=np.array(self.dims), **meta)

    @classmethod
    def load(cls, path=DEFAULT_WEIGHTS) -> "MLP":
        with np.load(path, allow_pickle=False) as z:
            dims = tuple(int(v) for v in z["dims"])
            m = cls(*dims)
            m.set_params(z["params"])
        return m


def near_pairs(cluster: IterativeSphereCluster, k_nearest: int = K_NEAREST) -> int:
    npairs = ctypes.c_int64()
    check(lib().tg_cluster_near_pairs(cluster._h, k_nearest, ctypes.byref(npairs)))
    return npairs.value


def pair_list(cluster: IterativeSphereCluster, npairs: int) -> np.ndarray:
    out = np.zeros(2 * npairs, np.int32)
    check(lib().tg_cluster_pair_list(cluster._h, out.ctypes.data))
    return out.reshape(npairs, 2)


def pair_features(cluster: IterativeSphereCluster, npairs: int) -> torch.Tensor:
    f = torch.empty((npairs, FEATURES), dtype=torch.float32, device=_device())
    check(lib().tg_cluster_pair_features(cluster._h, f.data_ptr()))
    return f


def set_pair_output(cluster: IterativeSphereCluster, out: torch.Tensor) -> None:
    out = out.to(_device(), torch.float32).contiguous()
    check(lib().tg_cluster_set_pair_output(cluster._h, out.data_ptr()))


def learned_apply(cluster: IterativeSphereCluster, x, adjoint=False, inner=False):
    """M^{-1} x (M^{-H} x with adjoint); inner: as the mixed-precision inner solves apply
    it (same-helicity near blocks stored as bfloat16 pairs, FP32 products)."""
    cm, vector = _columns(x, cluster.n)
    y = torch.empty_like(cm)
    mode = 2 if inner else int(adjoint)
    check(lib().tg_cluster_learned_apply(cluster._h, cm.data_ptr(), y.data_ptr(), cm.shape[0], mode))
    return y[0] if vector else y.T.contiguous()


def learned_loss(cluster: IterativeSphereCluster, probes, npairs: int, grad: bool = True):
    cm, _ = _columns(probes, cluster.n)
    loss = ctypes.c_double()
    dout = torch.empty((npairs, OUTPUTS), dtype=torch.float32, device=_device()) if grad else None
    check(lib().tg_cluster_learned_loss(cluster._h, cm.data_ptr(), cm.shape[0], ctypes.byref(loss),
                                        dout.data_ptr() if grad else None))
    return loss.value, dout


def attach(cluster: IterativeSphereCluster, mlp: MLP | None, k_nearest: int = K_NEAREST) -> None:
    """Use `mlp` for TG_PC_LEARNED solves of `cluster`."""
    if mlp is not None:
        near_pairs(cluster, k_nearest)
    check(lib().tg_cluster_set_mlp(cluster._h, mlp._h if mlp is not None else None))
    cluster._mlp = mlp  # keep the network alive as long as the cluster uses it


def training_cluster(seed: int, rng: np.random.Generator | None = None, n_range=(64, 1024)):
    """One training cluster of the training-set distribution: N uniform in
    {64..1024}, lmax in {3..6}, k0 U[0.8, 2.0], radii U[0.15, 0.35], eps U[2,6] +
    i U[0,0.2], volume fill fraction U[0.05, 0.25]."""
    rng = rng or np.random.default_rng(seed)
    N = int(rng.integers(n_range[0], n_range[1] + 1))
    lmax = int(rng.integers(3, 7))
    k0 = float(rng.uniform(0.8, 2.0))
    fill = float(rng.uniform(0.05, 0.25))
    mean_volume = 4.0 / 3.0 * np.pi * np.mean(np.array([0.15, 0.35]) ** 3)
    side = (N * mean_volume / fill) ** (1.0 / 3.0)
    radii, eps, pos = random_cluster(N, side, 0.05, (0.15, 0.35), None, seed=seed)
    return lmax, k0, radii, eps, pos


def probe_block(cluster: IterativeSphereCluster, probes: int, seed: int) -> torch.Tensor:
    """Probe columns in the range of T (T times complex Gaussians): GMRES on
    (I - T C) x = T a only visits that range, where high-degree source modes carry
    the tiny weights of their Mie coefficients."""
    gen = torch.Generator(device=_device())
    gen.manual_seed(seed)
    G = torch.randn((cluster.n, probes), dtype=torch.complex128, device=_device(), generator=gen)
    V = cluster.local(G)
    return V / torch.linalg.norm(V, dim=0, keepdim=True)


def dominant_probes(cluster: IterativeSphereCluster, probes: int, seed: int, steps: int = 4) -> torch.Tensor:
    """Power iteration of E = A M^{-1} - I from T-range probes: columns close to the
    dominant error modes of the current preconditioner, which limit GMRES."""
    V = probe_block(cluster, probes, seed)
    for _ in range(steps):
        W = cluster.apply(learned_apply(cluster, V)) - V
        V = W / torch.linalg.norm(W, dim=0, keepdim=True)
    return V


def error_radius(cluster: IterativeSphereCluster, probes: int = 4, seed: int = 0, steps: int = 12) -> float:
    """Power-iteration estimate of the spectral radius of A M^{-1} - I for the weights
    currently set on `cluster`."""
    V = dominant_probes(cluster, probes, seed, steps)
    W = cluster.apply(learned_apply(cluster, V)) - V
    return float(torch.linalg.norm(W, dim=0).max())


def train(n_clusters=48, epochs=6, probes=4, lr=1e-3, seed=1000, log=print, mlp: MLP | None = None,
          n_range=(64, 1024), lr_final=None, log_every=256) -> MLP:
    """Train on clusters the method generates (seeds seed..seed+n_clusters-1, drawn
    from the training-set distribution; a new order every epoch), one Adam step per
    cluster on the dominant error modes of the current preconditioner.  The learning
    rate decays on a cosine from `lr` to `lr_final` (default lr / 10) over all steps."""
    mlp = mlp or MLP(seed=seed)
    clusters = [training_cluster(seed + c, n_range=n_range) for c in range(n_clusters)]
    lr_final = lr / 10 if lr_final is None else lr_final
    total = max(1, epochs * n_clusters)
    order_rng = np.random.default_rng(seed)
    done = 0
    for epoch in range(epochs):
        t0 = time.perf_counter()
        losses, base = [], []
        order = order_rng.permutation(n_clusters) if epoch else np.arange(n_clusters)
        for k, c in enumerate(order):
            lmax, k0, radii, eps, pos = clusters[c]
            cl = IterativeSphereCluster(lmax, k0, radii, eps, pos)
            npairs = near_pairs(cl)
            feats = pair_features(cl, npairs)
            out = mlp.forward(feats)
            set_pair_output(cl, out)
            V = dominant_probes(cl, probes, seed * 7919 + int(c) + 104729 * epoch)
            loss, dout = learned_loss(cl, V, npairs)
            mlp.zero_grad()
            mlp.backward(dout)
            rate = lr_final + 0.5 * (lr - lr_final) * (1.0 + np.cos(np.pi * done / total))
            mlp.adam(lr=float(rate))
            done += 1
            losses.append(loss)
            if epoch == 0:
                set_pair_output(cl, torch.zeros_like(out))
                base.append(learned_loss(cl, V, npairs, grad=False)[0])
            del cl
            if log_every and (k + 1) % log_every == 0:
                log(f"epoch {epoch + 1} cluster {k + 1}/{n_clusters}: mean loss {np.mean(losses[-log_every:]):.4e}"
                    + (f", alpha = 1: {np.mean(base[-log_every:]):.4e}" if base else "")
                    + f" ({time.perf_counter() - t0:.1f} s)")
        msg = f"epoch {epoch + 1}/{epochs}: mean loss {np.mean(losses):.4e} ({time.perf_counter() - t0:.1f} s)"
        if base:
            msg += f"; plain Neumann (alpha = 1) {np.mean(base):.4e}"
        log(msg)
    return mlp
