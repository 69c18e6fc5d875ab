"""CPU oracle of the learned preconditioner -- test infrastructure only.

* MLP forward/backward in NumPy with the same BF16 rounding (nearest even) of both
  operands of every product as the GPU kernels, accumulated in FP64.
* The dense preconditioner M^{-1} = I + T sum_p D(alpha_p) C_p and its training loss
  |A M^{-1} V - V|^2 / |V|^2 with A = I - T C from oracle.Cluster (FP64, LAPACK-free).

There is no reference counterpart (the reference GMRES is unpreconditioned); the
gradient checks use central finite differences of these functions.
"""
from __future__ import annotations

import numpy as np


def bf16(x: np.ndarray) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params: np.ndarray, dims):
    """Layer list [(W (out x in), b)] from the flat parameter vector."""
    layers, off = [], 0
    for din, dout in zip(dims[:-1], dims[1:]):
        W = params[off:off + din * dout].reshape(dout, din)
        off += din * dout
        b = params[off:off + dout]
        off += dout
        layers.append((W.astype(np.float64), b.astype(np.float64)))
    return layers


def forward(params, dims, X):
    layers = unpack(params, dims)
    H = [np.asarray(X, np.float64)]
    Z = []
    for l, (W, b) in enumerate(layers):
        z = bf16(H[-1].astype(np.float32)) @ bf16(W.astype(np.float32)).T + b
        Z.append(z)
        if l < len(layers) - 1:
            H.append(np.maximum(z, 0.0).astype(np.float32).astype(np.float64))
    return Z[-1], (H, Z)


def backward(params, dims, cache, dY):
    layers = unpack(params, dims)
    H, Z = cache
    grads = []
    dz = np.asarray(dY, np.float64)
    for l in range(len(layers) - 1, -1, -1):
        W, _ = layers[l]
        dW = bf16(dz.astype(np.float32)).T @ bf16(H[l].astype(np.float32))
        db = dz.sum(axis=0)
        grads.append((dW, db))
        if l > 0:
            dh = bf16(dz.astype(np.float32)) @ bf16(W.astype(np.float32))
            dz = dh * (Z[l - 1] > 0)
    grads.reverse()
    return np.concatenate([np.concatenate([dW.reshape(-1), db]) for dW, db in grads])


def learned_inverse(cluster, pairs, out16):
    """Dense M^{-1} = I + T sum_p D(alpha_p) C_p for pairs (i, j) and network outputs."""
    import oracle

    lmax, M, n = cluster.lmax, cluster.M, cluster.n
    l_of_mode = oracle.modes(lmax)[0]
    S = np.zeros((n, n), np.complex128)
    for p, (i, j) in enumerate(pairs):
        C = oracle.translation_block(lmax, cluster.k0, cluster.positions[i] - cluster.positions[j], False)
        alpha = (1.0 + 0.5 * np.tanh(out16[p, 0:2 * lmax:2].astype(np.float64))
                 + 0.5j * np.tanh(out16[p, 1:2 * lmax:2].astype(np.float64)))
        S[i * M:(i + 1) * M, j * M:(j + 1) * M] += alpha[l_of_mode - 1][:, None] * C
    return np.eye(n) + cluster.T @ S


def learned_loss(cluster, pairs, out16, V):
    Minv = learned_inverse(cluster, pairs, out16)
    R = cluster.operator() @ (Minv @ V) - V
    return float(np.sum(np.abs(R) ** 2) / np.sum(np.abs(V) ** 2))
