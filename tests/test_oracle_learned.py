"""CPU pins of the oracle parts that the reference has no counterpart for: the MLP
oracle of the learned preconditioner (BF16 rounding against torch's own bfloat16
conversion, backward against torch autograd in FP64), the learned preconditioner in
its closed-form special case, and the matrix-free coupling apply of the CPU
baseline against the dense coupling matrix."""
import numpy as np
import torch

import mlp_oracle


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(size=4000).astype(np.float32) * 10.0 ** rng.integers(-6, 6, 4000),
                        np.array([1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -(1.0 + 2 ** -8)], np.float32)])
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(mlp_oracle.bf16(x), ref)


def _torch_mlp(params, dims, X):
    """The same network in torch FP64 with the weights and the layer inputs rounded to
    BF16 as values (no rounding inside the differentiated graph)."""
    off, h = 0, torch.tensor(X, dtype=torch.float64)
    Ws = []
    for l, (din, dout) in enumerate(zip(dims[:-1], dims[1:])):
        W = torch.tensor(mlp_oracle.bf16(params[off:off + din * dout].reshape(dout, din)), requires_grad=True)
        off += din * dout
        b = torch.tensor(params[off:off + dout].astype(np.float64), requires_grad=True)
        off += dout
        Ws += [W, b]
        hr = h + (torch.from_numpy(mlp_oracle.bf16(h.detach().float().numpy())) - h.detach())
        h = hr @ W.T + b
        if l < len(dims) - 2:
            h = torch.relu(h)
    return h, Ws


def test_mlp_oracle_against_torch_autograd():
    rng = np.random.default_rng(1)
    dims = [8, 16, 16, 4]
    count = sum(a * b + b for a, b in zip(dims[:-1], dims[1:]))
    params = (0.4 * rng.normal(size=count)).astype(np.float32)
    X = rng.normal(size=(33, 8)).astype(np.float32)
    y, cache = mlp_oracle.forward(params, dims, X)
    yt, Ws = _torch_mlp(params, dims, X)
    assert np.abs(y - yt.detach().numpy()).max() <= 1e-5 * np.abs(y).max()
    dY = rng.normal(size=y.shape)
    g = mlp_oracle.backward(params, dims, cache, dY)
    (yt * torch.tensor(dY)).sum().backward()
    gt = np.concatenate([w.grad.numpy().reshape(-1) for w in Ws])
    # the oracle rounds dY and the activations to BF16 inside the products: 1e-2
    assert np.linalg.norm(g - gt) <= 2e-2 * np.linalg.norm(gt)
    # layer by layer, weights and biases separately
    off = 0
    for din, dout in zip(dims[:-1], dims[1:]):
        for size in (din * dout, dout):
            blk = slice(off, off + size)
            assert np.linalg.norm(g[blk] - gt[blk]) <= 3e-2 * np.linalg.norm(gt[blk]) + 1e-12
            off += size


def test_learned_inverse_closed_form(oracle):
    """With zero network output (alpha = 1) on every ordered pair, the learned
    preconditioner is the first-order Neumann inverse I + T C."""
    rng = np.random.default_rng(2)
    N, lmax = 4, 2
    pos = rng.uniform(-1.5, 1.5, size=(N, 3)) + 3.0 * np.arange(N)[:, None]
    cl = oracle.Cluster(lmax, 1.3, rng.uniform(0.2, 0.3, N), 4 + 0.1j, pos)
    pairs = [(i, j) for i in range(N) for j in range(N) if i != j]
    out = np.zeros((len(pairs), 16), np.float32)
    Minv = mlp_oracle.learned_inverse(cl, pairs, out)
    assert np.abs(Minv - (np.eye(cl.n) + cl.T @ cl.C)).max() <= 1e-13 * np.abs(Minv).max()
    # a pair's weight multiplies exactly its block: alpha = 1 + 0.5 tanh(o) + 0.5 i tanh(o')
    out[0, 0] = 0.75  # exact in float32
    M2 = mlp_oracle.learned_inverse(cl, pairs, out)
    i, j = pairs[0]
    D = M2 - Minv
    M = cl.M
    mask = np.zeros_like(D, bool)
    mask[:, j * M:(j + 1) * M] = True
    assert np.abs(D[~mask]).max() == 0.0
    deg = oracle.modes(lmax)[0]
    Cij = cl.C[i * M:(i + 1) * M, j * M:(j + 1) * M] * (deg == 1)[:, None]
    expect = cl.T[:, i * M:(i + 1) * M] @ (0.5 * np.tanh(0.75) * Cij)
    assert np.abs(D[:, j * M:(j + 1) * M] - expect).max() <= 1e-13 * np.abs(expect).max()


def test_coupling_apply_against_dense(oracle):
    rng = np.random.default_rng(3)
    N, lmax = 5, 3
    pos = rng.uniform(-2, 2, size=(N, 3)) + 2.5 * np.arange(N)[:, None]
    C = oracle.pair_matrix(lmax, 1.3, pos, False)
    x = rng.normal(size=(C.shape[0], 3)) + 1j * rng.normal(size=(C.shape[0], 3))
    y = oracle.coupling_apply(lmax, 1.3, pos, x)
    assert np.abs(y - C @ x).max() <= 1e-12 * np.abs(C @ x).max()
    M = 2 * lmax * (lmax + 2)
    yr = oracle.coupling_apply(lmax, 1.3, pos, x, rows=(1, 3))
    assert np.abs(yr[M:3 * M] - (C @ x)[M:3 * M]).max() <= 1e-12 * np.abs(C @ x).max()
