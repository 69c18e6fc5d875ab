"""GPU checks of the learned preconditioner: the MFMA MLP against the NumPy oracle
(same BF16 operand rounding, FP64 accumulation), the preconditioner and its training
loss against the dense oracle, the loss gradient against central finite differences,
and the network's gain over the same preconditioner without it."""
import numpy as np
import pytest
import torch

from __graft_entry__ import load_package

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pkg():
    return load_package()


@pytest.fixture(scope="module")
def learned(pkg):
    import importlib

    return importlib.import_module("treams_rs_amd.learned")


def test_mlp_forward_backward_against_oracle(learned):
    import mlp_oracle

    rng = np.random.default_rng(0)
    mlp = learned.MLP(n_in=64, width=256, n_out=16, hidden_layers=3, seed=5)
    params = mlp.params()
    # give the zero-initialised last layer values so its gradient path is exercised
    params[-(256 * 16 + 16):] = rng.normal(scale=0.05, size=256 * 16 + 16).astype(np.float32)
    mlp.set_params(params)
    X = rng.normal(size=(300, 64)).astype(np.float32)
    dims = [64, 256, 256, 256, 16]
    y = mlp.forward(torch.from_numpy(X)).cpu().numpy()
    y_ref, cache = mlp_oracle.forward(params, dims, X)
    assert np.linalg.norm(y - y_ref) <= 2e-3 * np.linalg.norm(y_ref)
    dY = rng.normal(size=y.shape).astype(np.float32)
    mlp.zero_grad()
    mlp.backward(torch.from_numpy(dY))
    g = mlp.grads()
    g_ref = mlp_oracle.backward(params, dims, cache, dY)
    assert np.linalg.norm(g - g_ref) <= 1e-2 * np.linalg.norm(g_ref)
    # per-layer: every block of the gradient is close, not only the total
    off = 0
    for din, dout in zip(dims[:-1], dims[1:]):
        blk = slice(off, off + din * dout + dout)
        assert np.linalg.norm(g[blk] - g_ref[blk]) <= 2e-2 * np.linalg.norm(g_ref[blk]) + 1e-6
        off += din * dout + dout


def test_adam_moves_parameters(learned):
    mlp = learned.MLP(seed=3)
    p0 = mlp.params()
    x = torch.randn(50, 64)
    y = mlp.forward(x)
    mlp.zero_grad()
    mlp.backward(torch.ones_like(y))
    mlp.adam(lr=1e-2)
    p1 = mlp.params()
    # Adam's first step moves every parameter with a nonzero gradient by ~lr
    g = mlp.grads()
    moved = np.abs(p1 - p0)[np.abs(g) > 1e-6]
    assert moved.size > 0 and np.all(np.abs(moved - 1e-2) <= 2e-3)


def _small_cluster(pkg, oracle, seed=4, N=10, lmax=3):
    radii, eps, pos = pkg.random_cluster(N, 2.8, 0.05, (0.15, 0.35), None, seed=seed)
    return pkg.IterativeSphereCluster(lmax, 1.3, radii, eps, pos), oracle.Cluster(lmax, 1.3, radii, eps, pos)


@pytest.mark.parametrize("k", [4, 8])
def test_learned_apply_and_loss_against_dense_oracle(pkg, oracle, learned, k):
    """k = 8: pair lists built on the device, their column lists (the adjoint's) on
    first use; k = 4: both on the host."""
    import mlp_oracle

    cl, ref = _small_cluster(pkg, oracle)
    npairs = learned.near_pairs(cl, k)
    pairs = learned.pair_list(cl, npairs)
    rng = np.random.default_rng(1)
    out = (0.3 * rng.normal(size=(npairs, 16))).astype(np.float32)
    learned.set_pair_output(cl, torch.from_numpy(out))
    Minv = mlp_oracle.learned_inverse(ref, pairs, out)
    x = rng.normal(size=(ref.n, 2)) + 1j * rng.normal(size=(ref.n, 2))
    y = learned.learned_apply(cl, x).cpu().numpy()
    assert np.linalg.norm(y - Minv @ x) <= 1e-12 * np.linalg.norm(Minv @ x)
    ya = learned.learned_apply(cl, x, adjoint=True).cpu().numpy()
    assert np.linalg.norm(ya - Minv.conj().T @ x) <= 1e-12 * np.linalg.norm(x)
    V = rng.normal(size=(ref.n, 3)) + 1j * rng.normal(size=(ref.n, 3))
    loss, dout = learned.learned_loss(cl, V, npairs)
    assert abs(loss - mlp_oracle.learned_loss(ref, pairs, out, V)) <= 1e-10 * loss
    # gradient: central differences of the oracle loss in a few output coordinates
    dout = dout.cpu().numpy()
    h = 1e-3
    for p, t in [(0, 0), (0, 1), (3, 2), (npairs - 1, 4), (7, 5)]:
        o1, o2 = out.copy(), out.copy()
        o1[p, t] += h
        o2[p, t] -= h
        fd = (mlp_oracle.learned_loss(ref, pairs, o1, V) - mlp_oracle.learned_loss(ref, pairs, o2, V)) / (2 * h)
        assert abs(fd - dout[p, t]) <= 1e-4 * max(1.0, abs(fd)) + 1e-5, (p, t, fd, dout[p, t])
    # outputs beyond lmax carry no gradient
    assert np.all(dout[:, 2 * ref.lmax:] == 0)


@pytest.mark.parametrize("lmax,cols", [(2, 5), (6, 6)])
def test_learned_apply_many_columns(pkg, oracle, learned, lmax, cols):
    """The forward apply reads each pair block once for 4 columns at a time: column
    counts past one chunk, and block rows past one 64-lane slot (lmax = 6, 96 rows),
    against the dense oracle."""
    import mlp_oracle

    cl, ref = _small_cluster(pkg, oracle, seed=11, N=7, lmax=lmax)
    npairs = learned.near_pairs(cl, 5)
    pairs = learned.pair_list(cl, npairs)
    rng = np.random.default_rng(lmax)
    out = (0.3 * rng.normal(size=(npairs, 16))).astype(np.float32)
    learned.set_pair_output(cl, torch.from_numpy(out))
    Minv = mlp_oracle.learned_inverse(ref, pairs, out)
    x = rng.normal(size=(ref.n, cols)) + 1j * rng.normal(size=(ref.n, cols))
    y = learned.learned_apply(cl, x).cpu().numpy()
    assert np.linalg.norm(y - Minv @ x) <= 1e-12 * np.linalg.norm(Minv @ x)


@pytest.mark.parametrize("lmax,cols", [(2, 5), (6, 6)])
def test_learned_apply_inner_path(pkg, oracle, learned, lmax, cols):
    """The mixed-precision inner solves' apply (same-helicity near blocks stored as
    bfloat16 pairs, FP32 products and sums, 8 block loads in flight) against the dense
    oracle: bfloat16 rounds each block value to 2^-9 of itself, and the near-pair
    terms are part of the result, so the result's error stays below 2^-8 of it."""
    import mlp_oracle

    cl, ref = _small_cluster(pkg, oracle, seed=12, N=9, lmax=lmax)
    npairs = learned.near_pairs(cl, 8)
    pairs = learned.pair_list(cl, npairs)
    rng = np.random.default_rng(20 + lmax)
    out = (0.3 * rng.normal(size=(npairs, 16))).astype(np.float32)
    learned.set_pair_output(cl, torch.from_numpy(out))
    Minv = mlp_oracle.learned_inverse(ref, pairs, out)
    x = rng.normal(size=(ref.n, cols)) + 1j * rng.normal(size=(ref.n, cols))
    want = Minv @ x
    y = learned.learned_apply(cl, x, inner=True).cpu().numpy()
    err = np.linalg.norm(y - want) / np.linalg.norm(want)
    print(f"inner-path relative error lmax={lmax}: {err:.3e}", flush=True)
    assert err <= 2.0 ** -8
    # and not merely the c128 path again: the two differ at the rounded blocks' precision
    y128 = learned.learned_apply(cl, x).cpu().numpy()
    assert np.linalg.norm(y128 - want) <= 1e-12 * np.linalg.norm(want)
    assert np.linalg.norm(y - y128) > 1e-9 * np.linalg.norm(want)


def test_learned_preconditioned_solve_same_answer(pkg, oracle, learned):
    cl, ref = _small_cluster(pkg, oracle, seed=6, N=24, lmax=4)
    mlp = learned.MLP(seed=2)
    learned.attach(cl, mlp)
    a = cl.plane_wave([0, 0, 1], 1)
    x0, c0 = cl.solve(a, rtol=1e-10)
    x1, c1 = cl.solve(a, rtol=1e-10, preconditioner="learned")
    assert np.linalg.norm((x1 - x0).cpu().numpy()) <= 1e-8 * np.linalg.norm(x0.cpu().numpy())
    assert c1[0][1] <= 1e-10 * c1[0][2]
    # first-order near-field Neumann (untrained network: alpha = 1) already helps here
    assert c1[0][0] <= c0[0][0]
    # mixed precision with the learned preconditioner certifies the FP64 residual too
    x2, c2 = cl.solve(a, rtol=1e-9, mixed_precision=True, preconditioner="learned", restart=50)
    assert c2[0][1] <= 1e-9 * c2[0][2]
    # inner solves on the split-FP16 matrix-free coupling with the near blocks streamed
    # as c64; the FP64 residual certifies, and the answer is the FP64 one
    x3, c3 = cl.solve(a, rtol=1e-9, mixed_precision="matrix_free", preconditioner="learned", restart=50)
    assert c3[0][1] <= 1e-9 * c3[0][2]
    assert np.linalg.norm((x3 - x0).cpu().numpy()) <= 5e-8 * np.linalg.norm(x0.cpu().numpy())


def test_training_reduces_error_radius(pkg, learned):
    """A short GPU training run (the library's own forward, backward and Adam
    kernels) lowers the estimated spectral radius of A M^{-1} - I on its training
    clusters below that of the same preconditioner without the network (alpha = 1)."""
    log = []
    mlp = learned.train(n_clusters=4, epochs=10, probes=2, lr=1e-3, seed=2000, log=log.append)
    trained, plain = [], []
    for c in range(4):
        lmax, k0, radii, eps, pos = learned.training_cluster(2000 + c)
        cl = pkg.IterativeSphereCluster(lmax, k0, radii, eps, pos)
        npairs = learned.near_pairs(cl)
        out = mlp.forward(learned.pair_features(cl, npairs))
        learned.set_pair_output(cl, out)
        trained.append(learned.error_radius(cl, seed=c))
        learned.set_pair_output(cl, torch.zeros_like(out))
        plain.append(learned.error_radius(cl, seed=c))
    assert np.mean(trained) < np.mean(plain), (trained, plain, log)


def _knn_reference(pos, k):
    """k nearest other spheres by (squared distance, index), the distances rounded
    as ((t0^2 + t1^2) + t2^2) without fused multiply-adds; neighbours listed by index."""
    out = []
    for i in range(len(pos)):
        t = pos[i] - pos
        d = (t[:, 0] * t[:, 0] + t[:, 1] * t[:, 1]) + t[:, 2] * t[:, 2]
        j = np.arange(len(pos))
        keep = j != i
        order = np.lexsort((j[keep], d[keep]))[:k]
        out.extend((i, int(x)) for x in sorted(j[keep][order]))
    return np.array(out, np.int32)


@pytest.mark.parametrize("case", ["random", "lattice"])
@pytest.mark.parametrize("k", [8, 5])
def test_near_pairs_match_brute_force(pkg, learned, case, k):
    """Near-pair lists (k = 8 on the device, other k on the host) equal a brute-force
    selection, including exact distance ties broken by sphere index (cubic lattice)."""
    if case == "random":
        radii, e, pos = pkg.random_cluster(300, 9.0, 0.05, (0.15, 0.35), None, seed=3)
    else:
        g = np.arange(6, dtype=float)
        pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3) * 1.0
        radii = np.full(len(pos), 0.3)
        e = np.full(len(pos), 4 + 0.1j)
    cl = pkg.IterativeSphereCluster(2, 1.3, radii, e, pos)
    npairs = learned.near_pairs(cl, k)
    got = learned.pair_list(cl, npairs)
    assert np.array_equal(got, _knn_reference(np.asarray(pos, float), k))
