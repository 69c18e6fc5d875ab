"""The LU contract of the reference (linalg/mod.rs:137-438) on the GPU: whole-column
partial pivoting when pivoting inside 64-row blocks meets a zero pivot, balancing of
badly scaled operators, the automatic c64 -> c128 refactorisation of a stalled solve,
and the dense interaction / sphere-cluster / particle-cluster entry points of
python/treams_rs/diff.py built on these factors.

Reference answers: LAPACK (numpy.linalg.solve), the FP64 oracle, and central
differences (step 1e-6).  Tolerances are stated per test.
"""
import numpy as np
import pytest
import torch

from __graft_entry__ import load_package

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pkg():
    p = load_package()
    assert torch.cuda.is_available()
    return p


def _rand(rng, *shape):
    return rng.normal(size=shape) + 1j * rng.normal(size=shape)


def _singular_leading_block(n, seed):
    """A = P W: W = I + 0.1 R well conditioned with W[64:128, 0:64] = 0, and P moving
    rows 64..127 of W to the top, so A[0:64, 0:64] = 0 exactly while A is regular."""
    rng = np.random.default_rng(seed)
    R = _rand(rng, n, n) / np.sqrt(n)
    R[64:128, 0:64] = 0.0
    W = np.eye(n) + 0.1 * R
    perm = np.r_[64:128, 0:64, 128:n]
    return W[perm]


@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_whole_column_pivoting_on_singular_leading_block(pkg, precision):
    n = 200
    A = _singular_leading_block(n, 3)
    assert np.abs(A[:64, :64]).max() == 0.0 and np.linalg.cond(A) < 10
    rng = np.random.default_rng(4)
    b = _rand(rng, n, 3)
    f = pkg.DenseFactor(A, precision, rtol=1e-12)
    assert f.info["pivoting"] == "column"
    x, conv = f.solve(b)
    ref = np.linalg.solve(A, b)
    assert np.linalg.norm(x.cpu().numpy() - ref) <= 1e-11 * np.linalg.norm(ref)
    xa, _ = f.solve(b, adjoint=True)
    refa = np.linalg.solve(A.conj().T, b)
    assert np.linalg.norm(xa.cpu().numpy() - refa) <= 1e-11 * np.linalg.norm(refa)
    for c in conv:
        assert c.residual_norm <= 1e-12 * c.rhs_norm
    f.close()


def test_block_pivoting_kept_when_it_works(pkg):
    rng = np.random.default_rng(5)
    n = 300
    A = np.eye(n) + 0.2 * _rand(rng, n, n) / np.sqrt(n)
    f = pkg.DenseFactor(A, "c128")
    assert f.info == {"precision": "c128", "pivoting": "block", "fallback": False}
    b = _rand(rng, n, 2)
    x, _ = f.solve(b)
    assert np.linalg.norm(x.cpu().numpy() - np.linalg.solve(A, b)) <= 1e-12 * np.linalg.norm(x.cpu().numpy())
    f.close()


def test_c64_stall_refactors_in_c128(pkg):
    # condition number 1e9: c64 factors cannot refine it (cond * 2^-24 >> 1), c128
    # factors reach the FP64 backward error, a residual of about 1e-16 cond |b|
    rng = np.random.default_rng(6)
    n = 160
    U, _ = np.linalg.qr(_rand(rng, n, n))
    V, _ = np.linalg.qr(_rand(rng, n, n))
    s = np.logspace(0, -9, n)
    A = (U * s) @ V.conj().T
    b = _rand(rng, n, 2)
    f = pkg.DenseFactor(A, "c64", rtol=1e-6)
    assert f.info["precision"] == "c64" and not f.info["fallback"]
    x, conv = f.solve(b)
    assert f.info == {"precision": "c128", "pivoting": "column", "fallback": True}
    for c in conv:
        assert c.residual_norm <= 1e-6 * c.rhs_norm
    r = A @ x.cpu().numpy() - b
    assert np.linalg.norm(r) <= 2e-6 * np.linalg.norm(b)
    ref = np.linalg.solve(A, b)
    assert np.linalg.norm(x.cpu().numpy() - ref) <= 1e-5 * np.linalg.norm(ref)
    f.close()


def test_badly_scaled_operator_is_equilibrated_and_balanced(pkg):
    rng = np.random.default_rng(7)
    n = 180
    W = np.eye(n) + 0.3 * _rand(rng, n, n) / np.sqrt(n)
    dr = 2.0 ** rng.integers(-30, 30, n)
    dc = 2.0 ** rng.integers(-30, 30, n)
    A = dr[:, None] * W * dc[None, :]
    b = _rand(rng, n, 2)
    for precision in ("c128", "c64"):
        f = pkg.DenseFactor(A, precision, rtol=1e-12)
        x, conv = f.solve(b)
        ref = np.linalg.solve(A, b)
        # the scaled system W (dc x) = b / dr is well conditioned
        assert np.linalg.norm(dc[:, None] * (x.cpu().numpy() - ref)) <= 1e-10 * np.linalg.norm(dc[:, None] * ref)
        f.close()


def test_interaction_entry_point(pkg):
    rng = np.random.default_rng(8)
    n = 96
    T = 0.3 * _rand(rng, n, n) / np.sqrt(n)
    C = 0.5 * _rand(rng, n, n) / np.sqrt(n)
    X, ctx = pkg.diff.interaction(T, C)
    ref = np.linalg.solve(np.eye(n) - T @ C, T)
    assert np.linalg.norm(X.cpu().numpy() - ref) <= 1e-12 * np.linalg.norm(ref)
    dT, dC, G = _rand(rng, n, n), _rand(rng, n, n), _rand(rng, n, n)
    dX = ctx.pushforward(dT, dC).cpu().numpy()
    h = 1e-6
    fd = (np.linalg.solve(np.eye(n) - (T + h * dT) @ (C + h * dC), T + h * dT) -
          np.linalg.solve(np.eye(n) - (T - h * dT) @ (C - h * dC), T - h * dT)) / (2 * h)
    assert np.linalg.norm(dX - fd) <= 1e-7 * np.linalg.norm(fd)
    tbar, cbar = ctx.pullback(G)
    lhs = np.real(np.vdot(G, dX))
    rhs = np.real(np.vdot(tbar.cpu().numpy(), dT) + np.vdot(cbar.cpu().numpy(), dC))
    assert abs(lhs - rhs) <= 1e-10 * np.linalg.norm(G) * np.linalg.norm(dX)


def test_sphere_cluster_entry_point(pkg, oracle):
    radii, eps, pos = pkg.random_cluster(4, 2.2, 0.1, (0.2, 0.3), None, seed=12)
    lmax, k0 = 2, 1.3
    X, ctx = pkg.diff.sphere_cluster(lmax, k0, radii, eps, pos)
    ref = oracle.Cluster(lmax, k0, radii, eps, pos)
    Xr = ref.interaction_tmatrix(equilibrate=False)
    assert np.linalg.norm(X.cpu().numpy() - Xr) <= 1e-12 * np.linalg.norm(Xr)
    rng = np.random.default_rng(13)
    G = _rand(rng, *Xr.shape)
    gk, gr, ge, gp = ctx.pullback(G)
    t = dict(k0=0.17, radii=rng.normal(size=4) * 0.05, epsilon=_rand(rng, 4), positions=rng.normal(size=(4, 3)))
    dX = ctx.pushforward(t["k0"], t["radii"], t["epsilon"], t["positions"]).cpu().numpy()
    h = 1e-6

    def tm(s):
        return oracle.Cluster(lmax, k0 + s * t["k0"], radii + s * t["radii"], eps + s * t["epsilon"],
                              pos + s * t["positions"]).interaction_tmatrix(equilibrate=False)
    fd = (tm(h) - tm(-h)) / (2 * h)
    assert np.linalg.norm(dX - fd) <= 1e-7 * np.linalg.norm(fd)
    lhs = np.real(np.vdot(G, dX))
    rhs = gk * t["k0"] + gr @ t["radii"] + np.real(np.vdot(ge, t["epsilon"])) + np.sum(gp * t["positions"])
    assert abs(lhs - rhs) <= 1e-9 * np.linalg.norm(G) * np.linalg.norm(dX)


def test_particle_cluster_derivatives_against_oracle(pkg, oracle):
    """Particles with random local T-matrices (lmax 2, 2 and a lmax-1 particle): the
    pushforward against central differences of the oracle's dense X = (I - T C)^{-1} T,
    and the pullback to the local blocks, positions and k0 by the adjoint identity."""
    rng = np.random.default_rng(21)
    pos = np.array([[0.0, 0.0, 0.0], [1.4, 0.3, -0.2], [-0.5, 1.2, 0.6]])
    lm = [2, 2, 1]
    sizes = [2 * l * (l + 2) for l in lm]
    local = [0.2 * _rand(rng, s, s) / np.sqrt(s) for s in sizes]
    k0 = 1.1
    X, ctx = pkg.diff.particle_cluster(local, pos, k0, lmax=lm)

    def dense(loc, p, k):
        L, M = 2, 16
        full = np.zeros((3, M, M), complex)
        keep = []
        for i, (t, l) in enumerate(zip(loc, lm)):
            m = 2 * l * (l + 2)
            full[i][:m, :m] = t  # the reference mode order puts degree <= l first
            keep.extend(range(i * M, i * M + m))
        Tm = np.zeros((3 * M, 3 * M), complex)
        for i in range(3):
            Tm[i * M:(i + 1) * M, i * M:(i + 1) * M] = full[i]
        Cm = oracle.pair_matrix(L, k, p, False)
        Xf = np.linalg.solve(np.eye(3 * M) - Tm @ Cm, Tm)
        return Xf[np.ix_(keep, keep)]

    Xr = dense(local, pos, k0)
    assert np.linalg.norm(X.cpu().numpy() - Xr) <= 1e-12 * np.linalg.norm(Xr)
    dl = [_rand(rng, s, s) * 0.1 for s in sizes]
    dp = rng.normal(size=(3, 3))
    dk = 0.13
    dX = ctx.pushforward(dl, dp, dk).cpu().numpy()
    h = 1e-6
    fd = (dense([a + h * b for a, b in zip(local, dl)], pos + h * dp, k0 + h * dk) -
          dense([a - h * b for a, b in zip(local, dl)], pos - h * dp, k0 - h * dk)) / (2 * h)
    assert np.linalg.norm(dX - fd) <= 1e-7 * np.linalg.norm(fd)
    G = _rand(rng, *Xr.shape)
    gl, gp, gk = ctx.pullback(G)
    lhs = np.real(np.vdot(G, dX))
    rhs = (sum(np.real(np.vdot(g.cpu().numpy(), d)) for g, d in zip(gl, dl)) + np.sum(gp * dp) + gk * dk)
    assert abs(lhs - rhs) <= 1e-9 * np.linalg.norm(G) * np.linalg.norm(dX)
