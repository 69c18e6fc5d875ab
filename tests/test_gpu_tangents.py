"""Forward-mode derivatives (pushforward) on the GPU.

* Translation-block tangents against central differences of the oracle's blocks and,
  exactly, against the oracle's block pullback (<G, J t> = <J^T G, t>).
* Cluster pushforward (GMRES and LU) against central differences of the oracle's
  dense solve, and against the GPU pullback through the adjoint identity
  Re<G, dX> = g_k0 dk0 + g_r . dr + Re(conj(g_eps) . deps) + g_pos . dpos + Re<g_B, dB>.
* The interacting T-matrix's tangent (B = I, all n columns) against central
  differences of the oracle's interaction_tmatrix.

Tolerances: 1e-12 for exact identities between FP64 quantities; central differences
(step 1e-6 relative) agree to 1e-7 relative, their truncation and rounding error.
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


@pytest.mark.parametrize("regular", [False, True])
@pytest.mark.parametrize("lmax", [3, 6])
def test_translation_tangent_against_oracle(pkg, oracle, lmax, regular):
    rng = np.random.default_rng(lmax + 10 * regular)
    disp = rng.normal(size=(4, 3)) * 1.5
    disp[0] = [0.0, 0.0, 0.9]
    tang = rng.normal(size=(4, 3))
    k, dk = 1.3, 0.37
    got = pkg.translation_blocks_tangent(lmax, k, disp, tang, dk, regular).cpu().numpy()
    h = 1e-6
    for i in range(4):
        fd = (oracle.translation_block(lmax, k + h * dk, disp[i] + h * tang[i], regular) -
              oracle.translation_block(lmax, k - h * dk, disp[i] - h * tang[i], regular)) / (2 * h)
        scale = np.abs(fd).max()
        assert np.abs(got[i] - fd).max() <= 1e-7 * scale, i
        # exact: the oracle's block pullback is the transpose of this tangent
        G = rng.normal(size=got[i].shape) + 1j * rng.normal(size=got[i].shape)
        gd, gk = oracle.translation_pullback(lmax, k, disp[i], regular, G)
        lhs = np.real(np.vdot(G, got[i]))
        rhs = float(np.dot(gd, tang[i]) + np.real(gk) * dk)
        assert abs(lhs - rhs) <= 1e-12 * (np.abs(G).sum() * scale), i


def _problem(pkg, N=6, lmax=3, seed=2):
    radii, eps, pos = pkg.random_cluster(N, 3.0, 0.15, (0.2, 0.3), None, seed=seed)
    rng = np.random.default_rng(seed + 100)
    t = dict(k0=0.21, radii=rng.normal(size=N) * 0.1, epsilon=rng.normal(size=N) + 1j * rng.normal(size=N),
             positions=rng.normal(size=(N, 3)))
    return lmax, 1.3, radii, eps, pos, t


def _oracle_fd(oracle, lmax, k0, radii, eps, pos, a, da, t, h=1e-6):
    def solve(s):
        cl = oracle.Cluster(lmax, k0 + s * t["k0"], radii + s * t["radii"], eps + s * t["epsilon"],
                            pos + s * t["positions"])
        return np.linalg.solve(cl.operator(), oracle.apply_local(cl.blocks, a + s * da, False))
    return (solve(h) - solve(-h)) / (2 * h)


def test_cluster_pushforward_against_oracle_and_pullback(pkg, oracle):
    lmax, k0, radii, eps, pos, t = _problem(pkg)
    cl = pkg.IterativeSphereCluster(lmax, k0, radii, eps, pos)
    ref = oracle.Cluster(lmax, k0, radii, eps, pos)
    a = np.stack([ref.plane_wave([0, 0, 1], 1), ref.plane_wave([1, 0, 0], 0)], axis=1)
    rng = np.random.default_rng(7)
    da = rng.normal(size=a.shape) + 1j * rng.normal(size=a.shape)
    x, ctx, conv = cl.record(a, rtol=1e-13)
    dx, dconv = ctx.pushforward(t["k0"], t["radii"], t["epsilon"], t["positions"], da)
    dx = dx.cpu().numpy()
    for it, res, bn in dconv:
        assert res <= 1e-13 * bn
    fd = _oracle_fd(oracle, lmax, k0, radii, eps, pos, a, da, t)
    assert np.linalg.norm(dx - fd) <= 1e-7 * np.linalg.norm(fd)
    # each tangent alone (a term dropped from the right-hand side would fail one of these)
    for key in ("k0", "radii", "epsilon", "positions"):
        one = {k: (v if k == key else 0 * v) for k, v in t.items()}
        d1, _ = ctx.pushforward(one["k0"], one["radii"], one["epsilon"], one["positions"], None)
        f1 = _oracle_fd(oracle, lmax, k0, radii, eps, pos, a, 0 * da, one)
        assert np.linalg.norm(d1.cpu().numpy() - f1) <= 1e-7 * np.linalg.norm(f1), key
    # adjoint identity with the GPU pullback
    G = rng.normal(size=a.shape) + 1j * rng.normal(size=a.shape)
    gk, gr, ge, gp, gb, _ = ctx.pullback(G)
    lhs = np.real(np.vdot(G, dx))
    rhs = (gk * t["k0"] + gr @ t["radii"] + np.real(np.vdot(ge, t["epsilon"])) + np.sum(gp * t["positions"]) +
           np.real(np.vdot(gb.cpu().numpy(), da)))
    assert abs(lhs - rhs) <= 1e-9 * np.linalg.norm(G) * np.linalg.norm(dx)


def test_factor_pushforward_matches_gmres(pkg):
    lmax, k0, radii, eps, pos, t = _problem(pkg, N=7, lmax=4, seed=5)
    cl = pkg.IterativeSphereCluster(lmax, k0, radii, eps, pos)
    a = torch.stack([cl.plane_wave([0, 0, 1], 1), cl.plane_wave([0, 1, 0], 0)], dim=1)
    for prec in ("c128", "c64"):
        f = pkg.SphereClusterFactor(cl, prec, rtol=1e-13)
        x, ctx, _ = f.record(a)
        dxf, conv = ctx.pushforward(t["k0"], t["radii"], t["epsilon"], t["positions"])
        xg, ctxg, _ = cl.record(a, rtol=1e-13)
        dxg, _ = ctxg.pushforward(t["k0"], t["radii"], t["epsilon"], t["positions"])
        f.close()
        assert torch.linalg.norm(dxf - dxg) <= 1e-10 * torch.linalg.norm(dxg), prec
        for c in conv:
            assert c.residual_norm <= 1e-13 * c.rhs_norm


def test_interacting_tmatrix_tangent_against_oracle(pkg, oracle):
    lmax, k0, radii, eps, pos, t = _problem(pkg, N=4, lmax=2, seed=9)
    cl = pkg.IterativeSphereCluster(lmax, k0, radii, eps, pos)
    f = pkg.SphereClusterFactor(cl, "c128", rtol=1e-13)
    eye = torch.eye(cl.n, dtype=torch.complex128, device="cuda")
    X, ctx, _ = f.record(eye)
    dX, _ = ctx.pushforward(t["k0"], t["radii"], t["epsilon"], t["positions"])
    f.close()
    h = 1e-6

    def tm(s):
        return oracle.Cluster(lmax, k0 + s * t["k0"], radii + s * t["radii"], eps + s * t["epsilon"],
                              pos + s * t["positions"]).interaction_tmatrix(equilibrate=False)
    fd = (tm(h) - tm(-h)) / (2 * h)
    assert np.linalg.norm(dX.cpu().numpy() - fd) <= 1e-7 * np.linalg.norm(fd)
