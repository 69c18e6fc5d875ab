"""Particle clusters (arbitrary local T-matrices, heterogeneous bases) against the
FP64 oracle's dense systems and the reference's invariances
(cluster/particles.rs:65-123, python/treams_rs/diff.py:1473-1523,
tests/tmatrix/test_particle_cluster.py:133-143, 343-366)."""
import os

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


def _rel(a, b):
    a = a.detach().cpu().numpy() if torch.is_tensor(a) else np.asarray(a)
    b = b.detach().cpu().numpy() if torch.is_tensor(b) else np.asarray(b)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _mie_full(cl):
    """The sphere cluster's Mie blocks as full (N, M, M) T-matrices (reference order)."""
    blocks = cl.mie_blocks()
    N, lmax, M = cl.N, cl.lmax, cl.modes_per_sphere
    full = np.zeros((N, M, M), np.complex128)
    for a in range(M // 2):
        l = int(np.sqrt(a + 1 + 0.5))
        for h in range(2):
            for hp in range(2):
                full[:, 2 * a + h, 2 * a + hp] = blocks[:, l - 1, h, hp]
    return full


def _random_local(rng, n, scale=0.15):
    t = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    return scale * t / np.sqrt(n)


def test_particle_cluster_of_mie_blocks_equals_sphere_cluster(pkg):
    radii, e, pos = pkg.random_cluster(14, 3.6, 0.1, (0.2, 0.3), None, seed=5)
    sph = pkg.IterativeSphereCluster(3, 1.3, radii, e, pos)
    par = pkg.IterativeSphereCluster.from_tmatrices(3, 1.3, pos, _mie_full(sph))
    rng = np.random.default_rng(2)
    x = rng.normal(size=(sph.n, 3)) + 1j * rng.normal(size=(sph.n, 3))
    for adjoint in (False, True):
        assert _rel(par.apply(x, adjoint=adjoint), sph.apply(x, adjoint=adjoint)) <= 1e-13
        assert _rel(par.local(x, adjoint=adjoint), sph.local(x, adjoint=adjoint)) <= 1e-13
    a = sph.plane_wave([0, 0, 1], 1)
    xs, _ = sph.solve(a, rtol=1e-11)
    xp, cp = par.solve(a, rtol=1e-11, convergence="full")
    assert cp[0].residual_norm <= 1e-11 * cp[0].rhs_norm
    assert _rel(xp, xs) <= 1e-9
    for prec in ("c128", "c64"):
        fs, fp = pkg.SphereClusterFactor(sph, prec), pkg.SphereClusterFactor(par, prec)
        eye = torch.eye(sph.n, dtype=torch.complex128, device="cuda")[:, :24]
        assert _rel(fp.solve(eye)[0], fs.solve(eye)[0]) <= 1e-11
        fs.close()
        fp.close()


def test_particle_cluster_random_tmatrices_against_oracle(pkg, oracle):
    rng = np.random.default_rng(11)
    radii, e, pos = pkg.random_cluster(6, 2.6, 0.1, (0.2, 0.3), 4 + 0.1j, seed=11)
    lmax = 2
    M = 2 * lmax * (lmax + 2)
    local = [_random_local(rng, M) for _ in range(len(pos))]
    C = oracle.Cluster(lmax, 1.1, radii, e, pos).C
    T = np.zeros_like(C)
    for i, t in enumerate(local):
        T[i * M:(i + 1) * M, i * M:(i + 1) * M] = t
    expected = np.linalg.solve(np.eye(len(C)) - T @ C, T)
    got = pkg.particle_cluster(local, pos, 1.1, lmax=lmax)
    assert _rel(got, expected) <= 1e-10
    got64 = pkg.particle_cluster(local, pos, 1.1, lmax=lmax, precision="c64")
    assert _rel(got64, expected) <= 1e-10


def test_particle_cluster_heterogeneous_bases_against_reduced_oracle(pkg, oracle):
    """lmax 1, 2 and 3 and a mode subset: the embedded system's columns equal the
    reduced system's (oracle, reduced rows and columns of the full coupling)."""
    rng = np.random.default_rng(3)
    pos = np.array([[0.0, 0.0, 0.0], [1.3, 0.2, -0.1], [0.1, 1.4, 0.5], [-0.9, -0.8, 1.1]])
    subset = [(1, 0, 1), (1, 0, 0), (2, 1, 1), (3, -2, 0), (3, 3, 1)]
    modes = [
        [(l, m, p) for l in range(1, 2) for m in range(-l, l + 1) for p in (1, 0)],
        [(l, m, p) for l in range(1, 3) for m in range(-l, l + 1) for p in (1, 0)],
        [(l, m, p) for l in range(1, 4) for m in range(-l, l + 1) for p in (1, 0)],
        subset,
    ]
    local = [_random_local(rng, len(ms)) for ms in modes]
    got = pkg.particle_cluster(local, pos, 1.2, modes=modes)
    L, Mf = 3, 30
    C = oracle.Cluster(L, 1.2, np.full(4, 0.2), np.full(4, 2.0 + 0j), pos).C
    keep = np.concatenate([[i * Mf + pkg.mode_index(*md) for md in ms] for i, ms in enumerate(modes)])
    Cr = C[np.ix_(keep, keep)]
    T = np.zeros_like(Cr)
    o = 0
    for t in local:
        T[o:o + len(t), o:o + len(t)] = t
        o += len(t)
    expected = np.linalg.solve(np.eye(len(Cr)) - T @ Cr, T)
    assert got.shape == expected.shape
    assert _rel(got, expected) <= 1e-10
    # a lone particle does not couple: X = T
    single = pkg.particle_cluster([local[1]], pos[:1], 1.2, modes=[modes[1]])
    assert _rel(single, local[1]) <= 1e-13


def test_particle_cluster_rigid_translation_and_scale(pkg):
    """tests/tmatrix/test_particle_cluster.py:133-143: positions s p + shift with
    wavenumber k / s give the same coupled T-matrix."""
    rng = np.random.default_rng(8)
    pos = np.array([[0.0, 0.0, 0.0], [1.1, -0.3, 0.4], [0.2, 0.9, -0.7]])
    local = [_random_local(rng, 16) for _ in range(3)]
    base = pkg.particle_cluster(local, pos, 1.3, lmax=2)
    for s, shift in ((0.7, (0.3, -1.2, 2.0)), (2.5, (-3.0, 0.5, 1.0))):
        moved = pkg.particle_cluster(local, pos * s + np.array(shift), 1.3 / s, lmax=2)
        assert _rel(moved, base) <= 1e-12


def test_particle_cluster_rejects_sphere_only_operations(pkg):
    pos = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]])
    t = np.zeros((2, 16, 16), np.complex128)
    par = pkg.IterativeSphereCluster.from_tmatrices(2, 1.0, pos, t)
    with pytest.raises(ValueError, match="sphere"):
        par.mie_blocks()
    a = par.plane_wave([0, 0, 1], 1)
    with pytest.raises(ValueError, match="sphere"):
        par.solve(a, preconditioner="block_jacobi")
    with pytest.raises(ValueError, match="distinct"):
        pkg.IterativeSphereCluster.from_tmatrices(2, 1.0, np.zeros((2, 3)), t)


@pytest.mark.parametrize("N,side,scale", [(12, 5.0, 1.0), (24, 8.0, 0.5)])
@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_particle_cluster_strong_coupling_exchanges_rows(pkg, oracle, N, side, scale, precision):
    """T-matrices of norm ~0.5-1 on close particles: A = I - T C has off-diagonal
    entries up to ~100-150 against its unit diagonal, so the factorisation exchanges
    rows in many pivot steps of the 64-row diagonal blocks (121 of 192 steps at N = 12,
    61 of 384 at N = 24, in a partial-pivoting LU of the reference-order blocks) and
    equilibrates.  n = 192 is a block pair and a single block, whose exchanges reach
    the pair's columns at the end; n = 384 is three pairs with look-ahead.  Checked
    with the oracle's operator: the residual |A X - T| and the distance to its dense
    solve (cond(A) ~ 3e4 bounds the latter)."""
    rng = np.random.default_rng(5)
    radii, e, pos = pkg.random_cluster(N, side, 0.1, (0.2, 0.3), 4 + 0.1j, seed=7)
    lmax = 2
    M = 2 * lmax * (lmax + 2)
    local = [_random_local(rng, M, scale=scale) for _ in range(N)]
    C = oracle.Cluster(lmax, 1.1, radii, e, pos).C
    T = np.zeros_like(C)
    for i, t in enumerate(local):
        T[i * M:(i + 1) * M, i * M:(i + 1) * M] = t
    A = np.eye(len(C)) - T @ C
    expected = np.linalg.solve(A, T)
    got = pkg.particle_cluster(local, pos, 1.1, lmax=lmax, precision=precision)
    X = got.detach().cpu().numpy() if torch.is_tensor(got) else np.asarray(got)
    resid = np.linalg.norm(A @ X - T) / np.linalg.norm(T)
    err = _rel(X, expected)
    print(f"N={N} {precision}: residual {resid:.2e}, distance to the dense solve {err:.2e}")
    assert resid <= 1e-11  # observed 4e-14 - 1.3e-13
    assert err <= 1e-10    # observed 8e-14 - 2.9e-13


def test_particle_cluster_parity_basis_matches_mie_spheres(pkg, golden_dir):
    """Spheres given in the parity basis as diag(-a_l (N, pol 1), -b_l (M, pol 0)) with
    the Bohren-Huffman a_l, b_l of the mpmath fixtures (treams' T = -a (N), -b (M)) give
    the sphere cluster's interacting T-matrix in the parity basis: pins the helicity /
    parity convention (H = (N +- M) / sqrt(2)) through the Mie physics."""
    import json

    rows = json.load(open(os.path.join(golden_dir, "special_functions.json")))["mie_bohren_huffman"]
    cases = [(0.3, complex(4, 0.1)), (0.455, complex(6, 0.2)), (1.1, complex(2.5, 0))]
    lmax = 2
    modes = [(l, m, p) for l in range(1, lmax + 1) for m in range(-l, l + 1) for p in (1, 0)]
    local = []
    for x, eps in cases:
        ab = {r["n"]: (complex(*r["a"]), complex(*r["b"])) for r in rows
              if abs(r["x"] - x) < 1e-12 and complex(*r["eps"]) == eps}
        local.append(np.diag([-(ab[l][0] if p == 1 else ab[l][1]) for (l, m, p) in modes]))
    pos = np.array([[0.0, 0.0, 0.0], [3.0, 0.5, 0.0], [0.4, -0.3, 3.2]])
    Xp, ctx = pkg.diff.particle_cluster(local, pos, 1.0, modes=[modes] * 3, poltype="parity")
    Xh, _ = pkg.diff.sphere_cluster(lmax, 1.0, [c[0] for c in cases], [c[1] for c in cases], pos)
    U = pkg.diff.parity_change(modes)
    big = np.kron(np.eye(3), U)
    expected = big @ Xh.cpu().numpy() @ big
    assert np.linalg.norm(Xp.cpu().numpy() - expected) <= 1e-11 * np.linalg.norm(expected)
    # derivatives in the parity basis: adjoint identity of pushforward and pullback
    rng = np.random.default_rng(2)
    dl = [rng.normal(size=(16, 16)) + 1j * rng.normal(size=(16, 16)) for _ in range(3)]
    dp = rng.normal(size=(3, 3))
    dX = ctx.pushforward(dl, dp, 0.07).cpu().numpy()
    G = rng.normal(size=dX.shape) + 1j * rng.normal(size=dX.shape)
    gl, gp, gk = ctx.pullback(G)
    lhs = np.real(np.vdot(G, dX))
    rhs = sum(np.real(np.vdot(g.cpu().numpy(), d)) for g, d in zip(gl, dl)) + np.sum(gp * dp) + gk * 0.07
    assert abs(lhs - rhs) <= 1e-9 * np.linalg.norm(G) * np.linalg.norm(dX)
