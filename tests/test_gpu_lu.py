"""GPU parity of the dense LU path (assembled I - T C, blocked LU with MFMA trailing
updates, FP64 refinement) against the FP64 oracle (LAPACK solves of the oracle's
dense operator) and the reference goldens.

Tolerances: solutions 1e-10 relative to the oracle for both precisions (the c64
factors are refined with the FP64 residual to rtol 1e-12); certificates are the
FP64 residual itself; gradients 1e-9 relative like the GMRES pullback, and central
finite differences (step 1e-6) to 1e-5 relative.
"""
import json
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


def _cluster(pkg, oracle, N=8, lmax=3, seed=1, side=2.0, eps=4 + 0.1j, k0=1.3):
    radii, e, pos = pkg.random_cluster(N, side, 0.1, (0.2, 0.3), eps, seed=seed)
    return pkg.IterativeSphereCluster(lmax, k0, radii, e, pos), oracle.Cluster(lmax, k0, radii, e, pos)


@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_factor_solve_small_cluster(pkg, oracle, precision):
    """8 spheres, lmax 3 (240 unknowns: three full 64-blocks and a ragged one)."""
    cl, ref = _cluster(pkg, oracle, N=8, lmax=3, seed=5)
    f = pkg.SphereClusterFactor(cl, precision)
    a = torch.stack([cl.plane_wave([0, 0, 1], 1), cl.plane_wave([0, 0, -1], 0)], dim=1)
    x, conv = f.solve(a)
    an = a.cpu().numpy()
    for c in range(2):
        assert _rel(x[:, c], ref.solve(an[:, c])) <= 1e-10
        assert conv[c].residual_norm <= 1e-12 * conv[c].rhs_norm
        r = np.linalg.norm(ref.operator() @ x[:, c].cpu().numpy() - oracle.apply_local(ref.blocks, an[:, c]))
        assert abs(r - conv[c].residual_norm) <= 1e-3 * conv[c].residual_norm + 1e-14 * conv[c].rhs_norm
    if precision == "c128":
        assert max(c.refinements for c in conv) <= 1
    else:
        assert max(c.refinements for c in conv) >= 1
    f.close()


@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_factor_adjoint_solve(pkg, oracle, precision):
    cl, ref = _cluster(pkg, oracle, N=10, lmax=3, seed=7, side=2.6)
    f = pkg.SphereClusterFactor(cl, precision)
    rng = np.random.default_rng(3)
    g = rng.normal(size=(ref.n, 3)) + 1j * rng.normal(size=(ref.n, 3))
    y, conv = f.solve_adjoint(g)
    yr = np.linalg.solve(ref.operator().conj().T, g)
    assert _rel(y, yr) <= 1e-10
    # pullback_incident = T^H A^{-H} G (cluster/interaction.rs:220-226)
    assert _rel(f.pullback_incident(g), oracle.apply_local(ref.blocks, yr, True)) <= 1e-10


@pytest.mark.parametrize("N,lmax,precision", [(4, 2, "c128"), (6, 3, "c64"), (9, 3, "c128"), (16, 3, "c128"),
                                             (16, 3, "c64"), (19, 3, "c128")])
def test_interacting_tmatrix_many_columns(pkg, oracle, N, lmax, precision):
    """All n columns (the interacting T-matrix X = (I - T C)^{-1} T of sphere_cluster,
    cluster/spheres.rs:56-80): forward solves of more than 8 columns go block by block
    with MFMA GEMM updates (c64 factors widened to FP64), refined in FP64; n = 64, 180
    (c64, ragged last block), 270, 480 (c128 and c64, 8 blocks) and 570 (ragged)."""
    cl, ref = _cluster(pkg, oracle, N=N, lmax=lmax, seed=2, side=1.6 * (N / 4) ** (1 / 3))
    f = pkg.SphereClusterFactor(cl, precision)
    eye = np.eye(ref.n, dtype=np.complex128)
    X, conv = f.solve(eye)
    Xr = np.linalg.solve(ref.operator(), ref.T)
    assert _rel(X, Xr) <= 1e-11
    assert len(conv) == ref.n


@pytest.mark.parametrize("N,precision", [(40, "c64"), (33, "c128")])
def test_factor_larger_cluster(pkg, oracle, N, precision):
    """N = 40, lmax = 4 (1,920 unknowns, 30 blocks: 15 block pairs with K = 128
    trailing updates), c64 factors with FP64 refinement; N = 33 (1,584 unknowns, 24
    full blocks and a partial one left unpaired) in c128; 4 right-hand sides."""
    cl, ref = _cluster(pkg, oracle, N=N, lmax=4, seed=21, side=5.0 * (N / 40) ** (1 / 3))
    f = pkg.SphereClusterFactor(cl, precision)
    a = torch.stack([cl.plane_wave(d, p) for d in ([0, 0, 1], [0, 0, -1]) for p in (1, 0)], dim=1)
    x, conv = f.solve(a)
    xr = np.linalg.solve(ref.operator(), oracle.apply_local(ref.blocks, a.cpu().numpy()))
    assert _rel(x, xr) <= 1e-10
    for c in conv:
        assert c.residual_norm <= 1e-12 * c.rhs_norm
        assert (1 <= c.refinements <= 6) if precision == "c64" else c.refinements <= 1
    # the GMRES path gives the same answer
    xg, _ = cl.solve(a, rtol=1e-12)
    assert _rel(x, xg) <= 1e-10


def test_factor_cluster_example_golden(pkg, golden_dir):
    g = json.load(open(os.path.join(golden_dir, "reference_examples.json")))["cluster_example"]
    pos = (220 / np.sqrt(24)) * np.array(
        [[np.sqrt(8), 0, -1], [-np.sqrt(2), np.sqrt(6), -1], [-np.sqrt(2), -np.sqrt(6), -1], [0, 0, 3]])
    k0 = 2 * np.pi / g["k0_over_2pi_inverse"]
    f = pkg.sphere_cluster_factor(g["lmax"], k0, g["radii"], complex(*g["epsilon"]), pos)
    a = f.cluster.plane_wave(g["direction"], g["pol"])
    x, _ = f.solve(a)
    sca, ext = f.cluster.cross_sections(a, x)
    e = g["expected"]
    assert float(f"{sca:.6g}") == e["scattering"]
    assert float(f"{ext:.6g}") == e["extinction"]
    assert float(f"{ext - sca:.6g}") == e["absorption"]


def test_factor_pullback_matches_oracle_and_gmres(pkg, oracle):
    cl, ref = _cluster(pkg, oracle, N=6, lmax=3, seed=13, side=2.5, eps=None)
    f = pkg.SphereClusterFactor(cl, "c128")
    rng = np.random.default_rng(2)
    a = np.stack([ref.plane_wave([0, 0, 1], 1), ref.plane_wave([1, 0, 0], 0)], axis=1)
    x, ctx, _ = f.record(a)
    G = rng.normal(size=a.shape) + 1j * rng.normal(size=a.shape)
    gk, gr, ge, gp, ginc, conv = ctx.pullback(G)
    xr = np.stack([ref.solve(a[:, c]) for c in range(2)], axis=1)
    r = ref.pullback(a, xr, G)
    assert np.abs(gr - r["radii"]).max() <= 1e-9 * np.abs(r["radii"]).max()
    assert np.abs(ge - r["eps"]).max() <= 1e-9 * np.abs(r["eps"]).max()
    assert np.abs(gp - r["positions"]).max() <= 1e-9 * np.abs(r["positions"]).max()
    assert abs(gk - r["k0"]) <= 1e-9 * abs(r["k0"])
    assert _rel(ginc, r["incident"]) <= 1e-9
    # the GMRES pullback of the same record agrees
    _, ctx_g, _ = cl.record(a, rtol=1e-13)
    gk2, gr2, ge2, gp2, ginc2, _ = ctx_g.pullback(G)
    assert np.abs(gp - gp2).max() <= 1e-9 * np.abs(gp).max()
    assert np.abs(gr - gr2).max() <= 1e-9 * np.abs(gr).max()


def test_factor_scattering_gradient_finite_differences(pkg):
    """d sigma_sca / d (radii, positions) through the LU adjoint against central
    differences with step 1e-6 (each a fresh factorisation)."""
    radii, eps, pos = pkg.random_cluster(5, 2.4, 0.1, (0.2, 0.3), None, seed=31)
    k0, lmax = 1.3, 3

    def sigma(r, p):
        f = pkg.sphere_cluster_factor(lmax, k0, r, eps, p)
        a = f.cluster.plane_wave([0, 0, 1], 1)
        x, _ = f.solve(a)
        return f.cluster.cross_sections(a, x, flux=0.5)[0]

    f = pkg.sphere_cluster_factor(lmax, k0, radii, eps, pos)
    sca, gr, gp, _ = pkg.scattering_gradient(f, [0, 0, 1], 1)
    assert abs(sca - sigma(radii, pos)) <= 1e-12 * sca
    h = 1e-6
    for i in (0, 3):
        rp, rm = radii.copy(), radii.copy()
        rp[i] += h
        rm[i] -= h
        fd = (sigma(rp, pos) - sigma(rm, pos)) / (2 * h)
        assert abs(fd - gr[i]) <= 1e-5 * max(abs(fd), 1e-3 * np.abs(gr).max())
    for i, ax in ((1, 0), (2, 2), (4, 1)):
        pp, pm = pos.copy(), pos.copy()
        pp[i, ax] += h
        pm[i, ax] -= h
        fd = (sigma(radii, pp) - sigma(radii, pm)) / (2 * h)
        assert abs(fd - gp[i, ax]) <= 1e-5 * max(abs(fd), 1e-3 * np.abs(gp).max())


def test_factor_invalid_input(pkg, oracle):
    cl, _ = _cluster(pkg, oracle, N=3, lmax=2, seed=1, side=1.5)
    with pytest.raises(KeyError):
        pkg.SphereClusterFactor(cl, "c32")
    f = pkg.SphereClusterFactor(cl, "c128")
    with pytest.raises(ValueError):
        f.solve(np.zeros((cl.n + 1, 1), np.complex128))
    f.close()
    with pytest.raises(ValueError):
        f.solve(np.zeros(cl.n, np.complex128))


def test_global_tmatrix_matches_oracle(pkg, oracle):
    """T-matrix of a 5-sphere cluster in one basis at the origin (regular translations,
    all M columns solved with the factors) against the dense oracle; an operator
    within a factor 10 of its diagonal is not equilibrated."""
    cl, ref = _cluster(pkg, oracle, N=5, lmax=3, seed=12, side=2.4)
    f = pkg.SphereClusterFactor(cl, "c128")
    assert not f.equilibrated
    Tg, conv = f.global_tmatrix()
    assert len(conv) == cl.modes_per_sphere
    assert _rel(Tg, ref.global_tmatrix()) <= 1e-10
    f.close()


def test_global_tmatrix_sic_chain_golden(pkg):
    """The reference's golden (tests/tmatrix/test_particle_cluster.py:538-559): four
    250 nm SiC spheres with 20 nm gaps at 3 THz, lmax 10, whose operator I - T C has
    unit diagonal and entries above 1e22.  The factors are those of the equilibrated
    operator; the absorption of the global T-matrix matches to rtol 2e-11 and distinct
    azimuthal orders do not couple."""
    k0 = 2 * np.pi * 3e12 / 299792458.0
    pos = np.array([[0, 0, z * 520e-9] for z in (-1.5, -0.5, 0.5, 1.5)])
    cl = pkg.IterativeSphereCluster(10, k0, np.full(4, 250e-9), np.full(4, 12.87707323 + 0.19225647000000007j), pos)
    f = pkg.SphereClusterFactor(cl, "c128", rtol=1e-13)
    assert f.equilibrated
    Tg, conv = f.global_tmatrix()
    T = Tg.cpu().numpy()
    absorption = -2 * np.pi * (np.trace(T).real + np.vdot(T, T).real) / k0 ** 2
    assert abs(absorption * 1e12 / 0.0001754433586134972 - 1) <= 2e-11, absorption * 1e12
    m = np.array([mm for l in range(1, 11) for mm in range(-l, l + 1) for _ in (0, 1)])
    assert np.abs(T[m[:, None] != m[None, :]]).max() <= 2e-20
    f.close()
