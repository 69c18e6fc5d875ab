"""The torch.autograd adapter: gradients of cross sections through the device solve
(LU factors reused by the backward, or adjoint GMRES) against central finite
differences (step 1e-6) and against the library's own cross-section pullback."""
import importlib

import numpy as np
import pytest
import torch

from __graft_entry__ import load_package

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ag():
    load_package()
    return importlib.import_module("treams_rs_amd.autograd")


@pytest.fixture(scope="module")
def pkg():
    return load_package()


def _inputs(pkg, seed=31, N=5):
    radii, eps, pos = pkg.random_cluster(N, 2.4, 0.1, (0.2, 0.3), None, seed=seed)
    return radii, eps, pos


def _sigma(ag, lmax, k0, radii, eps, pos, method="lu", which=0):
    c = ag.TorchCluster(lmax, k0, torch.tensor(radii), torch.tensor(eps), torch.tensor(pos))
    a = c.plane_wave([0, 0, 1], 1)
    x = c.solve(a, method=method)
    return float(c.cross_sections(a, x)[which])


@pytest.mark.parametrize("method", ["lu", "gmres"])
def test_cross_section_gradients_finite_differences(pkg, ag, method):
    radii, eps, pos = _inputs(pkg)
    lmax, k0 = 3, 1.3
    r = torch.tensor(radii, requires_grad=True)
    p = torch.tensor(pos, requires_grad=True)
    e = torch.tensor(eps, requires_grad=True)
    c = ag.TorchCluster(lmax, k0, r, e, p)
    a = c.plane_wave([0, 0, 1], 1)
    opts = {} if method == "lu" else dict(rtol=1e-13)
    x = c.solve(a, method=method, **opts)
    sca, ext = c.cross_sections(a, x)
    (sca + 0.5 * ext).backward()
    h = 1e-6

    def f(rr, ee, pp):
        return _sigma(ag, lmax, k0, rr, ee, pp, method) + 0.5 * _sigma(ag, lmax, k0, rr, ee, pp, method, 1)

    scale_r, scale_p = r.grad.abs().max().item(), p.grad.abs().max().item()
    for i in (0, 2):
        rp, rm = radii.copy(), radii.copy()
        rp[i] += h
        rm[i] -= h
        fd = (f(rp, eps, pos) - f(rm, eps, pos)) / (2 * h)
        assert abs(fd - r.grad[i].item()) <= 1e-5 * max(abs(fd), 1e-2 * scale_r)
    for i, axis in ((1, 0), (3, 2)):
        pp, pm = pos.copy(), pos.copy()
        pp[i, axis] += h
        pm[i, axis] -= h
        fd = (f(radii, eps, pp) - f(radii, eps, pm)) / (2 * h)
        assert abs(fd - p.grad[i, axis].item()) <= 1e-5 * max(abs(fd), 1e-2 * scale_p)
    for i in (1, 4):
        for dz in (h, 1j * h):
            ep, em = eps.copy(), eps.copy()
            ep[i] += dz
            em[i] -= dz
            fd = (f(radii, ep, pos) - f(radii, em, pos)) / (2 * h)
            g = e.grad[i].real.item() if dz == h else e.grad[i].imag.item()
            assert abs(fd - g) <= 1e-5 * max(abs(fd), 1e-2 * e.grad.abs().max().item())


def test_adapter_matches_library_pullback(pkg, ag):
    """sigma_sca gradients through autograd = scattering_gradient (the library's
    composed pullback, checked against the reference's golden gradient table)."""
    radii, eps, pos = _inputs(pkg, seed=7, N=6)
    lmax, k0 = 3, 1.3
    r = torch.tensor(radii, requires_grad=True)
    p = torch.tensor(pos, requires_grad=True)
    c = ag.TorchCluster(lmax, k0, r, torch.tensor(eps), p)
    a = c.plane_wave([0, 0, 1], 1)
    x = c.solve(a, method="lu")
    sca, _ = c.cross_sections(a, x)
    sca.backward(retain_graph=True)
    f = pkg.sphere_cluster_factor(lmax, k0, radii, eps, pos)
    s2, gr, gp, _ = pkg.scattering_gradient(f, [0, 0, 1], 1)
    assert abs(sca.item() - s2) <= 1e-12 * s2
    assert np.abs(r.grad.numpy() - gr).max() <= 1e-10 * np.abs(gr).max()
    assert np.abs(p.grad.numpy() - gp).max() <= 1e-10 * np.abs(gp).max()
    # a second backward reuses the saved factors and accumulates the same gradient
    g1 = r.grad.clone()
    sca.backward()
    assert torch.allclose(r.grad, 2 * g1, rtol=1e-12, atol=0)


@pytest.mark.parametrize("method", ["lu", "gmres"])
def test_forward_mode_matches_reverse_mode(pkg, ag, method):
    """torch forward mode (the record's pushforward) against reverse mode (its
    pullback): Re<G, J t> = <J^T G, t> through the plane wave and the solve."""
    import torch.autograd.forward_ad as fwAD

    radii, eps, pos = _inputs(pkg, seed=11, N=5)
    lmax, k0 = 3, 1.3
    rng = np.random.default_rng(0)
    tr = torch.tensor(rng.normal(size=5) * 0.05)
    tp = torch.tensor(rng.normal(size=(5, 3)))
    opts = {"rtol": 1e-13}
    with fwAD.dual_level():
        r = fwAD.make_dual(torch.tensor(radii), tr)
        p = fwAD.make_dual(torch.tensor(pos), tp)
        c = ag.TorchCluster(lmax, k0, r, torch.tensor(eps), p)
        a = c.plane_wave([0, 0, 1], 1)
        x = c.solve(a, method=method, **opts)
        dx = fwAD.unpack_dual(x).tangent
    assert dx is not None
    G = torch.tensor(rng.normal(size=dx.shape[0]) + 1j * rng.normal(size=dx.shape[0]), device=dx.device)
    r2 = torch.tensor(radii, requires_grad=True)
    p2 = torch.tensor(pos, requires_grad=True)
    c2 = ag.TorchCluster(lmax, k0, r2, torch.tensor(eps), p2)
    x2 = c2.solve(c2.plane_wave([0, 0, 1], 1), method=method, **opts)
    torch.real(torch.vdot(G, x2)).backward()
    lhs = float(torch.real(torch.vdot(G, dx)))
    rhs = float(r2.grad @ tr + torch.sum(p2.grad * tp))
    assert abs(lhs - rhs) <= 1e-9 * float(torch.linalg.norm(G) * torch.linalg.norm(dx))
