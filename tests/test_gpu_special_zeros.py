"""Regular waves at the zeros of j_0 and j_1 (k d = pi, 2 pi, 3 pi, 4.4934...).

A regular sequence j_p scaled by j_0 = sin x / x alone loses its accuracy next to
x = k pi; these cases pin the GPU's regular translations, overlap, cross sections and
Mie blocks there.  The oracle's j_p is pinned at the same points by the mpmath values
of tests/golden/special_functions.json (test_oracle_golden.py), and the Mie blocks
directly by the Bohren-Huffman mpmath rows at size parameters pi and 4.4934.
Tolerances: 1e-12 relative (FP64 paths).
"""
import json
import os

import numpy as np
import pytest
import torch

from __graft_entry__ import load_package

pytestmark = pytest.mark.gpu

ZEROS = [np.pi, 2 * np.pi, 3 * np.pi, 4.493409457909064, 7.725251836937707]


@pytest.fixture(scope="module")
def pkg():
    p = load_package()
    assert torch.cuda.is_available()
    return p


@pytest.mark.parametrize("lmax", [3, 6])
def test_regular_blocks_at_bessel_zeros(pkg, oracle, lmax):
    k = 1.3
    disp = []
    for kd in ZEROS:
        d = kd / k
        disp.append([0.0, 0.0, d])
        disp.append(list(d * np.array([0.48, -0.6, 0.64])))
    got = pkg.translation_blocks(lmax, k, np.array(disp), True).cpu().numpy()
    for i, d in enumerate(disp):
        ref = oracle.translation_block(lmax, k, np.array(d), True)
        assert np.abs(got[i] - ref).max() <= 1e-12 * np.abs(ref).max(), (i, d)


def test_mie_at_bessel_zeros_against_bohren_huffman(pkg, golden_dir):
    rows = json.load(open(os.path.join(golden_dir, "special_functions.json")))["mie_bohren_huffman"]
    for x in (np.pi, 4.493409457909064):
        sel = [r for r in rows if abs(r["x"] - x) < 1e-12]
        assert sel, x
        eps = complex(*sel[0]["eps"])
        # one sphere of radius x / k0 gives size parameter x
        cl = pkg.IterativeSphereCluster(6, 1.0, [x], eps, [[0.0, 0.0, 0.0]])
        blocks = cl.mie_blocks()[0]
        for r in sel:
            a, b = complex(*r["a"]), complex(*r["b"])
            expected = -np.array([[a + b, a - b], [a - b, a + b]]) / 2
            got = blocks[r["n"] - 1]
            assert np.abs(got - expected).max() <= 1e-12 * np.abs(expected).max(), (x, r["n"])


def test_two_spheres_half_wavelength_cross_sections(pkg, oracle):
    k0 = 1.3
    d = np.pi / k0  # k0 d = pi: j_0(k0 d) = 0 in every regular pair translation
    radii = [0.3, 0.25]
    eps = [4 + 0.1j, 2.5 + 0.0j]
    pos = np.array([[0.0, 0.0, -0.5 * d], [0.0, 0.0, 0.5 * d]])
    lmax = 4
    cl = pkg.IterativeSphereCluster(lmax, k0, radii, eps, pos)
    ref = oracle.Cluster(lmax, k0, radii, eps, pos)
    rng = np.random.default_rng(5)
    x = rng.normal(size=(ref.n, 2)) + 1j * rng.normal(size=(ref.n, 2))
    got = cl.overlap(x).cpu().numpy()
    exp = ref.R @ x
    assert np.linalg.norm(got - exp) <= 1e-12 * np.linalg.norm(exp)
    for direction in ([0, 0, 1], [1, 0, 0]):
        a = cl.plane_wave(direction, 1)
        xs, _ = cl.solve(a, rtol=1e-13)
        a_ref = ref.plane_wave(direction, 1)
        x_ref = ref.solve(a_ref)
        sca, ext = cl.cross_sections(a, xs)
        sca_ref, ext_ref = ref.cross_sections(a_ref, x_ref)
        assert abs(sca - sca_ref) <= 1e-11 * abs(sca_ref), direction
        assert abs(ext - ext_ref) <= 1e-11 * abs(ext_ref), direction
