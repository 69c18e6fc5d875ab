"""CPU-side checks of the product library (no GPU calls): it loads, exports every
symbol of include/treams_gpu.h, and its host-built translation plan equals the
oracle's plan term by term."""
import os
import re

import numpy as np
import pytest

from __graft_entry__ import ROOT, load_package


@pytest.fixture(scope="module")
def pkg():
    return load_package()


def test_header_symbols_exported(pkg):
    header = open(os.path.join(ROOT, "include", "treams_gpu.h")).read()
    declared = set(re.findall(r"\b(tg_[a-z0-9_]+)\s*\(", header))
    lib = pkg.lib()
    for name in declared:
        assert hasattr(lib, name), name
    assert declared == set(pkg._lib.EXPORTED)


def test_wigner3j_against_sympy(pkg, golden_dir):
    import json

    rows = json.load(open(os.path.join(golden_dir, "special_functions.json")))["wigner3j"]
    for j1, j2, j3, m1, m2, m3, expected in rows:
        actual = pkg.wigner3j(int(j1), int(j2), int(j3), int(m1), int(m2), int(m3))
        assert abs(actual - expected) <= 1e-13 * max(abs(expected), 1e-3)


def test_rotation_matrix_closed_forms_and_group(pkg):
    """Rotation matrices of z-y-z Euler angles (host): the degree-1 small-d elements
    in closed form (d_11 = (1 + cos t) / 2, d_10 = -sin t / sqrt 2, d_00 = cos t,
    d_1-1 = (1 - cos t) / 2), e^(-i m phi) on the diagonal for rotations about z, the
    identity at zero angles, unitarity, zero coupling between degrees and
    polarisations, and the group law for successive rotations about y."""
    t = 0.7
    R = pkg.rotation_matrix(1, 0.0, t, 0.0)

    def idx(l, m, pol):
        return 2 * (l * l + l + m - 1) + (1 - pol)

    for pol in (0, 1):
        d = lambda mp, m: R[idx(1, mp, pol), idx(1, m, pol)]
        assert abs(d(1, 1) - (1 + np.cos(t)) / 2) <= 1e-15
        assert abs(d(1, 0) + np.sin(t) / np.sqrt(2)) <= 1e-15
        assert abs(d(0, 1) - np.sin(t) / np.sqrt(2)) <= 1e-15
        assert abs(d(0, 0) - np.cos(t)) <= 1e-15
        assert abs(d(1, -1) - (1 - np.cos(t)) / 2) <= 1e-15
    lmax, phi = 6, 0.4
    Rz = pkg.rotation_matrix(lmax, phi)
    for l in range(1, lmax + 1):
        for m in range(-l, l + 1):
            for pol in (0, 1):
                assert abs(Rz[idx(l, m, pol), idx(l, m, pol)] - np.exp(-1j * m * phi)) <= 1e-14
    assert np.abs(Rz - np.diag(np.diag(Rz))).max() == 0.0
    assert np.abs(pkg.rotation_matrix(lmax, 0.0, 0.0, 0.0) - np.eye(Rz.shape[0])).max() <= 1e-15
    Ra = pkg.rotation_matrix(lmax, 0.3, 1.1, -0.8)
    assert np.abs(Ra @ Ra.conj().T - np.eye(Ra.shape[0])).max() <= 1e-13
    pols = np.array([1 - (k % 2) for k in range(Ra.shape[0])])
    degs = np.array([int(np.sqrt(k // 2 + 1)) for k in range(Ra.shape[0])])
    mask = (pols[:, None] != pols[None, :]) | (degs[:, None] != degs[None, :])
    assert np.abs(Ra[mask]).max() == 0.0
    a, b = 0.5, 0.9
    assert np.abs(pkg.rotation_matrix(lmax, 0, a, 0) @ pkg.rotation_matrix(lmax, 0, b, 0)
                  - pkg.rotation_matrix(lmax, 0, a + b, 0)).max() <= 1e-13


@pytest.mark.parametrize("lmax", [1, 3, 6])
def test_plan_matches_oracle(pkg, oracle, lmax):
    starts, nsame, w, hidx = pkg.plan_terms(lmax)
    H = lmax * (lmax + 2)
    M = 2 * H
    o_starts, o_index, o_weight = oracle.plan(lmax)
    l, m, pol = oracle.modes(lmax)
    mode = {(int(a), int(b), int(c)): i for i, (a, b, c) in enumerate(zip(l, m, pol))}
    lm = [(ll, mm) for ll in range(1, lmax + 1) for mm in range(-ll, ll + 1)]
    for a in range(H):
        for b in range(H):
            e = a * H + b
            terms = list(zip(hidx[starts[e]:starts[e + 1]], w[starts[e]:starts[e + 1]]))
            ns = int(nsame[e])
            for p_ in (1, 0):
                sign = 2 * p_ - 1
                mine = {}
                for t, (h, wt) in enumerate(terms):
                    val = wt if t < ns else sign * 1j * wt
                    mine[int(h)] = mine.get(int(h), 0) + val
                s = mode[(lm[a][0], lm[a][1], p_)]
                d = mode[(lm[b][0], lm[b][1], p_)]
                oe = s * M + d
                ref = {}
                for t in range(o_starts[oe], o_starts[oe + 1]):
                    ref[int(o_index[t])] = ref.get(int(o_index[t]), 0) + o_weight[t]
                keys = set(mine) | set(ref)
                # the oracle's log-gamma factorials carry ~1e-15 relative rounding and
                # leave accidental zeros at ~1e-16 of the entry's largest weight
                scale = max([abs(v) for v in ref.values()] + [1.0])
                for k in keys:
                    diff = abs(mine.get(k, 0) - ref.get(k, 0))
                    assert diff <= 5e-14 * scale, (a, b, p_, k)
