"""Pins of the FP64 oracle (CPU only).

Every oracle function is checked here against something it does not compute
itself: the reference's committed data (spherical_hankel.txt, formal/golden
degrees/harmonics tables, the metallic Mie values, the two documented examples),
mpmath/sympy closed forms (tests/golden/special_functions.json), closed dipole
forms, and exact symmetries.

| oracle function                 | pinned by |
|---------------------------------|-----------|
| or_hankels                      | spherical_hankel.txt (90 digits), mpmath h_p(x) for x in [0.4, 40] |
| or_spherical_jn / series        | mpmath j_p(z), complex z incl. |z| < 0.5 |
| or_wigner3j (w3j_row)           | sympy exact 3j, j <= 12 |
| or_degrees                      | formal/golden/degrees.txt (2450 rows) |
| or_solid_table                  | harmonics.txt index layout; scipy sph_harm closed form |
| or_plan / or_translation_block  | closed dipole form (tests/waves/test_translation_coefficients.py:63-95), origin identity, parity, reciprocity, azimuth |
| or_translation_pullback         | central finite differences; Euler identity |
| or_mie / or_mie_pullback        | metallic golden (mie.rs:574-593), Bohren-Huffman mpmath, FD |
| or_plane_wave                   | cluster example golden (cross sections), single-sphere optical theorem |
| Cluster.solve / cross_sections  | docs/examples/output/cluster.txt |
| Cluster.scattering_gradient     | docs/examples/output/reverse_cluster_gradient.txt |
| gmres                           | dense LAPACK solve |
| gmres_batch / MatrixFreeCluster | dense LAPACK solve (4 columns), dense R and cross sections |
| or_pair_apply (regular)         | dense pair matrix R (or_pair_matrix), identity self blocks |
"""
import json
import math
import os

import numpy as np
import pytest


def _load(golden_dir, name):
    with open(os.path.join(golden_dir, name)) as f:
        return json.load(f)


def _table(path):
    rows = []
    for line in open(path):
        line = line.strip()
        if not line or line.startswith("#"):
            continue
        key, _, values = line.partition(":")
        rows.append((key.split(), values.split()))
    return rows


def test_hankel_against_reference_table(oracle, golden_dir):
    for key, value in _table(os.path.join(golden_dir, "spherical_hankel.txt")):
        p, x = int(key[0]), float(key[1])
        expected = complex(float(value[0]), float(value[1]))
        actual = oracle.hankels(p, x)[p]
        assert abs(actual - expected) <= 1e-13 * abs(expected), (p, x)


def test_hankel_against_mpmath(oracle, golden_dir):
    d = _load(golden_dir, "special_functions.json")["hankel"]
    for x, row in zip(d["x"], d["values"]):
        actual = oracle.hankels(d["pmax"], x)
        expected = np.array([complex(*v) for v in row])
        assert np.all(np.abs(actual - expected) <= 1e-13 * np.abs(expected)), x


def test_spherical_jn_against_mpmath(oracle, golden_dir):
    d = _load(golden_dir, "special_functions.json")["jn"]
    for z, row in zip(d["z"], d["values"]):
        actual = oracle.spherical_jn(d["pmax"], complex(*z))
        expected = np.array([complex(*v) for v in row])
        # relative 1e-13, except for a value at a zero of j_0 or j_1 (z = pi, 4.4934...),
        # which FP64 resolves only to the rounding of the sequence's scale
        scale = np.abs(expected[:2]).max()
        assert np.all(np.abs(actual - expected) <= 1e-13 * np.abs(expected) + 1e-15 * scale), z


def test_wigner3j_against_sympy(oracle, golden_dir):
    rows = _load(golden_dir, "special_functions.json")["wigner3j"]
    for j1, j2, j3, m1, m2, m3, expected in rows:
        actual = oracle.wigner3j(int(j1), int(j2), int(j3), int(m1), int(m2), int(m3))
        assert abs(actual - expected) <= 1e-13 * max(abs(expected), 1e-3), (j1, j2, j3, m1, m2, m3)


def test_degrees_against_lean_golden(oracle, golden_dir):
    rows = _table(os.path.join(golden_dir, "degrees.txt"))
    assert len(rows) == 2450
    for key, values in rows:
        l, m, lam, mu, cross = map(int, key)
        assert oracle.degrees(l, m, lam, mu, cross) == [int(v) for v in values], key


def test_harmonic_table_layout(oracle, golden_dir):
    for key, values in _table(os.path.join(golden_dir, "harmonics.txt")):
        lmax = int(key[0])
        pairs = [tuple(map(int, v.split(","))) for v in values]
        for index, (p, m) in enumerate(pairs):
            assert p * p + p + m == index
        assert len(pairs) == (lmax + 1) ** 2


def test_solid_table_against_associated_legendre(oracle):
    from scipy.special import lpmv

    theta, phi = 0.7, -1.1
    unit = [math.sin(theta) * math.cos(phi), math.sin(theta) * math.sin(phi), math.cos(theta)]
    table = oracle.solid_table(8, unit)
    for p in range(9):
        for m in range(-p, p + 1):
            expected = lpmv(m, p, math.cos(theta)) * np.exp(1j * m * phi)  # Condon-Shortley phase
            assert abs(table[p * p + p + m] - expected) <= 1e-12 * max(1.0, abs(expected)), (p, m)


def _block_entry(block, lmax, to, frm):
    l, m, pol = oracle_modes(lmax)
    idx = {(int(a), int(b), int(c)): i for i, (a, b, c) in enumerate(zip(l, m, pol))}
    return block[idx[to], idx[frm]]


def oracle_modes(lmax):
    import oracle as o

    return o.modes(lmax)


def test_translation_against_closed_dipole_form(oracle):
    # tests/waves/test_translation_coefficients.py:63-95 at a real argument: the same
    # closed forms hold for any z.  A = (C++ + C--)/2, B = (C++ - C--)/2.
    z = 1.4
    h1 = -np.exp(1j * z) * (z + 1j) / z ** 2
    h2 = np.exp(1j * z) * (1j / z - 3 / z ** 2 - 3j / z ** 3)
    j1 = np.sin(z) / z ** 2 - np.cos(z) / z
    j2 = (3 / z ** 3 - 1 / z) * np.sin(z) - 3 * np.cos(z) / z ** 2
    for theta in [1e-8, 1e-3, 0.4]:
        phi = 0.4
        d = np.array([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)])
        for regular, r1, r2 in [(False, h1, h2), (True, j1, j2)]:
            block = oracle.translation_block(1, 1.0, z * d, regular)
            pp = _block_entry(block, 1, (1, 0, 1), (1, 1, 1))
            mm = _block_entry(block, 1, (1, 0, 0), (1, 1, 0))
            A, B = (pp + mm) / 2, (pp - mm) / 2
            phase = np.exp(0.4j) * np.sin(theta)
            expected_a = -3 / (2 * np.sqrt(2)) * r2 * np.cos(theta) * phase
            expected_b = 3j / (2 * np.sqrt(2)) * r1 * phase
            assert abs(A - expected_a) <= 3e-13 * abs(expected_a), (theta, regular)
            assert abs(B - expected_b) <= 3e-13 * abs(expected_b), (theta, regular)


def test_translation_symmetries(oracle):
    lmax, k = 3, 1.3
    l, m, pol = oracle.modes(lmax)
    M = l.size
    d = np.array([0.7, -0.4, 1.1])
    for regular in (False, True):
        A = oracle.translation_block(lmax, k, d, regular)
        Ainv = oracle.translation_block(lmax, k, -d, regular)
        scale = np.abs(A).max()
        idx = {(int(a), int(b), int(c)): i for i, (a, b, c) in enumerate(zip(l, m, pol))}
        for t in range(M):
            for s in range(M):
                if pol[t] != pol[s]:
                    assert A[t, s] == 0
                    continue
                # parity: A(-d) = (-1)^(l+lambda) P A(d) P (P swaps helicity); bitwise
                flip_t = idx[(int(l[t]), int(m[t]), 1 - int(pol[t]))]
                flip_s = idx[(int(l[s]), int(m[s]), 1 - int(pol[s]))]
                sign = (-1) ** int(l[t] + l[s])
                assert Ainv[t, s] == sign * A[flip_t, flip_s]
                # reciprocity: A_(t,s)(d) = (-1)^(m+mu) A_(s',t')(-d) with m -> -m
                rs = idx[(int(l[s]), -int(m[s]), int(pol[s]))]
                rt = idx[(int(l[t]), -int(m[t]), int(pol[t]))]
                sgn = (-1) ** int(m[t] + m[s])
                assert abs(A[t, s] - sgn * Ainv[rs, rt]) <= 1e-12 * scale
        # azimuthal covariance
        alpha = 0.37
        c, s_ = np.cos(alpha), np.sin(alpha)
        Arot = oracle.translation_block(lmax, k, [d[0] * c - d[1] * s_, d[0] * s_ + d[1] * c, d[2]], regular)
        phase = np.exp(1j * (m[None, :] - m[:, None]) * alpha)
        assert np.abs(Arot - phase * A).max() <= 1e-12 * scale
    # regular block at the origin is the identity
    assert np.abs(oracle.translation_block(lmax, k, [0, 0, 0], True) - np.eye(M)).max() <= 1e-13


def test_translation_pullback_finite_differences(oracle):
    lmax, k = 2, 1.3
    d = np.array([0.9, -0.3, 0.6])
    M = 2 * lmax * (lmax + 2)
    rng = np.random.default_rng(3)
    G = rng.normal(size=(M, M)) + 1j * rng.normal(size=(M, M))
    for regular in (False, True):
        grad, gk = oracle.translation_pullback(lmax, k, d, regular, G)
        loss = lambda dd, kk: np.real(np.vdot(G, oracle.translation_block(lmax, kk, dd, regular)))
        h = 1e-6
        for a in range(3):
            e = np.zeros(3)
            e[a] = h
            fd = (loss(d + e, k) - loss(d - e, k)) / (2 * h)
            assert abs(fd - grad[a]) <= 1e-7 * max(1.0, abs(fd))
        fdk = (loss(d, k + h) - loss(d, k - h)) / (2 * h)
        assert abs(fdk - gk.real) <= 1e-7 * max(1.0, abs(fdk))
        # Euler identity d . g_d = Re(k conj(g_k))
        assert abs(np.dot(d, grad) - (k * np.conj(gk)).real) <= 1e-10 * np.abs(grad).sum()


def test_mie_metallic_golden(oracle, golden_dir):
    g = _load(golden_dir, "reference_examples.json")["metallic_mie"]
    eps = complex(*g["epsilon"])
    for l, dre, dim, ore, oim in g["rows"]:
        t = oracle.mie(l, g["x"], eps)
        diag, off = complex(dre, dim), complex(ore, oim)
        expected = np.array([[diag, off], [off, diag]])
        assert np.all(np.abs(t - expected) <= 2e-13 + 2e-11 * np.abs(expected)), l


def test_mie_against_bohren_huffman(oracle, golden_dir):
    # Helicity Mie matrix of an achiral sphere: diag -(a+b)/2, off-diagonal -(a-b)/2
    # (treams' T = -a_n (N), -b_n (M)); confirmed on the metallic golden above.
    for row in _load(golden_dir, "special_functions.json")["mie_bohren_huffman"]:
        a, b = complex(*row["a"]), complex(*row["b"])
        t = oracle.mie(row["n"], row["x"], complex(*row["eps"]))
        expected = -np.array([[a + b, a - b], [a - b, a + b]]) / 2
        assert np.all(np.abs(t - expected) <= 1e-12 * np.abs(expected).max()), row


def test_mie_pullback_finite_differences(oracle):
    rng = np.random.default_rng(0)
    G = rng.normal(size=(2, 2)) + 1j * rng.normal(size=(2, 2))
    for l, size, eps in [(1, 0.4, 4 + 0.1j), (3, 1.2, 2.5 + 0.05j)]:
        gs, ge = oracle.mie_pullback(l, size, eps, G)
        loss = lambda s, e: np.real(np.vdot(G, oracle.mie(l, s, e)))
        h = 1e-6
        assert abs((loss(size + h, eps) - loss(size - h, eps)) / (2 * h) - gs) <= 1e-7
        fr = (loss(size, eps + h) - loss(size, eps - h)) / (2 * h)
        fi = (loss(size, eps + 1j * h) - loss(size, eps - 1j * h)) / (2 * h)
        assert abs(complex(fr, fi) - ge) <= 1e-7


def test_single_sphere_optical_theorem(oracle):
    # One lossless sphere: sigma_ext = sigma_sca; the cluster is its Mie T-matrix.
    cl = oracle.Cluster(4, 1.3, [0.3], 4.0, [[0.1, -0.2, 0.05]])
    a = cl.plane_wave([0, 0, 1], 1)
    sca, ext = cl.cross_sections(a)
    assert abs(sca - ext) <= 1e-12 * abs(ext)


def _cluster_example(oracle, golden_dir):
    g = _load(golden_dir, "reference_examples.json")["cluster_example"]
    pos = (220 / np.sqrt(24)) * np.array(
        [[np.sqrt(8), 0, -1], [-np.sqrt(2), np.sqrt(6), -1], [-np.sqrt(2), -np.sqrt(6), -1], [0, 0, 3]])
    k0 = 2 * np.pi / g["k0_over_2pi_inverse"]
    cl = oracle.Cluster(g["lmax"], k0, g["radii"], complex(*g["epsilon"]), pos)
    return cl, g


def test_cluster_example_cross_sections(oracle, golden_dir):
    cl, g = _cluster_example(oracle, golden_dir)
    a = cl.plane_wave(g["direction"], g["pol"])
    sca, ext = cl.cross_sections(a)
    e = g["expected"]
    assert float(f"{sca:.6g}") == e["scattering"]
    assert float(f"{ext:.6g}") == e["extinction"]
    assert float(f"{ext - sca:.6g}") == e["absorption"]


def test_cluster_example_intensity_map(oracle, golden_dir):
    """The total-field intensity 0.5 |E_inc + E_sca|^2 of docs/examples/cluster.py on
    its 7 x 7 grid (local expansion) against docs/examples/output/cluster.txt at the
    6 digits it prints; points inside a sphere are NaN there and skipped here."""
    cl, g = _cluster_example(oracle, golden_dir)
    a = cl.plane_wave(g["direction"], g["pol"])
    x = cl.solve(a)
    gx = np.linspace(-300, 300, 7)
    pts = np.array([[u, 0.0, w] for u in gx for w in gx])
    inside = np.array([any(np.linalg.norm(p - cl.positions[i]) <= cl.radii[i] for i in range(cl.N)) for p in pts])
    expected = g["intensity_local"]["values"]
    assert [[v is None for v in row] for row in expected] == inside.reshape(7, 7).tolist()
    out = pts[~inside]
    e = oracle.plane_wave_field(cl.k0, g["direction"], g["pol"], out) + \
        oracle.cluster_field(cl.lmax, cl.k0, cl.positions, x, out)
    got = 0.5 * np.sum(np.abs(e) ** 2, axis=1)
    want = np.array([v for row in expected for v in row if v is not None])
    assert [float(f"{v:.6g}") for v in got] == want.tolist()


def _gradient_example():
    idx = np.arange(8)
    positions = np.column_stack((180 * (idx % 2), 190 * ((idx // 2) % 2), 200 * (idx // 4))).astype(float)
    positions += np.column_stack((9 * np.sin(idx), 7 * np.cos(idx), 11 * np.sin(2 * idx)))
    positions -= positions.mean(axis=0)
    radii = 45.0 + 3 * (idx % 4)
    return radii, positions, 2 * np.pi / 700


def test_reverse_gradient_example(oracle, golden_dir):
    g = _load(golden_dir, "reference_examples.json")["reverse_gradient_example"]
    radii, positions, k0 = _gradient_example()
    cl = oracle.Cluster(g["lmax"], k0, radii, complex(*g["epsilon"]), positions)
    sca, gr, gp = cl.scattering_gradient(g["direction"], g["pol"])
    assert round(sca, 8) == g["expected_scattering"]
    expected = np.array(g["expected_gradient_radius_xyz"])
    actual = np.column_stack((gr, gp))
    assert np.abs(np.round(actual, 6) - expected).max() <= 1.5e-6


def sic_chain(oracle):
    """Four 250 nm SiC spheres with 20 nm gaps on the z axis at 3 THz, lmax 10: the
    reference's test_small_wavenumber_high_order_chain_absorption_is_stable
    (tests/tmatrix/test_particle_cluster.py:538-559)."""
    k0 = 2 * np.pi * 3e12 / 299792458.0
    pos = np.array([[0, 0, z * 520e-9] for z in (-1.5, -0.5, 0.5, 1.5)])
    return oracle.Cluster(10, k0, [250e-9] * 4, 12.87707323 + 0.19225647000000007j, pos), k0


def test_sic_chain_global_absorption_golden(oracle):
    """The reference's golden absorption (rtol 2e-11) of the global T-matrix of a
    cluster whose operator has unit diagonal and entries above 1e22: only the
    equilibrated solve reaches it (the unscaled one misses by ~5e-4)."""
    cl, k0 = sic_chain(oracle)
    assert np.abs(cl.operator()).max() > 1e22
    Tg = cl.global_tmatrix()
    absorption = -2 * np.pi * (np.trace(Tg).real + np.vdot(Tg, Tg).real) / k0 ** 2
    assert abs(absorption * 1e12 / 0.0001754433586134972 - 1) <= 2e-11
    # axial symmetry: no coupling between distinct azimuthal orders
    m = np.array([mm for l in range(1, 11) for mm in range(-l, l + 1) for _ in (0, 1)])
    assert np.abs(Tg[m[:, None] != m[None, :]]).max() <= 2e-20
    unscaled = cl.global_tmatrix(equilibrate=False)
    a0 = -2 * np.pi * (np.trace(unscaled).real + np.vdot(unscaled, unscaled).real) / k0 ** 2
    assert abs(a0 * 1e12 / 0.0001754433586134972 - 1) > 1e-6


def test_gmres_matches_dense(oracle):
    rng = np.random.default_rng(1)
    N, lmax = 5, 2
    pos = rng.uniform(-1.5, 1.5, size=(N, 3))
    pos[:, 0] += 2.0 * np.arange(N)
    cl = oracle.Cluster(lmax, 1.3, rng.uniform(0.2, 0.3, N), 4 + 0.1j, pos)
    a = cl.plane_wave([0, 0, 1], 1)
    x_dense = cl.solve(a)
    x_it, conv = cl.solve_gmres(a, rtol=1e-12)
    assert conv.residual_norm <= 1e-12 * conv.rhs_norm
    assert np.linalg.norm(x_it - x_dense) <= 1e-10 * np.linalg.norm(x_dense)


def test_batched_gmres_and_matrix_free_cluster_match_dense(oracle):
    rng = np.random.default_rng(4)
    N, lmax = 9, 3
    pos = rng.uniform(-2.0, 2.0, size=(N, 3))
    pos[:, 0] += 1.6 * np.arange(N)
    radii = rng.uniform(0.2, 0.3, N)
    eps = 4 + 0.1j
    dense = oracle.Cluster(lmax, 1.3, radii, eps, pos)
    mf = oracle.MatrixFreeCluster(lmax, 1.3, radii, eps, pos)
    a = np.stack([dense.plane_wave(d, p) for d, p in (([0, 0, 1], 1), ([0, 0, -1], 0), ([1, 0, 0], 1),
                                                        ([0, 1, 1], 0))], axis=1)
    x_dense = np.linalg.solve(dense.operator(), oracle.apply_local(dense.blocks, a, False))
    X, conv = mf.solve(a, rtol=1e-12, restart=10)
    for c in conv:
        assert c.residual_norm <= 1e-12 * c.rhs_norm
    assert np.linalg.norm(X - x_dense) <= 1e-10 * np.linalg.norm(x_dense)
    # the batch changes when products run, not the arithmetic of a column
    x0, c0 = dense.solve_gmres(a[:, 1], rtol=1e-12, restart=10)
    assert conv[1].iterations == c0.iterations
    assert np.linalg.norm(mf.overlap(X) - dense.R @ X) <= 1e-14 * np.linalg.norm(dense.R @ X)
    sca, ext = mf.cross_sections(a, X)
    for c in range(4):
        s_ref, e_ref = dense.cross_sections(a[:, c], x_dense[:, c])
        assert abs(sca[c] - s_ref) <= 1e-11 * abs(s_ref) and abs(ext[c] - e_ref) <= 1e-11 * abs(e_ref)


def test_gradient_planted_mistake_is_caught(oracle, golden_dir):
    """A dropped term (the incident position phase) must break the golden."""
    g = _load(golden_dir, "reference_examples.json")["reverse_gradient_example"]
    radii, positions, k0 = _gradient_example()
    cl = oracle.Cluster(g["lmax"], k0, radii, complex(*g["epsilon"]), positions)
    kvec = k0 * np.array(g["direction"]) / np.linalg.norm(g["direction"])
    a = cl.plane_wave(g["direction"], g["pol"])
    x = cl.solve(a)
    G = (cl.R + cl.R.conj().T) @ x / k0 ** 2
    grads = cl.pullback(a, x, G)
    Rbar = np.outer(x, x.conj()) / k0 ** 2
    gR, _ = oracle.pair_matrix_pullback(g["lmax"], k0, positions, True, Rbar)
    wrong = grads["positions"] + gR  # phase term dropped
    expected = np.array(g["expected_gradient_radius_xyz"])[:, 1:]
    assert np.abs(wrong - expected).max() > 1e-3
    del kvec
