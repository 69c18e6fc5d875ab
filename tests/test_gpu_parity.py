"""GPU parity of the HIP path against the FP64 oracle and the reference goldens.

Tolerances (FP64 paths): 1e-12 relative on translation blocks, Mie blocks and
operator applications; 1e-10 on solutions; the reference example goldens at the
digits they print.  c64 stored paths: 2e-6 relative (single-precision storage).
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


def _cluster(pkg, oracle, N=6, lmax=3, seed=1, side=3.0, eps=4 + 0.1j, k0=1.3):
    radii, e, pos = pkg.random_cluster(N, side, 0.1, (0.2, 0.3), eps, seed=seed)
    return pkg.IterativeSphereCluster(lmax, k0, radii, e, pos), oracle.Cluster(lmax, k0, radii, e, pos)


def test_mie_blocks(pkg, oracle):
    radii = np.array([0.15, 0.25, 0.35, 0.3])
    eps = np.array([2.0 + 0.0j, 4 + 0.1j, 6 + 0.2j, 16 + 0.5j])
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1.2, 0], [0, 0, 1.5]])
    cl = pkg.IterativeSphereCluster(6, 1.3, radii, eps, pos)
    got = cl.mie_blocks()
    ref = oracle.tmatrix_blocks(6, 1.3, radii, eps)
    for i in range(4):
        for l in range(6):
            scale = np.abs(ref[i, l]).max()
            assert np.abs(got[i, l] - ref[i, l]).max() <= 1e-12 * max(scale, 1e-30) + 1e-300, (i, l)


def test_mie_metallic_golden(pkg, golden_dir):
    g = json.load(open(os.path.join(golden_dir, "reference_examples.json")))["metallic_mie"]
    # size x = k0 r = 80 with k0 = 1: one sphere, lmax up to 12 (rows l = 1, 3)
    cl = pkg.IterativeSphereCluster(12, 1.0, [80.0], complex(*g["epsilon"]), [[0, 0, 0]])
    blocks = cl.mie_blocks()[0]
    for l, dre, dim, ore, oim in g["rows"]:
        if l > 12:
            continue
        d, o = complex(dre, dim), complex(ore, oim)
        expected = np.array([[d, o], [o, d]])
        assert np.all(np.abs(blocks[l - 1] - expected) <= 2e-13 + 2e-11 * np.abs(expected)), l


@pytest.mark.parametrize("regular", [False, True])
@pytest.mark.parametrize("lmax", [1, 3, 6])
def test_translation_blocks(pkg, oracle, lmax, regular):
    rng = np.random.default_rng(lmax)
    disp = rng.normal(size=(5, 3))
    disp[0] = [0, 0, 0.7]
    disp[1] = [0, 0, -1.1]
    disp[2] *= 0.3  # kr ~ 0.4 (below the reference's recurrence domain)
    got = pkg.translation_blocks(lmax, 1.3, disp, regular).cpu().numpy()
    for i, d in enumerate(disp):
        ref = oracle.translation_block(lmax, 1.3, d, regular)
        assert np.abs(got[i] - ref).max() <= 1e-12 * np.abs(ref).max(), i
    if regular:
        origin = pkg.translation_blocks(lmax, 1.3, [[0.0, 0.0, 0.0]], True).cpu().numpy()[0]
        assert np.abs(origin - np.eye(origin.shape[0])).max() <= 1e-13


@pytest.mark.parametrize("N,lmax", [(8, 3), (5, 6)])
def test_dense_coupling_and_overlap(pkg, oracle, N, lmax):
    cl, ref = _cluster(pkg, oracle, N=N, lmax=lmax, seed=N + lmax)
    C = cl.coupling_matrix().cpu().numpy()
    assert _rel(C, ref.C) <= 1e-12
    for i in range(N):  # self blocks exactly zero
        M = ref.M
        assert np.all(C[i * M:(i + 1) * M, i * M:(i + 1) * M] == 0)
    R = cl.coupling_matrix(regular=True).cpu().numpy()
    assert _rel(R, ref.R) <= 1e-12


def test_operator_apply_matrix_free_and_stored(pkg, oracle):
    cl, ref = _cluster(pkg, oracle, N=7, lmax=4, seed=3)
    rng = np.random.default_rng(0)
    x = rng.normal(size=(ref.n, 3)) + 1j * rng.normal(size=(ref.n, 3))
    for adjoint in (False, True):
        got = cl.apply(x, adjoint=adjoint)
        exp = np.stack([ref.apply(x[:, c], adjoint) for c in range(3)], axis=1)
        assert _rel(got, exp) <= 1e-12
        gc = cl.coupling(x, adjoint=adjoint)
        ec = (ref.C.conj().T if adjoint else ref.C) @ x
        assert _rel(gc, ec) <= 1e-12
    assert _rel(cl.local(x), ref.T @ x) <= 1e-13
    assert _rel(cl.local(x, adjoint=True), ref.T.conj().T @ x) <= 1e-13
    cl.assemble("c128")
    assert _rel(cl.coupling(x, stored=True), ref.C @ x) <= 1e-12
    assert _rel(cl.coupling(x, adjoint=True, stored=True), ref.C.conj().T @ x) <= 1e-12
    cl.assemble("c64")
    assert _rel(cl.coupling(x, stored=True), ref.C @ x) <= 2e-6
    assert _rel(cl.apply(x[:, 0], stored=True), ref.apply(x[:, 0])) <= 2e-6


def test_operator_apply_matrix_free_c64(pkg, oracle):
    """The c64 matrix-free operator (c64 tables, FP32 MFMA, FP64 contraction) against
    the FP64 oracle: single-precision level, forward and adjoint, ragged and full
    GEMM tiles (N = 6 and N = 32)."""
    for N, seed in ((6, 3), (32, 9)):
        cl, ref = _cluster(pkg, oracle, N=N, lmax=3, seed=seed, side=0.9 * N ** (1 / 3) * 1.6)
        rng = np.random.default_rng(N)
        x = rng.normal(size=(ref.n, 3)) + 1j * rng.normal(size=(ref.n, 3))
        assert _rel(cl.apply(x, c64=True), ref.apply(x)) <= 2e-6
        assert _rel(cl.apply(x, adjoint=True, c64=True), ref.apply(x, adjoint=True)) <= 2e-6
        assert _rel(cl.coupling(x, c64=True), ref.C @ x) <= 1e-5


@pytest.mark.parametrize("N,lmax,cols", [(6, 3, 3), (45, 4, 5), (96, 6, 4), (130, 3, 12)])
def test_inner_gemm_split16_vs_fp32(pkg, N, lmax, cols):
    """The split-FP16 inner coupling (scaled hi + lo halves on FP16 MFMA) against the
    FP64 operator, beside the FP32 form: single-precision level for both, on ragged
    sizes (N not a multiple of 32, partial row and column tiles), several right-hand
    sides, and a source vector whose sphere amplitudes span 1e-9 .. 1 (the per-column
    shifts and the 2^17 full-precision window)."""
    radii, e, pos = pkg.random_cluster(N, 1.9 * N ** (1 / 3), 0.05, (0.15, 0.35), None, seed=N)
    cl = pkg.IterativeSphereCluster(lmax, 1.3, radii, e, pos)
    rng = np.random.default_rng(N)
    x = rng.normal(size=(cl.n, cols)) + 1j * rng.normal(size=(cl.n, cols))
    amp = 10.0 ** rng.uniform(-9, 0, size=N)
    xs = x * np.repeat(amp, cl.n // N)[:, None]
    errs = {}
    first = None
    for kind in ("split16", "fp32"):
        cl.set_inner_gemm(kind)
        for name, v in (("plain", x), ("spread", xs)):
            ref = cl.apply(v).cpu().numpy()
            got = cl.apply(v, c64=True)
            if kind == "split16" and name == "plain":
                first = got.cpu().numpy()
            errs[kind, name] = _rel(got, ref)
            assert errs[kind, name] <= 2e-6, (kind, name, errs[kind, name])
        errs[kind, "coupling"] = _rel(cl.coupling(x, c64=True), cl.coupling(x))
        assert errs[kind, "coupling"] <= 1e-5, (kind, errs[kind, "coupling"])
    print(N, lmax, {k: f"{v:.2e}" for k, v in errs.items()})
    # split tables rebuilt (row exponents and fragments again from the recurrences): the
    # same tables, the same result bit for bit
    cl.set_inner_gemm("split16")
    assert np.array_equal(cl.apply(x, c64=True).cpu().numpy(), first)


def _log2_table_bound(k0, pos, lmax):
    """log2 of max over p <= 2 lmax of |h_p(k0 dmin)| sqrt((2p)!), the bound on every
    singular pair-table entry, from scipy's spherical Bessel functions."""
    from scipy.special import gammaln, spherical_jn, spherical_yn

    pos = np.asarray(pos)
    d = np.linalg.norm(pos[:, None] - pos[None], axis=-1)[np.triu_indices(len(pos), 1)]
    x, p = k0 * d.min(), np.arange(2 * lmax + 1)
    h = np.hypot(spherical_jn(p, x), spherical_yn(p, x))
    return float(np.max(np.log2(h) + 0.5 * gammaln(2 * p + 1) / np.log(2)))


def test_high_degree_split16_in_range(pkg):
    """lmax 10 with well separated spheres (tables far inside FP32 range): the split-FP16
    coupling runs (its GEMM is timed) at single-precision level against FP64."""
    g = np.arange(2) * 6.0
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    assert _log2_table_bound(1.3, pos, 10) + np.log2(len(pos)) < 105
    cl = pkg.IterativeSphereCluster(10, 1.3, np.full(8, 0.3), 4 + 0.1j, pos)
    cl.set_inner_gemm("split16")
    cl.timing(True)
    x = np.random.default_rng(3).normal(size=(cl.n, 3)) + 0j
    err = _rel(cl.apply(x, c64=True), cl.apply(x))
    print("lmax 10 split16", f"{err:.2e}")
    assert err <= 2e-6
    assert cl.timing_get("gemm_split16")[1] > 0


def test_c64_paths_beyond_fp32_range(pkg):
    """lmax 10 with close spheres: singular tables beyond FP32 range.  Every c64-precision
    path (split-FP16 and FP32 matrix-free couplings, the stored c64 coupling, the c64 LU,
    the learned preconditioner's bfloat16 near blocks) falls back to FP64 tables and stays
    finite and correct; the split GEMM does not run."""
    lmax = 10
    radii, e, pos = pkg.random_cluster(12, 1.9 * 12 ** (1 / 3), 0.05, (0.15, 0.35), None, seed=12)
    assert _log2_table_bound(1.3, pos, lmax) > 125
    cl = pkg.IterativeSphereCluster(lmax, 1.3, radii, e, pos)
    cl.timing(True)
    rng = np.random.default_rng(12)
    x = rng.normal(size=(cl.n, 2)) + 1j * rng.normal(size=(cl.n, 2))
    ref, cref = cl.apply(x).cpu().numpy(), cl.coupling(x).cpu().numpy()
    errs = {}
    for kind in ("split16", "fp32"):
        cl.set_inner_gemm(kind)
        errs[kind] = _rel(cl.apply(x, c64=True), ref)
        errs[kind, "coupling"] = _rel(cl.coupling(x, c64=True), cref)
    cl.assemble("c64")
    errs["stored"] = _rel(cl.coupling(x, stored=True), cref)
    errs["stored apply"] = _rel(cl.apply(x, stored=True), ref)
    print("lmax 10 close", {k: f"{v:.2e}" for k, v in errs.items()})
    assert all(np.isfinite(v) and v <= 2e-6 for v in errs.values()), errs
    assert cl.timing_get("gemm_split16")[1] == 0
    a = cl.plane_wave([0, 0, 1], 1)
    xs = {}
    for prec in ("c128", "c64"):
        f = pkg.SphereClusterFactor(cl, prec, rtol=1e-12)
        xs[prec], conv = f.solve(a)
        f.close()
        print("lmax 10 close LU", prec, conv)
    assert np.isfinite(xs["c64"].cpu().numpy()).all()
    assert _rel(xs["c64"], xs["c128"]) <= 1e-10
    import importlib

    learned = importlib.import_module("treams_rs_amd.learned")
    learned.attach(cl, learned.MLP.load(learned.DEFAULT_WEIGHTS))
    xl, cl_conv = cl.solve(a, rtol=1e-9, mixed_precision="matrix_free", preconditioner="learned", restart=50)
    print("lmax 10 close learned", cl_conv)
    assert cl_conv[0][1] <= 1e-9 * cl_conv[0][2]
    assert _rel(xl, xs["c128"]) <= 1e-7


@pytest.mark.parametrize("N,lmax", [(45, 4), (96, 6), (130, 3)])
def test_one_pass_tables_equal_fp64_tables(pkg, N, lmax):
    """The one-pass table build (split FP16 and FP64 tables together, the path of a
    cluster whose first coupling is c64-precision) writes the same FP64 tables, bit for
    bit, as the FP64 table kernel (the path of a cluster whose first coupling is FP64):
    both evaluate the recurrences with the same explicit FMAs."""
    radii, e, pos = pkg.random_cluster(N, 1.9 * N ** (1 / 3), 0.05, (0.15, 0.35), None, seed=N)
    rng = np.random.default_rng(1)
    n = N * 2 * lmax * (lmax + 2)
    x = rng.normal(size=(n, 3)) + 1j * rng.normal(size=(n, 3))
    a = pkg.IterativeSphereCluster(lmax, 1.3, radii, e, pos)
    ya = a.coupling(x)
    b = pkg.IterativeSphereCluster(lmax, 1.3, radii, e, pos)
    yb16 = b.coupling(x, c64=True)
    yb = b.coupling(x)
    assert torch.equal(ya, yb)
    assert _rel(yb16, yb) <= 1e-6


@pytest.mark.parametrize("N,lmax,cols", [(6, 3, 3), (45, 4, 5), (96, 6, 4)])
def test_inner_relaxed_hi_only_coupling(pkg, oracle, N, lmax, cols):
    """The relaxed inner coupling (split-FP16 tables, hi halves only: 11 significant bits
    per value, FP32 accumulation) against the oracle's FP64 coupling on a row subset and
    against the library's FP64 coupling on every row: half-precision level (2^-11 per
    operand), forward and adjoint; and measurably coarser than the hi + lo form, so the
    hi-only kernel is the one that ran."""
    radii, e, pos = pkg.random_cluster(N, 1.9 * N ** (1 / 3), 0.05, (0.15, 0.35), None, seed=N + 1)
    cl = pkg.IterativeSphereCluster(lmax, 1.3, radii, e, pos)
    rng = np.random.default_rng(N)
    x = rng.normal(size=(cl.n, cols)) + 1j * rng.normal(size=(cl.n, cols))
    ref = cl.coupling(x).cpu().numpy()
    M = cl.n // N
    rows = (0, min(N, 3))
    yr = oracle.coupling_apply(lmax, 1.3, pos, x, rows=rows)[rows[0] * M:rows[1] * M]
    assert _rel(ref[rows[0] * M:rows[1] * M], yr) <= 1e-12
    hi = cl.coupling(x, c64="hi")
    full = cl.coupling(x, c64=True)
    err_hi, err_full = _rel(hi, ref), _rel(full, ref)
    assert _rel(hi.cpu().numpy()[rows[0] * M:rows[1] * M], yr) <= 2e-3
    assert err_hi <= 2e-3, err_hi
    assert err_hi >= 10 * err_full, (err_hi, err_full)
    assert _rel(cl.coupling(x, adjoint=True, c64="hi"), cl.coupling(x, adjoint=True)) <= 2e-3
    print(N, lmax, f"hi-only {err_hi:.2e}, hi + lo {err_full:.2e}")


def test_relaxed_inner_solves_certified(pkg, oracle):
    """Mixed solves on split-FP16 tables with relaxed inner applies give the answer of
    the unrelaxed solve: relaxation off, the default threshold, and every inner apply
    hi-only (threshold 1e-300: more FP64 refinements, same certified tolerance).  The
    timers show which GEMM ran."""
    cl, ref = _cluster(pkg, oracle, N=24, lmax=4, seed=8, side=4.5)
    B = torch.stack([cl.plane_wave([0, 0, 1], 1), cl.plane_wave([0, 0, -1], 0)], dim=1)
    xr = ref.solve(B.cpu().numpy())
    out = {}
    for th in (0.0, 1e-3, 1e-300):
        cl.set_inner_relaxation(th)
        cl.timing_reset()
        cl.timing(True)
        x, conv = cl.solve(B, rtol=1e-10, restart=50, mixed_precision="matrix_free", convergence="full")
        cl.timing(False)
        for c in conv:
            assert c.residual_norm <= 1e-10 * c.rhs_norm
        assert _rel(x, xr) <= 1e-8
        out[th] = (cl.timing_get("gemm_split16")[1], cl.timing_get("gemm_split16_hi")[1],
                   [c.iterations for c in conv], [c.refinements for c in conv])
    print(out)
    assert out[0.0][1] == 0 and out[0.0][0] > 0
    assert out[1e-300][0] == 0 and out[1e-300][1] > 0
    assert out[1e-3][1] > 0
    # the default relaxation leaves the iteration and refinement counts unchanged
    assert out[1e-3][2:] == out[0.0][2:]
    cl.set_inner_relaxation(1e-3)
    with pytest.raises(ValueError):
        cl.set_inner_relaxation(-1.0)


def test_solve_mixed_matrix_free_c64(pkg, oracle):
    cl, ref = _cluster(pkg, oracle, N=12, lmax=4, seed=8, side=3.5)
    a = cl.plane_wave([0, 0, 1], 1)
    cl.timing(True)
    x, conv = cl.solve(a, rtol=1e-10, restart=50, mixed_precision="matrix_free", convergence="full")
    cl.timing(False)
    assert conv[0].residual_norm <= 1e-10 * conv[0].rhs_norm
    assert conv[0].refinements >= 1
    assert _rel(x, ref.solve(a.cpu().numpy())) <= 1e-8
    # the inner solves ran on the c64 matrix-free operator, never on a stored coupling
    assert cl.timing_get("coupling_mf32")[1] > 0
    assert cl.timing_get("matvec_stored")[1] == 0


def test_matrix_free_apply_larger_cluster(pkg, oracle):
    """N = 96, lmax = 6: pair-table matrix-free apply vs the oracle on a row subset
    and vs the stored c128 product (an independent kernel) on every row."""
    radii, e, pos = pkg.random_cluster(96, 7.0, 0.05, (0.15, 0.35), None, seed=21)
    cl = pkg.IterativeSphereCluster(6, 1.3, radii, e, pos)
    rng = np.random.default_rng(5)
    x = rng.normal(size=(cl.n, 3)) + 1j * rng.normal(size=(cl.n, 3))
    y = cl.coupling(x).cpu().numpy()
    rows = (10, 14)
    yr = oracle.coupling_apply(6, 1.3, pos, x, rows=rows)
    sl = slice(rows[0] * 96, rows[1] * 96)
    assert _rel(y[sl], yr[sl]) <= 1e-12
    cl.assemble("c128")
    assert _rel(cl.coupling(x, stored=True), y) <= 1e-12
    ya = cl.coupling(x, adjoint=True).cpu().numpy()
    assert _rel(cl.coupling(x, adjoint=True, stored=True), ya) <= 1e-12


def test_overlap_apply(pkg, oracle):
    cl, ref = _cluster(pkg, oracle, N=5, lmax=3, seed=9)
    rng = np.random.default_rng(4)
    x = rng.normal(size=ref.n) + 1j * rng.normal(size=ref.n)
    assert _rel(cl.overlap(x), ref.R @ x) <= 1e-12
    assert _rel(cl.overlap(x, adjoint=True), ref.R.conj().T @ x) <= 1e-12


def test_overlap_apply_tables_larger(pkg, oracle):
    """Regular overlap from the regular pair tables (GEMM path) at N = 33 (ragged GEMM
    tiles and K tail) against the oracle's dense R, both directions, 2 columns."""
    cl, ref = _cluster(pkg, oracle, N=33, lmax=3, seed=13, side=5.0)
    rng = np.random.default_rng(6)
    x = rng.normal(size=(ref.n, 2)) + 1j * rng.normal(size=(ref.n, 2))
    assert _rel(cl.overlap(x), ref.R @ x) <= 1e-12
    assert _rel(cl.overlap(x, adjoint=True), ref.R.conj().T @ x) <= 1e-12


def test_table_limit_selects_per_pair_path(pkg, oracle):
    """With a table limit below one table set the matrix-free operators recompute the
    translation coefficients per pair: same results as the table GEMMs and the
    oracle (FP64), c64 solves included; a negative limit is rejected."""
    cl, ref = _cluster(pkg, oracle, N=20, lmax=4, seed=17, side=4.0)
    rng = np.random.default_rng(8)
    x = rng.normal(size=(ref.n, 2)) + 1j * rng.normal(size=(ref.n, 2))
    with_tables = [cl.coupling(x), cl.coupling(x, adjoint=True), cl.overlap(x), cl.apply(x, c64=True)]
    cl.set_table_limit(0)
    per_pair = [cl.coupling(x), cl.coupling(x, adjoint=True), cl.overlap(x), cl.apply(x, c64=True)]
    for a, b in zip(with_tables[:3], per_pair[:3]):
        assert _rel(a, b) <= 1e-12
    assert _rel(per_pair[0], ref.C @ x) <= 1e-12
    assert _rel(per_pair[2], ref.R @ x) <= 1e-12
    assert _rel(per_pair[3], ref.apply(x)) <= 1e-12  # c64 falls back to the FP64 per-pair path
    assert _rel(with_tables[3], ref.apply(x)) <= 2e-6
    a = cl.plane_wave([0, 0, 1], 1)
    xs, conv = cl.solve(a, rtol=1e-10, mixed_precision="matrix_free", convergence="full")
    assert _rel(xs, ref.solve(a.cpu().numpy())) <= 1e-8
    with pytest.raises(Exception):
        cl.set_table_limit(-1)


@pytest.mark.parametrize("direction", [[0, 0, 1], [0, 0, -1], [1, 0, 0], [0.2, -0.1, 1.0]])
@pytest.mark.parametrize("pol", [0, 1])
def test_plane_wave(pkg, oracle, direction, pol):
    cl, ref = _cluster(pkg, oracle, N=4, lmax=6, seed=2)
    assert _rel(cl.plane_wave(direction, pol), ref.plane_wave(direction, pol)) <= 1e-13


def test_solve_fp64_matches_dense(pkg, oracle):
    cl, ref = _cluster(pkg, oracle, N=8, lmax=3, seed=5, side=2.0)
    a = torch.stack([cl.plane_wave([0, 0, 1], 1), cl.plane_wave([0, 0, -1], 0)], dim=1)
    x, conv = cl.solve(a, rtol=1e-12)
    an = a.cpu().numpy()
    for c in range(2):
        xr = ref.solve(an[:, c])
        assert _rel(x[:, c], xr) <= 1e-10
        its, resid, rhs = conv[c]
        assert resid <= 1e-12 * rhs
        # the certificate is the true residual
        r = np.linalg.norm(ref.operator() @ x[:, c].cpu().numpy() - oracle.apply_local(ref.blocks, an[:, c]))
        assert abs(r - resid) <= 1e-3 * resid + 1e-14 * rhs


def test_solve_mixed_precision_reaches_fp64_tolerance(pkg, oracle):
    cl, ref = _cluster(pkg, oracle, N=12, lmax=4, seed=8, side=3.5)
    a = cl.plane_wave([0, 0, 1], 1)
    x, conv = cl.solve(a, rtol=1e-10, restart=50, mixed_precision=True, convergence="full")
    xr = ref.solve(a.cpu().numpy())
    assert conv[0].residual_norm <= 1e-10 * conv[0].rhs_norm
    assert conv[0].refinements >= 1
    assert _rel(x, xr) <= 1e-8


@pytest.mark.parametrize("pc", ["block_jacobi"])
def test_preconditioned_solve_same_answer(pkg, oracle, pc):
    cl, ref = _cluster(pkg, oracle, N=16, lmax=3, seed=11, side=3.0)
    a = cl.plane_wave([0, 0, 1], 1)
    x0, c0 = cl.solve(a, rtol=1e-10)
    x1, c1 = cl.solve(a, rtol=1e-10, preconditioner=pc)
    assert _rel(x1, x0) <= 1e-8
    assert c1[0][1] <= 1e-10 * c1[0][2]
    assert c1[0][0] <= c0[0][0]


def test_block_jacobi_ragged_groups(pkg, oracle):
    """11 spheres in groups of at most 4 (sizes differ): the batched LU factors
    matrices of different dimensions in one set of launches; the preconditioned
    solve matches the dense solution, for the forward and the adjoint system."""
    cl, ref = _cluster(pkg, oracle, N=11, lmax=3, seed=17, side=3.2)
    cl.set_group_size(4)
    a = cl.plane_wave([0, 0, 1], 1)
    x, conv = cl.solve(a, rtol=1e-12, preconditioner="block_jacobi")
    assert _rel(x, ref.solve(a.cpu().numpy())) <= 1e-10
    g = np.random.default_rng(5).normal(size=ref.n) + 0j
    y, _ = cl.solve(g, rtol=1e-12, preconditioner="block_jacobi", adjoint=True)
    assert _rel(y, np.linalg.solve(ref.operator().conj().T, g)) <= 1e-10


def test_not_converged_raises(pkg, oracle):
    cl, _ = _cluster(pkg, oracle, N=8, lmax=3, seed=5, side=2.0)
    a = cl.plane_wave([0, 0, 1], 1)
    with pytest.raises(pkg.NotConverged):
        cl.solve(a, rtol=1e-14, max_iterations=2)


def test_invalid_input_rejected(pkg):
    with pytest.raises(ValueError):
        pkg.IterativeSphereCluster(3, 1.3, [0.3, 0.3], 4.0, [[0, 0, 0], [0.5, 0, 0]])  # overlapping
    with pytest.raises(ValueError):
        pkg.IterativeSphereCluster(3, -1.0, [0.3], 4.0, [[0, 0, 0]])


def _example_cluster(pkg, golden_dir):
    g = json.load(open(os.path.join(golden_dir, "reference_examples.json")))["cluster_example"]
    pos = (220 / np.sqrt(24)) * np.array(
        [[np.sqrt(8), 0, -1], [-np.sqrt(2), np.sqrt(6), -1], [-np.sqrt(2), -np.sqrt(6), -1], [0, 0, 3]])
    k0 = 2 * np.pi / g["k0_over_2pi_inverse"]
    return pkg.IterativeSphereCluster(g["lmax"], k0, g["radii"], complex(*g["epsilon"]), pos), g


def test_cluster_example_golden(pkg, golden_dir):
    cl, g = _example_cluster(pkg, golden_dir)
    a = cl.plane_wave(g["direction"], g["pol"])
    x, _ = cl.solve(a, rtol=1e-12)
    sca, ext = cl.cross_sections(a, x)
    e = g["expected"]
    assert float(f"{sca:.6g}") == e["scattering"]
    assert float(f"{ext:.6g}") == e["extinction"]
    assert float(f"{ext - sca:.6g}") == e["absorption"]


def test_reverse_gradient_golden(pkg, oracle, golden_dir):
    g = json.load(open(os.path.join(golden_dir, "reference_examples.json")))["reverse_gradient_example"]
    idx = np.arange(8)
    positions = np.column_stack((180 * (idx % 2), 190 * ((idx // 2) % 2), 200 * (idx // 4))).astype(float)
    positions += np.column_stack((9 * np.sin(idx), 7 * np.cos(idx), 11 * np.sin(2 * idx)))
    positions -= positions.mean(axis=0)
    radii = 45.0 + 3 * (idx % 4)
    cl = pkg.IterativeSphereCluster(g["lmax"], 2 * np.pi / 700, radii, complex(*g["epsilon"]), positions)
    sca, gr, gp, _ = pkg.scattering_gradient(cl, g["direction"], g["pol"], rtol=1e-13)
    assert round(sca, 8) == g["expected_scattering"]
    expected = np.array(g["expected_gradient_radius_xyz"])
    assert np.abs(np.round(np.column_stack((gr, gp)), 6) - expected).max() <= 1.5e-6


def test_pullback_matches_oracle(pkg, oracle):
    cl, ref = _cluster(pkg, oracle, N=6, lmax=3, seed=13, side=2.5, eps=None)
    rng = np.random.default_rng(2)
    a = np.stack([ref.plane_wave([0, 0, 1], 1), ref.plane_wave([1, 0, 0], 0)], axis=1)
    x, ctx, _ = cl.record(a, rtol=1e-13)
    G = rng.normal(size=a.shape) + 1j * rng.normal(size=a.shape)
    gk, gr, ge, gp, ginc, conv = ctx.pullback(G)
    xr = np.stack([ref.solve(a[:, c]) for c in range(2)], axis=1)
    r = ref.pullback(a, xr, G)
    assert np.abs(gr - r["radii"]).max() <= 1e-9 * np.abs(r["radii"]).max()
    assert np.abs(ge - r["eps"]).max() <= 1e-9 * np.abs(r["eps"]).max()
    assert np.abs(gp - r["positions"]).max() <= 1e-9 * np.abs(r["positions"]).max()
    assert abs(gk - r["k0"]) <= 1e-9 * abs(r["k0"])
    assert _rel(ginc, r["incident"]) <= 1e-9
    # zero total position gradient (translation invariance, iterative.rs:832-835)
    assert np.abs(gp.sum(axis=0)).max() <= 1e-9 * np.abs(gp).max()


def test_pullback_many_columns_chunked(pkg, oracle):
    """20 illuminations at lmax 6: the pair-gradient kernel takes the columns in LDS-sized
    chunks and accumulates them in a fixed order; the result matches the oracle."""
    cl, ref = _cluster(pkg, oracle, N=3, lmax=6, seed=3, side=1.8, eps=None)
    rng = np.random.default_rng(4)
    a = rng.normal(size=(ref.n, 20)) + 1j * rng.normal(size=(ref.n, 20))
    f = pkg.SphereClusterFactor(cl, "c128")
    x, ctx, _ = f.record(a)
    G = rng.normal(size=a.shape) + 1j * rng.normal(size=a.shape)
    gk, gr, ge, gp, ginc, _ = ctx.pullback(G)
    r = ref.pullback(a, np.linalg.solve(ref.operator(), oracle.apply_local(ref.blocks, a)), G)
    assert np.abs(gp - r["positions"]).max() <= 1e-9 * np.abs(r["positions"]).max()
    assert np.abs(gr - r["radii"]).max() <= 1e-9 * np.abs(r["radii"]).max()
    assert abs(gk - r["k0"]) <= 1e-9 * abs(r["k0"])


def _example(pkg, oracle, golden_dir):
    g = json.load(open(os.path.join(golden_dir, "reference_examples.json")))["cluster_example"]
    pos = (220 / np.sqrt(24)) * np.array(
        [[np.sqrt(8), 0, -1], [-np.sqrt(2), np.sqrt(6), -1], [-np.sqrt(2), -np.sqrt(6), -1], [0, 0, 3]])
    k0 = 2 * np.pi / g["k0_over_2pi_inverse"]
    eps = np.full(4, complex(*g["epsilon"]))
    cl = pkg.IterativeSphereCluster(g["lmax"], k0, np.array(g["radii"]), eps, pos)
    return cl, oracle.Cluster(g["lmax"], k0, g["radii"], complex(*g["epsilon"]), pos), g


def test_fields_match_oracle(pkg, oracle, golden_dir):
    """Outgoing-wave field of a solved cluster (3 columns) and the plane-wave field at
    points around and between the spheres, against the oracle (1e-12)."""
    cl, ref, g = _example(pkg, oracle, golden_dir)
    rng = np.random.default_rng(2)
    pts = rng.uniform(-400, 400, size=(300, 3))
    pts = pts[[all(np.linalg.norm(p - ref.positions[i]) > 1e-3 for i in range(4)) for p in pts]]
    x = rng.normal(size=(ref.n, 3)) + 1j * rng.normal(size=(ref.n, 3))
    got = cl.field(x, pts).cpu().numpy()
    for c in range(3):
        assert _rel(got[:, :, c], oracle.cluster_field(ref.lmax, ref.k0, ref.positions, x[:, c], pts)) <= 1e-12
    for d, p in (([1, 0, 0], 0), ([0.3, -0.2, 1], 1), ([0, 0, -1], 1)):
        assert _rel(pkg.plane_wave_field(ref.k0, d, p, pts), oracle.plane_wave_field(ref.k0, d, p, pts)) <= 1e-13


def test_regular_field_reproduces_plane_wave(pkg):
    """Regular waves against a closed form: a plane wave's regular expansion about a
    centre (degrees to 10), evaluated as a regular field, is the plane wave itself at
    k |r - c| <= 1 (truncation below 1e-10), including the centre (the series of
    j_l(kr) / r^l) and points on both sides of k r = 1/2."""
    k0, c = 1.3, np.array([0.4, -0.2, 0.7])
    cl = pkg.IterativeSphereCluster(10, k0, [0.2], 4 + 0.1j, [c])
    rng = np.random.default_rng(5)
    u = rng.normal(size=(40, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    rad = np.concatenate([[0.0], rng.uniform(0.0, 1.0, size=39)]) / k0
    pts = c + u * rad[:, None]
    for d, pol in (([0, 0, 1], 1), ([0.3, -0.5, 0.8], 0), ([-1, 0.2, 0.1], 1)):
        a = cl.plane_wave(d, pol)
        got = cl.field(a, pts, regular=True).cpu().numpy()
        want = pkg.plane_wave_field(k0, d, pol, pts).cpu().numpy()
        assert np.isfinite(got).all()
        err = np.abs(got - want).max() / np.abs(want).max()
        print("regular field vs plane wave", d, pol, f"{err:.2e}")
        assert err <= 1e-9, (d, pol, err)


def test_expansion_field_matches_cluster_field(pkg):
    """One expansion about a centre (tg_expansion_field) against a one-sphere cluster
    there: the same waves, outgoing and regular, several columns; and the scattered
    field of a two-sphere cluster's global T-matrix (degree 12 about the origin)
    against the cluster's multi-centre scattered field outside the circumscribing
    sphere."""
    k0, c, lmax = 1.1, np.array([0.3, 0.1, -0.4]), 5
    cl = pkg.IterativeSphereCluster(lmax, k0, [0.3], 3 + 0.2j, [c])
    rng = np.random.default_rng(7)
    x = rng.normal(size=(cl.n, 3)) + 1j * rng.normal(size=(cl.n, 3))
    pts = c + rng.normal(size=(50, 3)) * 2.0
    for regular in (False, True):
        a = cl.field(x, pts, regular=regular)
        b = pkg.expansion_field(lmax, k0, x, pts, regular=regular, center=c)
        assert _rel(b, a) <= 1e-15, regular
    # global expansion of a solved two-sphere cluster
    pos = np.array([[0, 0, -0.6], [0, 0, 0.6]])
    two = pkg.IterativeSphereCluster(4, k0, [0.3, 0.35], 4 + 0.1j, pos)
    f = pkg.SphereClusterFactor(two, "c128")
    ainc = two.plane_wave([0.2, 0.1, 1.0], 1)
    xs, _ = f.solve(ainc)
    Tg, _ = f.global_tmatrix(12)
    f.close()
    one = pkg.IterativeSphereCluster(12, k0, [0.1], 4 + 0.1j, [[0, 0, 0]])
    ag = one.plane_wave([0.2, 0.1, 1.0], 1)
    pg = torch.as_tensor(Tg).to(ag.device) @ ag
    u = rng.normal(size=(30, 3))
    far = u / np.linalg.norm(u, axis=1, keepdims=True) * 4.0
    got = pkg.expansion_field(12, k0, pg, far).cpu().numpy()
    want = two.field(xs, far).cpu().numpy()
    err = np.abs(got - want).max() / np.abs(want).max()
    print("global expansion field", f"{err:.2e}")
    assert err <= 1e-8  # measured 3.7e-10


def test_cluster_example_intensity_map_golden(pkg, oracle, golden_dir):
    """docs/examples/cluster.py: total-field intensity 0.5 |E_inc + E_sca|^2 on the 7 x 7
    grid at y = 0 from the GPU solve and field kernels, against
    docs/examples/output/cluster.txt at its 6 printed digits (points inside a sphere
    are NaN there and skipped)."""
    cl, ref, g = _example(pkg, oracle, golden_dir)
    a = cl.plane_wave(g["direction"], g["pol"])
    x, _ = cl.solve(a, rtol=1e-12)
    gx = np.linspace(-300, 300, 7)
    pts = np.array([[u, 0.0, w] for u in gx for w in gx])
    inside = np.array([any(np.linalg.norm(p - ref.positions[i]) <= ref.radii[i] for i in range(4)) for p in pts])
    out = pts[~inside]
    e = pkg.plane_wave_field(ref.k0, g["direction"], g["pol"], out) + cl.field(x, out)
    got = 0.5 * torch.sum(torch.abs(e) ** 2, dim=1).cpu().numpy()
    want = [v for row in g["intensity_local"]["values"] for v in row if v is not None]
    assert [float(f"{v:.6g}") for v in got] == want


@pytest.mark.parametrize("ld,ls", [(3, 3), (6, 3), (3, 6), (2, 5)])
@pytest.mark.parametrize("regular", [False, True])
def test_translation_blocks_rect(pkg, ld, ls, regular):
    """Blocks between bases of different degree: the modes of degree <= min(ld, ls)
    come first in both bases, so that corner equals the square block of that degree
    (the selection rule and weights do not depend on lmax); regular blocks at the
    origin are [I 0; 0 0]."""
    rng = np.random.default_rng(ld * 10 + ls)
    d = rng.normal(size=(5, 3)) * 1.7
    k = 1.3
    rect = pkg.translation_blocks_rect(ld, ls, k, d, regular).cpu().numpy()
    lo = min(ld, ls)
    m = 2 * lo * (lo + 2)
    sq = pkg.translation_blocks(lo, k, d, regular).cpu().numpy()
    assert rect.shape == (5, 2 * ld * (ld + 2), 2 * ls * (ls + 2))
    assert _rel(rect[:, :m, :m], sq) <= 1e-13
    if ld == ls:
        assert _rel(rect, sq) <= 1e-14
    if regular:
        o = pkg.translation_blocks_rect(ld, ls, k, np.zeros((1, 3)), True).cpu().numpy()[0]
        eye = np.zeros_like(o)
        eye[:m, :m] = np.eye(m)
        assert np.abs(o - eye).max() <= 1e-15


def test_cluster_example_global_cross_sections_golden(pkg, oracle, golden_dir):
    """docs/examples/cluster.py: the solved cluster expanded about the origin with
    lmax 6 (spheres at lmax 3; rectangular regular translations), its cross sections
    for the same plane wave against the reference's printed "global" row."""
    cl, ref, g = _example(pkg, oracle, golden_dir)
    f = pkg.SphereClusterFactor(cl, "c128")
    Tg, conv = f.global_tmatrix(6)
    assert Tg.shape == (96, 96)
    # a one-sphere basis at the origin gives the plane wave's coefficients and the
    # cross sections of a single-origin T-matrix (its regular overlap is I)
    one = pkg.IterativeSphereCluster(6, ref.k0, np.array([1.0]), np.array([2.0 + 0j]), np.zeros((1, 3)))
    a = one.plane_wave(g["direction"], g["pol"])
    sca, ext = one.cross_sections(a, Tg @ a)
    e = g["expected_global"]
    assert float(f"{sca:.6g}") == e["scattering"]
    assert float(f"{ext:.6g}") == e["extinction"]
    assert float(f"{ext - sca:.6g}") == e["absorption"]
    f.close()


def test_cluster_example_global_intensity_map_golden(pkg, oracle, golden_dir):
    """docs/examples/cluster.py: total-field intensity 0.5 |E_inc + E_sca|^2 on the 7 x 7
    grid at y = 0 from the expansion about the origin (lmax 6): the scattered field of
    T_g a, one outgoing expansion at the origin (tg_expansion_field), against
    docs/examples/output/cluster.txt's "global" map at its 6 printed digits (points
    within radius 260 nm, where the expansion does not hold, are NaN there and
    skipped)."""
    cl, ref, g = _example(pkg, oracle, golden_dir)
    f = pkg.SphereClusterFactor(cl, "c128")
    Tg, _ = f.global_tmatrix(6)
    f.close()
    one = pkg.IterativeSphereCluster(6, ref.k0, np.array([1.0]), np.array([2.0 + 0j]), np.zeros((1, 3)))
    a = one.plane_wave(g["direction"], g["pol"])
    gx = np.linspace(-300, 300, 7)
    pts = np.array([[u, 0.0, w] for u in gx for w in gx])
    want = np.array([v for row in g["intensity_global"]["values"] for v in row], dtype=object)
    keep = np.array([v is not None for v in want])
    out = pts[keep]
    e = pkg.plane_wave_field(ref.k0, g["direction"], g["pol"], out) + pkg.expansion_field(6, ref.k0, Tg @ a, out)
    got = 0.5 * torch.sum(torch.abs(e) ** 2, dim=1).cpu().numpy()
    assert [float(f"{v:.6g}") for v in got] == [float(v) for v in want[keep]]


def test_cluster_example_rotated_golden(pkg, oracle, golden_dir):
    """docs/examples/cluster.py, third row: the expansion about the origin (lmax 6)
    rotated by z-y-z Euler angles (0, pi/2, 0) (R T R^H with Wigner D blocks) and lit
    along -z reproduces the first illumination in the rotated frame: its cross
    sections and its 7 x 7 intensity map against docs/examples/output/cluster.txt's
    "rotated" row and map at 6 printed digits."""
    cl, ref, g = _example(pkg, oracle, golden_dir)
    f = pkg.SphereClusterFactor(cl, "c128")
    Tg, _ = f.global_tmatrix(6)
    f.close()
    e = g["expected_rotated"]
    Tr = pkg.rotate_tmatrix(Tg, 6, *e["euler"])
    one = pkg.IterativeSphereCluster(6, ref.k0, np.array([1.0]), np.array([2.0 + 0j]), np.zeros((1, 3)))
    a = one.plane_wave(e["direction"], g["pol"])
    p = Tr @ a
    sca, ext = one.cross_sections(a, p)
    assert float(f"{sca:.6g}") == e["scattering"]
    assert float(f"{ext:.6g}") == e["extinction"]
    assert float(f"{ext - sca:.6g}") == e["absorption"]
    gx = np.linspace(-300, 300, 7)
    pts = np.array([[u, 0.0, w] for u in gx for w in gx])
    want = np.array([v for row in g["intensity_rotated"]["values"] for v in row], dtype=object)
    keep = np.array([v is not None for v in want])
    out = pts[keep]
    fld = pkg.plane_wave_field(ref.k0, e["direction"], g["pol"], out) + pkg.expansion_field(6, ref.k0, p, out)
    got = 0.5 * torch.sum(torch.abs(fld) ** 2, dim=1).cpu().numpy()
    assert [float(f"{v:.6g}") for v in got] == [float(v) for v in want[keep]]
