"""FP64 CPU oracle for the cluster-coupling path -- test infrastructure only.

This module is the checker, never the product: only ``tests/``, the smoke check of
``__graft_entry__`` and the CPU-baseline leg of ``bench.py`` import it.  It loads
``liboracle.so`` (built from ``treams_oracle.c`` by ``oracle/build.py``) and adds the
dense linear algebra and the cluster-level algebra in NumPy.

Each function cites the reference it restates (paths relative to
``crates/treams-core/src/`` of treams-rs unless they start with ``python/``).
Parity pins of every function are listed in ``tests/test_oracle_golden.py``.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_c128p = np.ctypeslib.ndpointer(dtype=np.complex128, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")


class _C128(ctypes.Structure):
    _fields_ = [("re", ctypes.c_double), ("im", ctypes.c_double)]


def lib():
    """The compiled oracle; build it with ``python oracle/build.py`` first."""
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run oracle/build.py)")
        L = ctypes.CDLL(_LIB_PATH)
        L.or_wigner3j.restype = ctypes.c_double
        L.or_wigner3j.argtypes = [ctypes.c_int] * 6
        L.or_modes.restype = ctypes.c_int
        L.or_modes.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.or_degrees.restype = ctypes.c_int
        L.or_degrees.argtypes = [ctypes.c_int] * 5 + [ctypes.c_void_p]
        L.or_plan.restype = ctypes.c_int
        L.or_plan.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.or_hankels.argtypes = [ctypes.c_int, ctypes.c_double, _c128p]
        L.or_spherical_jn.argtypes = [ctypes.c_int, _C128, _c128p]
        L.or_radial_jets.argtypes = [ctypes.c_int, _C128, ctypes.c_int, _c128p]
        L.or_solid_table.argtypes = [ctypes.c_int, _f64p, _c128p, ctypes.c_void_p]
        L.or_translation_block.argtypes = [ctypes.c_int, ctypes.c_double, _f64p, ctypes.c_int, _c128p]
        L.or_translation_pullback.restype = ctypes.c_int
        L.or_translation_pullback.argtypes = [ctypes.c_int, ctypes.c_double, _f64p, ctypes.c_int, _c128p,
                                              _f64p, _c128p]
        L.or_pair_matrix.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, _f64p, ctypes.c_int, _c128p]
        L.or_pair_matrix_pullback.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, _f64p, ctypes.c_int,
                                              _c128p, _f64p, _f64p]
        L.or_coupling_apply.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, _f64p, ctypes.c_int, _c128p,
                                        _c128p, ctypes.c_int, ctypes.c_int]
        L.or_pair_apply.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, _f64p, ctypes.c_int, _c128p,
                                    _c128p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.or_mie.restype = ctypes.c_int
        L.or_mie.argtypes = [ctypes.c_int, ctypes.c_double, _C128, _c128p]
        L.or_mie_pullback.restype = ctypes.c_int
        L.or_mie_pullback.argtypes = [ctypes.c_int, ctypes.c_double, _C128, _c128p, _f64p, _c128p]
        L.or_plane_wave.argtypes = [ctypes.c_int, ctypes.c_int, _f64p, ctypes.c_double, _f64p, ctypes.c_int,
                                    _c128p]
        _lib = L
    return _lib


def _c(z) -> _C128:
    z = complex(z)
    return _C128(z.real, z.imag)


# ------------------------------------------------------------------ special functions


def wigner3j(j1, j2, j3, m1, m2, m3) -> float:
    """special/wigner.rs:263-279 (two-sided recurrence)."""
    return lib().or_wigner3j(j1, j2, j3, m1, m2, m3)


def modes(lmax: int):
    """sw/mod.rs:114-125: (l, m, pol) arrays in mode order."""
    M = 2 * lmax * (lmax + 2)
    l = np.zeros(M, np.int32)
    m = np.zeros(M, np.int32)
    pol = np.zeros(M, np.int32)
    lib().or_modes(lmax, l.ctypes.data, m.ctypes.data, pol.ctypes.data)
    return l, m, pol


def degrees(l, m, lam, mu, cross) -> list[int]:
    """sw/coupling.rs:146-160."""
    out = np.zeros(600, np.int32)
    n = lib().or_degrees(l, m, lam, mu, int(cross), out.ctypes.data)
    return [int(v) for v in out[:n]]


def plan(lmax: int):
    """CSR plan of sw/plan.rs:71-170 (helicity): (starts, index, weight)."""
    M = 2 * lmax * (lmax + 2)
    starts = np.zeros(M * M + 1, np.int64)
    total = lib().or_plan(lmax, starts.ctypes.data, None, None)
    index = np.zeros(max(total, 1), np.int32)
    weight = np.zeros(max(total, 1), np.complex128)
    lib().or_plan(lmax, starts.ctypes.data, index.ctypes.data, weight.ctypes.data)
    return starts, index[:total], weight[:total]


def hankels(last: int, x: float) -> np.ndarray:
    """Outgoing h_p(x), p=0..last, real x (special/bessel.rs:495-516)."""
    out = np.zeros(last + 1, np.complex128)
    lib().or_hankels(last, float(x), out)
    return out


def spherical_jn(last: int, z) -> np.ndarray:
    """Regular j_p(z), p=0..last (series of special/bessel.rs:425-431; ratio recurrence)."""
    out = np.zeros(last + 1, np.complex128)
    lib().or_spherical_jn(last, _c(z), out)
    return out


def radial_jets(last: int, z, regular: bool) -> np.ndarray:
    """(last+1, 3) jets (value, first, second), special/bessel.rs:453-487."""
    out = np.zeros(3 * (last + 1), np.complex128)
    lib().or_radial_jets(last, _c(z), int(regular), out)
    return out.reshape(last + 1, 3)


def solid_table(order: int, position, gradient: bool = False):
    """special/harmonics.rs:147-305: values (and gradients) at index p^2+p+m."""
    n = (order + 1) ** 2
    pos = np.ascontiguousarray(position, np.float64)
    val = np.zeros(n, np.complex128)
    if gradient:
        grad = np.zeros(3 * n, np.complex128)
        lib().or_solid_table(order, pos, val, grad.ctypes.data)
        return val, grad.reshape(n, 3)
    lib().or_solid_table(order, pos, val, None)
    return val


# ---------------------------------------------------------------- translations


def translation_block(lmax: int, k: float, d, regular: bool = False) -> np.ndarray:
    """One M x M block (column-major memory -> returned as [dest, source]), sw/plan.rs:282-289."""
    M = 2 * lmax * (lmax + 2)
    out = np.zeros(M * M, np.complex128)
    lib().or_translation_block(lmax, float(k), np.ascontiguousarray(d, np.float64), int(regular), out)
    return out.reshape(M, M).T  # memory is column-major: [source][dest]


def translation_pullback(lmax: int, k: float, d, regular: bool, cotangent: np.ndarray):
    """sw/plan.rs:596-615; cotangent indexed [dest, source]. Returns (grad[3], gk)."""
    cot = np.ascontiguousarray(np.asarray(cotangent, np.complex128).T).reshape(-1)
    grad = np.zeros(3)
    gk = np.zeros(1, np.complex128)
    rc = lib().or_translation_pullback(lmax, float(k), np.ascontiguousarray(d, np.float64), int(regular), cot,
                                       grad, gk)
    if rc:
        raise ValueError("translation pullback at the origin")
    return grad, complex(gk[0])


def pair_matrix(lmax: int, k0: float, positions, regular: bool) -> np.ndarray:
    """Dense C (singular, self blocks 0; cluster/spheres.rs:174-220) or R (regular)."""
    pos = np.ascontiguousarray(positions, np.float64).reshape(-1, 3)
    N = pos.shape[0]
    n = N * 2 * lmax * (lmax + 2)
    out = np.zeros(n * n, np.complex128)
    lib().or_pair_matrix(lmax, float(k0), N, pos.reshape(-1), int(regular), out)
    return out.reshape(n, n).T


def pair_matrix_pullback(lmax: int, k0: float, positions, regular: bool, gbar: np.ndarray):
    """Position / k0 gradients of a pair matrix from its cotangent (cluster/iterative.rs:500-551)."""
    pos = np.ascontiguousarray(positions, np.float64).reshape(-1, 3)
    N = pos.shape[0]
    g = np.ascontiguousarray(np.asarray(gbar, np.complex128).T).reshape(-1)
    grad = np.zeros(3 * N)
    gk = np.zeros(1)
    lib().or_pair_matrix_pullback(lmax, float(k0), N, pos.reshape(-1), int(regular), g, grad, gk)
    return grad.reshape(N, 3), float(gk[0])


def coupling_apply(lmax: int, k0: float, positions, x: np.ndarray, rows=None, regular: bool = False) -> np.ndarray:
    """Matrix-free C x (cluster/iterative.rs:187-221); x is (n,) or (n, cols).  With
    rows = (i0, i1) only destination particles i0..i1-1 are computed (others zero).
    regular: the regular overlap R x instead (identity self blocks)."""
    pos = np.ascontiguousarray(positions, np.float64).reshape(-1, 3)
    N = pos.shape[0]
    vec = x.ndim == 1
    X = np.ascontiguousarray((x[:, None] if vec else x).T, np.complex128)  # column-major
    cols = X.shape[0]
    Y = np.zeros_like(X)
    i0, i1 = rows if rows is not None else (0, N)
    lib().or_pair_apply(lmax, float(k0), N, pos.reshape(-1), cols, X.reshape(-1), Y.reshape(-1), i0, i1,
                        int(regular))
    Y = Y.T
    return Y[:, 0] if vec else Y


# ------------------------------------------------------------------------- Mie


def mie(l: int, size: float, eps) -> np.ndarray:
    """2x2 Mie matrix, helicity order (negative, positive): coeffs/mie.rs:209-255."""
    out = np.zeros(4, np.complex128)
    if lib().or_mie(l, float(size), _c(eps), out):
        raise ValueError("singular Mie matrix")
    return out.reshape(2, 2)


def mie_pullback(l: int, size: float, eps, cotangent_helicity: np.ndarray):
    """coeffs/mie.rs:306-349 for one boundary: (g_size, g_eps)."""
    cot = np.ascontiguousarray(cotangent_helicity, np.complex128).reshape(-1)
    gs = np.zeros(1)
    ge = np.zeros(1, np.complex128)
    if lib().or_mie_pullback(l, float(size), _c(eps), cot, gs, ge):
        raise ValueError("singular Mie matrix")
    return float(gs[0]), complex(ge[0])


def to_mode_order(m2: np.ndarray) -> np.ndarray:
    """coeffs/mie.rs:123-130: swap both axes (pol 1 first)."""
    return m2[::-1, ::-1].copy()


def tmatrix_blocks(lmax: int, k0: float, radii, eps) -> np.ndarray:
    """(N, lmax, 2, 2) Mie blocks in mode order for each sphere and degree
    (cluster/iterative.rs:66-75; tmatrix/sphere.rs:43-71)."""
    radii = np.asarray(radii, np.float64)
    eps = np.broadcast_to(np.asarray(eps, np.complex128), radii.shape)
    out = np.zeros((radii.size, lmax, 2, 2), np.complex128)
    for i, (r, e) in enumerate(zip(radii, eps)):
        for l in range(1, lmax + 1):
            out[i, l - 1] = to_mode_order(mie(l, k0 * r, e))
    return out


def block_degrees(lmax: int) -> np.ndarray:
    """tmatrix/sphere.rs:76-78: degree index of each 2x2 block (one per (l, m))."""
    return np.concatenate([np.full(2 * l + 1, l - 1) for l in range(1, lmax + 1)])


def tmatrix_dense(blocks: np.ndarray) -> np.ndarray:
    """Block-diagonal dense T (n x n) from (N, lmax, 2, 2) mode-order blocks."""
    N, lmax = blocks.shape[:2]
    deg = block_degrees(lmax)
    M = 2 * deg.size
    T = np.zeros((N * M, N * M), np.complex128)
    for i in range(N):
        for b, d in enumerate(deg):
            o = i * M + 2 * b
            T[o:o + 2, o:o + 2] = blocks[i, d]
    return T


def apply_local(blocks: np.ndarray, x: np.ndarray, adjoint: bool = False) -> np.ndarray:
    """T x or T^H x with 2x2 blocks per (particle, l, m): cluster/iterative.rs:99-136."""
    N, lmax = blocks.shape[:2]
    deg = block_degrees(lmax)
    M = 2 * deg.size
    B = blocks[:, deg]  # (N, nb, 2, 2)
    if adjoint:
        B = np.conj(np.swapaxes(B, -1, -2))
    xs = x.reshape(N, deg.size, 2, -1)
    y = np.einsum("nbij,nbjc->nbic", B, xs)
    return y.reshape(x.shape)


# ------------------------------------------------------------------- plane wave


def plane_wave(lmax: int, positions, k0: float, direction, pol: int) -> np.ndarray:
    """Local regular coefficients of a unit plane wave (pw/expand.rs:23-66, pw/field.rs:20-26)."""
    pos = np.ascontiguousarray(positions, np.float64).reshape(-1, 3)
    N = pos.shape[0]
    M = 2 * lmax * (lmax + 2)
    out = np.zeros(N * M, np.complex128)
    lib().or_plane_wave(lmax, N, pos.reshape(-1), float(k0), np.ascontiguousarray(direction, np.float64), int(pol),
                        out)
    return out


# ------------------------------------------------------------------------ fields


def spherical_waves(lmax: int, k: float, r, regular: bool = False) -> np.ndarray:
    """(M, 3) Cartesian electric fields of one sphere's helicity vector spherical waves
    (sphere mode order) at displacement r = point - centre (fields.rs:144-320):
    psi = c S with c = z_l(k r) / r^l and the solid harmonic S = r^l P_l^m e^{i m phi};
    n = d grad S + l e S r (d = (l + 1) c / k - r^2 e, e = z_{l+1} / r^{l+1}),
    m = c (r x grad S); wave = i sqrt((2l+1) / (4 pi l (l+1))) sqrt((l-m)!/(l+m)!)
    (n - (2 pol - 1) m) / sqrt(2).  Coordinates are scaled by the largest component
    (and by 1/k for regular waves), which leaves the waves unchanged."""
    r = np.asarray(r, np.float64)
    scale = np.abs(r).max()
    if regular or scale == 0.0:
        scale = max(scale, 1.0 / abs(k))
    ks, rs = k * scale, r / scale
    r2 = float(rs @ rs)
    rr = np.sqrt(r2)
    val, grad = solid_table(lmax, rs, gradient=True)
    if regular:
        z = spherical_jn(lmax + 1, ks * rr) if rr > 0 else np.array([1.0] + [0.0] * (lmax + 1), np.complex128)
        rad = np.array([z[l] / rr ** l if rr > 0 else 0.0 for l in range(lmax + 2)], np.complex128)
        if rr == 0 or abs(ks * rr) < 0.5:  # the origin series (fields.rs:146-160)
            rad = np.zeros(lmax + 2, np.complex128)
            for l in range(lmax + 2):
                den = np.prod([2.0 * i + 1 for i in range(l + 1)])
                coef, tot = ks ** l / den, 0j
                for q in range(12):
                    tot += coef * (ks * rr) ** (2 * q)
                    coef /= -2.0 * (q + 1) * (2 * l + 2 * q + 3)
                rad[l] = tot
    else:
        h = hankels(lmax + 1, ks * rr)
        rad = np.array([h[l] / rr ** l for l in range(lmax + 2)], np.complex128)
    M = 2 * lmax * (lmax + 2)
    out = np.zeros((M, 3), np.complex128)
    for l in range(1, lmax + 1):
        c, e = rad[l], rad[l + 1]
        d = (l + 1) / ks * c - r2 * e
        for m in range(-l, l + 1):
            S, G = val[l * l + l + m], grad[l * l + l + m]
            n = d * G + l * e * S * rs
            mv = c * np.cross(rs, G)
            norm = 1j * np.sqrt((2 * l + 1) / (4 * np.pi * l * (l + 1))) * np.sqrt(
                float(mp_factorial_ratio(l, m)))
            for pol in (1, 0):
                out[2 * (l * l + l + m - 1) + (1 - pol)] = norm * (n - (2 * pol - 1) * mv) / np.sqrt(2)
    return out


def mp_factorial_ratio(l: int, m: int) -> float:
    """(l - m)! / (l + m)!"""
    from math import factorial
    return factorial(l - m) / factorial(l + m)


def cluster_field(lmax: int, k: float, positions, x, points, regular: bool = False) -> np.ndarray:
    """(P, 3) field sum_i sum_modes x[i M + mode] wave_mode(point - p_i)
    (fields.rs:947-1010 for a multi-centre helicity basis)."""
    pos = np.asarray(positions, np.float64).reshape(-1, 3)
    M = 2 * lmax * (lmax + 2)
    out = np.zeros((len(points), 3), np.complex128)
    for p, pt in enumerate(points):
        for i in range(pos.shape[0]):
            out[p] += x[i * M:(i + 1) * M] @ spherical_waves(lmax, k, np.asarray(pt) - pos[i], regular)
    return out


def plane_wave_field(k0: float, direction, pol: int, points) -> np.ndarray:
    """(P, 3) field of a unit helicity plane wave: (N + (2 pol - 1) M) / sqrt(2)
    exp(i k.r) with M = i (k_y, -k_x, 0) / k_t and N = (-k_x k_z, -k_y k_z, k_t^2) /
    (k k_t); on the z axis M = (0, -i, 0) and N = (-sign k_z, 0, 0)
    (pw/polarization.rs:138-203)."""
    d = np.asarray(direction, np.float64)
    kv = k0 * d / np.linalg.norm(d)
    kt = np.hypot(kv[0], kv[1])
    if kt == 0.0:
        Mv = np.array([0, -1j, 0])
        Nv = np.array([-np.sign(kv[2]), 0, 0], np.complex128)
    else:
        xy = kv[:2] / kt
        Mv = np.array([1j * xy[1], -1j * xy[0], 0])
        Nv = np.array([-xy[0] * kv[2] / k0, -xy[1] * kv[2] / k0, kt / k0], np.complex128)
    e = (Nv + (2 * pol - 1) * Mv) / np.sqrt(2)
    phase = np.exp(1j * (np.asarray(points, np.float64) @ kv))
    return phase[:, None] * e[None, :]


# ------------------------------------------------------------------------- GMRES


@dataclass
class Convergence:
    """linalg/gmres.rs:81-93."""

    iterations: int
    residual_norm: float
    rhs_norm: float


class NotConverged(RuntimeError):
    pass


def gmres(apply, rhs: np.ndarray, rtol=1e-10, atol=0.0, restart=30, max_iterations=300):
    """Restarted GMRES with two MGS passes, complex Givens rotations and true-residual
    certification, restating linalg/gmres.rs:119-347 for one column."""
    n = rhs.size
    restart = min(restart, n, max_iterations)
    rhs_norm = float(np.linalg.norm(rhs))
    tol = max(atol, rtol * rhs_norm)
    x = np.zeros(n, np.complex128)
    residual = rhs.copy()
    iterations = 0
    while True:
        rn = float(np.linalg.norm(residual))
        if np.isfinite(rn) and rn <= tol:
            return x, Convergence(iterations, rn, rhs_norm)
        if not np.isfinite(rn) or iterations >= max_iterations:
            raise NotConverged(f"GMRES did not converge after {iterations} iterations: residual {rn:e}")
        V = [residual / rn]
        H = np.zeros((restart + 1, restart), np.complex128)
        rot = []
        g = np.zeros(restart + 1, np.complex128)
        g[0] = rn
        used = 0
        for j in range(restart):
            w = apply(V[j])
            for _ in range(2):
                for i in range(j + 1):
                    dot = np.vdot(V[i], w)
                    H[i, j] += dot
                    w = w - dot * V[i]
            nxt = float(np.linalg.norm(w))
            H[j + 1, j] = nxt
            for i, (c, s) in enumerate(rot):
                a, b = H[i, j], H[i + 1, j]
                H[i, j] = c * a + s * b
                H[i + 1, j] = -np.conj(s) * a + c * b
            a, b = H[j, j], H[j + 1, j]
            length = float(np.hypot(abs(a), abs(b)))
            if not np.isfinite(length) or length == 0.0:
                raise NotConverged("GMRES Arnoldi breakdown before convergence")
            phase = 1.0 if abs(a) == 0 else a / abs(a)
            c = abs(a) / length
            s = phase * np.conj(b) / length
            H[j, j] = phase * length
            H[j + 1, j] = 0.0
            rot.append((c, s))
            g[j + 1] = -np.conj(s) * g[j]
            g[j] *= c
            used = j + 1
            iterations += 1
            if abs(g[j + 1]) <= tol or nxt == 0.0 or iterations == max_iterations:
                break
            V.append(w / nxt)
        for i in range(used - 1, -1, -1):
            tail = sum(H[i, k] * g[k] for k in range(i + 1, used))
            g[i] = (g[i] - tail) / H[i, i]
            x = x + g[i] * V[i]
        residual = rhs - apply(x)


class _Krylov:
    """One column's restarted Arnoldi state (linalg/gmres.rs:200-347): two MGS passes,
    complex Givens rotations, convergence only on the true residual."""

    def __init__(self, rhs, rtol, atol, max_iterations):
        self.rhs = rhs
        self.rhs_norm = float(np.linalg.norm(rhs))
        self.tol = max(atol, rtol * self.rhs_norm)
        self.max_iterations = max_iterations
        self.x = np.zeros_like(rhs)
        self.residual = rhs.copy()
        self.iterations = 0
        self.rn = self.rhs_norm

    def converged(self):
        self.rn = float(np.linalg.norm(self.residual))
        if np.isfinite(self.rn) and self.rn <= self.tol:
            return True
        if not np.isfinite(self.rn) or self.iterations >= self.max_iterations:
            raise NotConverged(f"GMRES did not converge after {self.iterations} iterations: residual {self.rn:e}")
        return False

    def start(self, restart):
        self.V = [self.residual / self.rn]
        self.H = np.zeros((restart + 1, restart), np.complex128)
        self.rot = []
        self.g = np.zeros(restart + 1, np.complex128)
        self.g[0] = self.rn
        self.used = 0

    def step(self, w, j):
        """Extend the basis with w = A V[j]; True while the cycle should go on."""
        H = self.H
        for _ in range(2):
            for i in range(j + 1):
                dot = np.vdot(self.V[i], w)
                H[i, j] += dot
                w = w - dot * self.V[i]
        nxt = float(np.linalg.norm(w))
        H[j + 1, j] = nxt
        for i, (c, s) in enumerate(self.rot):
            a, b = H[i, j], H[i + 1, j]
            H[i, j] = c * a + s * b
            H[i + 1, j] = -np.conj(s) * a + c * b
        a, b = H[j, j], H[j + 1, j]
        length = float(np.hypot(abs(a), abs(b)))
        if not np.isfinite(length) or length == 0.0:
            raise NotConverged("GMRES Arnoldi breakdown before convergence")
        phase = 1.0 if abs(a) == 0 else a / abs(a)
        c = abs(a) / length
        s = phase * np.conj(b) / length
        H[j, j] = phase * length
        H[j + 1, j] = 0.0
        self.rot.append((c, s))
        self.g[j + 1] = -np.conj(s) * self.g[j]
        self.g[j] *= c
        self.used = j + 1
        self.iterations += 1
        if abs(self.g[j + 1]) <= self.tol or nxt == 0.0 or self.iterations == self.max_iterations:
            return False
        self.V.append(w / nxt)
        return True

    def update(self):
        g, H = self.g, self.H
        for i in range(self.used - 1, -1, -1):
            tail = sum(H[i, k] * g[k] for k in range(i + 1, self.used))
            g[i] = (g[i] - tail) / H[i, i]
            self.x = self.x + g[i] * self.V[i]


def gmres_batch(apply_columns, rhs: np.ndarray, rtol=1e-10, atol=0.0, restart=30, max_iterations=300):
    """Independent GMRES solves of the columns of rhs (n x cols) that share operator
    applications, restating linalg/gmres.rs:119-185: each column keeps its own Krylov
    basis, rotations and true-residual certificate; apply_columns(X) multiplies the
    n x k block of the columns that still need that application."""
    n, cols = rhs.shape
    restart = min(restart, n, max_iterations)
    states = [_Krylov(rhs[:, c].copy(), rtol, atol, max_iterations) for c in range(cols)]
    while True:
        cycle = [c for c, s in enumerate(states) if not s.converged()]
        if not cycle:
            break
        for c in cycle:
            states[c].start(restart)
        extending = list(cycle)
        for j in range(restart):
            W = apply_columns(np.stack([states[c].V[j] for c in extending], axis=1))
            extending = [c for k, c in enumerate(extending) if states[c].step(W[:, k].copy(), j)]
            if not extending:
                break
        for c in cycle:
            states[c].update()
        AX = apply_columns(np.stack([states[c].x for c in cycle], axis=1))
        for k, c in enumerate(cycle):
            states[c].residual = states[c].rhs - AX[:, k]
    X = np.stack([s.x for s in states], axis=1)
    return X, [Convergence(s.iterations, s.rn, s.rhs_norm) for s in states]


class MatrixFreeCluster:
    """Matrix-free FP64 oracle of a sphere cluster too large for dense matrices
    (cluster/iterative.rs:53-300): (I - T C) x with the per-pair plan application
    (or_pair_apply), T from the Mie blocks, batched GMRES, and the cross sections with
    the regular overlap applied matrix-free (python/treams_rs/_framework_tmatrix.py:
    231-289)."""

    def __init__(self, lmax, k0, radii, eps, positions):
        self.lmax = int(lmax)
        self.k0 = float(k0)
        self.radii = np.asarray(radii, np.float64).copy()
        self.eps = np.broadcast_to(np.asarray(eps, np.complex128), self.radii.shape).copy()
        self.positions = np.asarray(positions, np.float64).reshape(-1, 3).copy()
        self.N = self.radii.size
        self.M = 2 * self.lmax * (self.lmax + 2)
        self.n = self.N * self.M
        self.blocks = tmatrix_blocks(self.lmax, self.k0, self.radii, self.eps)

    def coupling(self, x, rows=None):
        return coupling_apply(self.lmax, self.k0, self.positions, x, rows=rows)

    def overlap(self, x, rows=None):
        return coupling_apply(self.lmax, self.k0, self.positions, x, rows=rows, regular=True)

    def apply(self, x):
        """(I - T C) x, x of shape (n,) or (n, cols)."""
        return x - apply_local(self.blocks, self.coupling(x), False)

    def rhs(self, incident):
        return apply_local(self.blocks, incident, False)

    def solve(self, incident, rtol=1e-10, atol=0.0, restart=30, max_iterations=300):
        """X = (I - T C)^{-1} T A by batched GMRES (cluster/iterative.rs:264-284)."""
        A = incident.reshape(self.n, -1)
        return gmres_batch(self.apply, self.rhs(A), rtol, atol, restart, max_iterations)

    def cross_sections(self, incident, x, flux=0.5):
        """(sca, ext) per column, with weights 1/k^2 and R x matrix-free."""
        X = x.reshape(self.n, -1)
        A = incident.reshape(self.n, -1)
        w = 1.0 / self.k0 ** 2
        RX = self.overlap(X)
        sca = np.real(np.sum(np.conj(X) * RX, axis=0)) * w * 0.5 / flux
        ext = -np.real(np.sum(np.conj(A) * X, axis=0)) * w * 0.5 / flux
        return sca, ext


# ----------------------------------------------------------------------- cluster


class Cluster:
    """Dense FP64 oracle of a sphere cluster in vacuum (helicity basis).

    Restates cluster/spheres.rs:56-220 (T blocks, C, ``I - T C``),
    cluster/interaction.rs:207-364 (solve), cluster/iterative.rs:99-592 (matrix-free
    apply, pair and Mie gradients) and the cross sections of
    python/treams_rs/_tmatrix.py:404-440 / _framework_tmatrix.py:231-289.
    The dense solve uses LAPACK (numpy.linalg), an independent library.
    """

    def __init__(self, lmax, k0, radii, eps, positions):
        self.lmax = int(lmax)
        self.k0 = float(k0)
        self.radii = np.asarray(radii, np.float64).copy()
        self.eps = np.broadcast_to(np.asarray(eps, np.complex128), self.radii.shape).copy()
        self.positions = np.asarray(positions, np.float64).reshape(-1, 3).copy()
        self.N = self.radii.size
        self.M = 2 * self.lmax * (self.lmax + 2)
        self.n = self.N * self.M
        self.blocks = tmatrix_blocks(self.lmax, self.k0, self.radii, self.eps)
        self._C = None
        self._R = None

    @property
    def C(self):
        if self._C is None:
            self._C = pair_matrix(self.lmax, self.k0, self.positions, False)
        return self._C

    @property
    def R(self):
        if self._R is None:
            self._R = pair_matrix(self.lmax, self.k0, self.positions, True)
        return self._R

    @property
    def T(self):
        return tmatrix_dense(self.blocks)

    def operator(self):
        """A = I - T C (cluster/interaction.rs:207-218)."""
        return np.eye(self.n) - self.T @ self.C

    def apply(self, x, adjoint=False):
        """(I - T C) x or (I - C^H T^H) x (cluster/iterative.rs:225-232)."""
        if adjoint:
            return x - self.C.conj().T @ apply_local(self.blocks, x, True)
        return x - apply_local(self.blocks, self.C @ x, False)

    def plane_wave(self, direction, pol):
        return plane_wave(self.lmax, self.positions, self.k0, direction, pol)

    def solve(self, incident):
        """x = (I - T C)^{-1} T a (cluster/interaction.rs:263-266)."""
        return np.linalg.solve(self.operator(), apply_local(self.blocks, incident, False))

    def interaction_tmatrix(self, equilibrate=True):
        """X = (I - T C)^{-1} T, all n columns (cluster/spheres.rs:56-80), by LAPACK on the
        operator equilibrated as in linalg/mod.rs:137-236: when its entries span more
        than a factor 10 and a diagonal entry is below a tenth of the largest, rows
        and columns are scaled by the reciprocal of their largest magnitude
        (max(|re|, |im|)), the more disparate axis first."""
        A = self.operator()
        rhs = self.T
        if not equilibrate:
            return np.linalg.solve(A, rhs)
        mag = np.maximum(np.abs(A.real), np.abs(A.imag))
        big = mag.max()
        nz = mag[mag > 0]
        if nz.min() / big >= 0.1 or np.all(np.diag(mag) / big >= 0.1):
            return np.linalg.solve(A, rhs)
        rmax, cmax = mag.max(axis=1), mag.max(axis=0)
        if rmax.min() / big >= 0.1 and cmax.min() / big >= 0.1:
            return np.linalg.solve(A, rhs)
        if cmax.min() < rmax.min():
            c = 1.0 / cmax
            r = 1.0 / (mag * c[None, :]).max(axis=1)
        else:
            r = 1.0 / rmax
            c = 1.0 / (mag * r[:, None]).max(axis=0)
        S = A * r[:, None] * c[None, :]
        return c[:, None] * np.linalg.solve(S, r[:, None] * rhs)

    def global_tmatrix(self, equilibrate=True):
        """The cluster's T-matrix in one spherical basis at the origin with the same
        lmax: T_g = sum_ij R(0 <- p_i) X_ij R(p_j <- 0) with regular translations
        (python/treams_rs/_tmatrix.py:243-257, sw/expansion.rs:116-132)."""
        X = self.interaction_tmatrix(equilibrate)
        M = self.M
        Tg = np.zeros((M, M), np.complex128)
        rout = [translation_block(self.lmax, self.k0, -self.positions[i], True) for i in range(self.N)]
        rin = [translation_block(self.lmax, self.k0, self.positions[j], True) for j in range(self.N)]
        for i in range(self.N):
            for j in range(self.N):
                Tg += rout[i] @ X[i * M:(i + 1) * M, j * M:(j + 1) * M] @ rin[j]
        return Tg

    def solve_gmres(self, incident, **options):
        """Matrix-free GMRES per column (cluster/iterative.rs:236-284)."""
        rhs = apply_local(self.blocks, incident, False)
        if rhs.ndim == 1:
            return gmres(self.apply, rhs, **options)
        out = np.zeros_like(rhs)
        conv = []
        for c in range(rhs.shape[1]):
            out[:, c], cv = gmres(self.apply, rhs[:, c].copy(), **options)
            conv.append(cv)
        return out, conv

    def cross_sections(self, incident, x=None, flux=0.5):
        """(sca, ext) with weights 1/k^2 and the regular overlap R
        (python/treams_rs/_framework_tmatrix.py:231-289)."""
        if x is None:
            x = self.solve(incident)
        w = 1.0 / self.k0 ** 2
        weighted = x * w
        radiated = self.R @ weighted if self.N > 1 else weighted
        sca = float(np.real(np.vdot(x, radiated))) * 0.5 / flux
        ext = float(-np.real(np.vdot(incident, weighted))) * 0.5 / flux
        return sca, ext

    def pullback(self, incident, x, G):
        """Gradients of a real loss with d L = Re<G, dx> through x = A^{-1} T a:
        cluster/interaction.rs:533-558 (Y = A^{-H} G, B = T^H Y, C = T^H Y X^H,
        T = Y (a + C x)^H on the 2x2 blocks) with the pair pullback of
        cluster/iterative.rs:500-551 and the Mie pullback of iterative.rs:564-592.
        Returns dict(radii, eps, positions, k0, incident)."""
        A = self.operator()
        Y = np.linalg.solve(A.conj().T, G)
        inc_grad = apply_local(self.blocks, Y, True)
        X = x.reshape(self.n, -1)
        Yb = inc_grad.reshape(self.n, -1)
        Cbar = Yb @ X.conj().T
        gpos, gk = pair_matrix_pullback(self.lmax, self.k0, self.positions, False, Cbar)
        response = (incident.reshape(self.n, -1) + self.C @ X)
        Ym = Y.reshape(self.n, -1)
        deg = block_degrees(self.lmax)
        g_r = np.zeros(self.N)
        g_e = np.zeros(self.N, np.complex128)
        for i in range(self.N):
            dc = np.zeros((self.lmax, 2, 2), np.complex128)
            for b, d in enumerate(deg):
                o = i * self.M + 2 * b
                dc[d] += Ym[o:o + 2] @ response[o:o + 2].conj().T
            for d in range(self.lmax):
                gs, ge = mie_pullback(d + 1, self.k0 * self.radii[i], self.eps[i], to_mode_order(dc[d]))
                g_r[i] += self.k0 * gs
                gk += self.radii[i] * gs
                g_e[i] += ge
        return dict(radii=g_r, eps=g_e, positions=gpos, k0=gk, incident=inc_grad)

    def scattering_gradient(self, direction, pol):
        """sigma_sca of one plane wave and its gradient w.r.t. radii and positions,
        through the solve, the regular overlap R and the incident position phases."""
        kvec = self.k0 * np.asarray(direction, np.float64) / np.linalg.norm(direction)
        a = self.plane_wave(direction, pol)
        x = self.solve(a)
        w = 1.0 / self.k0 ** 2
        R = self.R
        sca = float(np.real(np.vdot(x, R @ x))) * w
        G = (R + R.conj().T) @ x * w
        grads = self.pullback(a, x, G)
        gpos = grads["positions"].copy()
        Rbar = np.outer(x, x.conj()) * w
        gR, _ = pair_matrix_pullback(self.lmax, self.k0, self.positions, True, Rbar)
        gpos += gR
        abar = grads["incident"].reshape(self.N, self.M)
        ar = a.reshape(self.N, self.M)
        phase_grad = np.real(np.sum(np.conj(abar) * 1j * ar, axis=1))
        gpos += phase_grad[:, None] * kvec[None, :]
        return sca, grads["radii"], gpos
