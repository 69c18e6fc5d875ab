"""Sphere clusters on the GPU, with the reference's Python-facing interface.

``IterativeSphereCluster`` mirrors ``treams_rs._native.IterativeSphereCluster``
(treams-py/src/iterative.rs:45-209, python/treams_rs/_native.pyi:910-1170): the
same constructor arguments, ``solve`` with ``rtol/atol/restart/max_iterations``
returning ``(X, [(iterations, residual_norm, rhs_norm)])``, ``record`` and
``IterativeContext.pullback``.  Vectors follow the reference layout: particle-major,
``sw::modes`` order inside a particle, one column per illumination.

Arrays may be NumPy or torch; results are torch tensors on the device.  All
numerical work runs in libtreams_gpu.so; there is no CPU path.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from ._lib import GmresOptions as _Options, Convergence as _Conv, check, lib

#: linalg/gmres.rs:34-43 / python/treams_rs/iterative.py:302-310
DEFAULT_RTOL = 1e-10
DEFAULT_ATOL = 0.0
DEFAULT_RESTART = 30
DEFAULT_MAX_ITERATIONS = 300

_PC = {"none": _lib.TG_PC_NONE, "block_jacobi": _lib.TG_PC_BLOCK_JACOBI, "learned": _lib.TG_PC_LEARNED}


def _device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("treams-rs_amd needs a ROCm GPU (torch.cuda is not available)")
    return torch.device("cuda", torch.cuda.current_device())


def _stream_ptr() -> int:
    return torch.cuda.current_stream().cuda_stream


def _columns(a, n: int) -> tuple[torch.Tensor, bool]:
    """(cols, n) contiguous complex128 device tensor = column-major n x cols."""
    t = torch.as_tensor(a)
    vector = t.ndim == 1
    if vector:
        t = t[:, None]
    if t.ndim != 2 or t.shape[0] != n:
        raise ValueError(f"expected shape ({n},) or ({n}, columns), got {tuple(t.shape)}")
    t = t.to(device=_device(), dtype=torch.complex128)
    return t.T.contiguous(), vector


def _back(cm: torch.Tensor, vector: bool) -> torch.Tensor:
    out = cm.T
    return out[:, 0].contiguous() if vector else out.contiguous()


@dataclass(frozen=True)
class Convergence:
    """linalg/gmres.rs:81-93 (plus the mixed-precision refinement count)."""

    iterations: int
    residual_norm: float
    rhs_norm: float
    refinements: int = 0

    def as_tuple(self):
        return (self.iterations, self.residual_norm, self.rhs_norm)


class IterativeSphereCluster:
    """Homogeneous nonmagnetic spheres in vacuum, solved on the GPU
    (cluster/iterative.rs:53-86)."""

    def __init__(self, lmax: int, k0: float, radii, epsilon, positions):
        radii = np.ascontiguousarray(np.asarray(radii, np.float64).reshape(-1))
        eps = np.ascontiguousarray(np.broadcast_to(np.asarray(epsilon, np.complex128), radii.shape))
        pos = np.ascontiguousarray(np.asarray(positions, np.float64).reshape(-1, 3))
        if pos.shape[0] != radii.size:
            raise ValueError("require one position per sphere")
        self.lmax, self.k0 = int(lmax), float(k0)
        self.radii, self.epsilon, self.positions = radii.copy(), eps.copy(), pos.copy()
        self._dev = _device()
        h = ctypes.c_void_p()
        check(lib().tg_cluster_create(self.lmax, self.k0, radii.size, radii.ctypes.data, eps.ctypes.data,
                                      pos.ctypes.data, ctypes.c_void_p(_stream_ptr()), ctypes.byref(h)))
        self._h = h
        self._destroy = lib().tg_cluster_destroy  # bound now: module globals are gone at exit
        self.N = radii.size
        self.modes_per_sphere = lib().tg_modes_per_sphere(self.lmax)
        self.n = int(lib().tg_cluster_dimension(h))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._destroy(h)
            self._h = None

    @classmethod
    def from_tmatrices(cls, lmax: int, k0: float, positions, tmatrices) -> "IterativeSphereCluster":
        """Particle cluster: arbitrary local T-matrices (N, M, M) in the reference mode
        order at distinct positions (cluster/particles.rs:65-85).  Solves, factors,
        the interacting and global T-matrices and cross sections work as for spheres;
        the preconditioners and the radius / permittivity pullbacks need spheres."""
        pos = np.ascontiguousarray(np.asarray(positions, np.float64).reshape(-1, 3))
        N, M = pos.shape[0], int(lib().tg_modes_per_sphere(int(lmax)))
        t = torch.as_tensor(np.asarray(tmatrices) if not torch.is_tensor(tmatrices) else tmatrices)
        t = t.to(_device(), torch.complex128).reshape(N, M, M)
        tcol = t.transpose(1, 2).contiguous()  # column-major blocks
        self = cls.__new__(cls)
        self.lmax, self.k0 = int(lmax), float(k0)
        self.radii = self.epsilon = None
        self.positions = pos.copy()
        self._dev = _device()
        h = ctypes.c_void_p()
        check(lib().tg_cluster_create_particles(self.lmax, self.k0, N, pos.ctypes.data, tcol.data_ptr(),
                                                ctypes.c_void_p(_stream_ptr()), ctypes.byref(h)))
        self._h = h
        self._destroy = lib().tg_cluster_destroy
        self.N = N
        self.modes_per_sphere = M
        self.n = int(lib().tg_cluster_dimension(h))
        self.tmatrices = t
        return self

    # ---------------------------------------------------------------- queries
    @property
    def dimension(self) -> int:
        return self.n

    def mie_blocks(self) -> np.ndarray:
        """(N, lmax, 2, 2) Mie blocks in mode order."""
        out = np.zeros(self.N * self.lmax * 4, np.complex128)
        check(lib().tg_cluster_mie_blocks(self._h, out.ctypes.data))
        return out.reshape(self.N, self.lmax, 2, 2)

    def coupling_matrix(self, regular: bool = False) -> torch.Tensor:
        """Dense C (or the regular overlap R) as an (n, n) device tensor
        (cluster/spheres.rs:188-220)."""
        out = torch.empty((self.n, self.n), dtype=torch.complex128, device=self._dev)
        check(lib().tg_pair_matrix(self._h, int(regular), out.data_ptr()))
        return out.T  # memory is column-major

    def assemble(self, storage: str = "c64") -> None:
        """Keep the coupling on the device (helicity-split, c64 or c128)."""
        s = {"c64": _lib.TG_STORE_C64, "c128": _lib.TG_STORE_C128}[storage]
        check(lib().tg_cluster_assemble(self._h, s))

    def set_group_size(self, spheres: int) -> None:
        """Spheres per block-Jacobi group."""
        check(lib().tg_cluster_set_group_size(self._h, int(spheres)))

    def set_table_limit(self, nbytes: int) -> None:
        """Largest pair-table set the matrix-free operators keep (bytes); larger
        clusters recompute every pair's translation coefficients per apply."""
        check(lib().tg_cluster_set_table_limit(self._h, int(nbytes)))

    def set_inner_gemm(self, kind: str) -> None:
        """Arithmetic of the c64-precision coupling GEMMs of the inner solves:
        "split16" (default; scaled FP16 hi + lo pairs on FP16 MFMA) or "fp32" (c64
        tables on FP32 MFMA)."""
        kinds = {"fp32": 0, "split16": 1}
        if kind not in kinds:
            raise ValueError(f"inner GEMM kind must be one of {sorted(kinds)}, got {kind!r}")
        check(lib().tg_cluster_set_inner_gemm(self._h, kinds[kind]))

    def set_inner_relaxation(self, threshold: float) -> None:
        """Relaxed inner applies of mixed-precision solves (inexact Krylov): an inner
        Arnoldi step whose allowed relative operator error (inner tolerance over the
        step's residual estimate) is at least `threshold` multiplies the FP16 hi halves
        only.  0 switches it off; default 1e-3.  The FP64 residual certifies the answer
        either way."""
        check(lib().tg_cluster_set_inner_relaxation(self._h, float(threshold)))

    def release(self) -> None:
        """Free the stored coupling, the pair tables and the preconditioner (rebuilt on
        demand)."""
        check(lib().tg_cluster_release(self._h))

    def reset(self) -> None:
        """Mark the stored coupling, the pair tables and the preconditioner stale while
        keeping their device memory: the next solve recomputes them in place."""
        check(lib().tg_cluster_reset(self._h))

    # -------------------------------------------------------------- operators
    def _apply(self, fn, x, *args):
        cm, vector = _columns(x, self.n)
        out = torch.empty_like(cm)
        check(fn(self._h, cm.data_ptr(), out.data_ptr(), cm.shape[0], *args))
        return _back(out, vector)

    @staticmethod
    def _op(stored: bool, c64) -> int:
        """tg_operator: FP64 matrix-free, stored coupling, or matrix-free c64 (c64=True;
        c64="hi" the relaxed arithmetic of split-FP16 tables, hi halves only)."""
        if stored:
            return _lib.TG_OP_STORED
        if c64 == "hi":
            return _lib.TG_OP_MATRIX_FREE_C64_HI
        return _lib.TG_OP_MATRIX_FREE_C64 if c64 else _lib.TG_OP_MATRIX_FREE_F64

    def apply(self, x, adjoint: bool = False, stored: bool = False, c64: bool = False):
        """(I - T C) x or (I - C^H T^H) x (cluster/iterative.rs:225-232)."""
        return self._apply(lib().tg_cluster_apply, x, int(adjoint), self._op(stored, c64))

    def coupling(self, x, adjoint: bool = False, stored: bool = False, c64: bool = False):
        return self._apply(lib().tg_cluster_coupling, x, int(adjoint), self._op(stored, c64))

    def local(self, x, adjoint: bool = False):
        return self._apply(lib().tg_cluster_local, x, int(adjoint))

    def overlap(self, x, adjoint: bool = False):
        """R x or R^H x, the regular overlap (python/treams_rs/_tmatrix.py:404-411)."""
        return self._apply(lib().tg_cluster_overlap, x, int(adjoint))

    def plane_wave(self, direction, pol) -> torch.Tensor:
        """Regular local coefficients of a unit plane wave (pw/expand.rs:23-66)."""
        if isinstance(pol, str):
            pol = {"positive_helicity": 1, "negative_helicity": 0}[pol]
        d = np.ascontiguousarray(np.asarray(direction, np.float64).reshape(3))
        out = torch.empty(self.n, dtype=torch.complex128, device=self._dev)
        check(lib().tg_cluster_plane_wave(self._h, d.ctypes.data, int(pol), out.data_ptr()))
        return out

    def field(self, x, points, regular: bool = False) -> torch.Tensor:
        """Electric field of helicity vector spherical waves with coefficients x (n or
        n x P, reference layout) about every sphere, at points (Q x 3): (Q, 3) or
        (Q, 3, P) complex128 (fields.rs:947).  Outgoing waves (the efield of a
        scattered wave; NaN at a sphere centre), or regular waves with ``regular``
        (an incident expansion; finite everywhere)."""
        cm, vector = _columns(x, self.n)
        pts = torch.as_tensor(points, dtype=torch.float64).reshape(-1, 3).to(self._dev).contiguous()
        q = pts.shape[0]
        out = torch.empty((cm.shape[0], 3 * q), dtype=torch.complex128, device=self._dev)
        check(lib().tg_cluster_field(self._h, cm.data_ptr(), cm.shape[0], q, pts.data_ptr(), int(regular),
                                     out.data_ptr()))
        out = out.reshape(cm.shape[0], q, 3).permute(1, 2, 0)
        return out[..., 0].contiguous() if vector else out.contiguous()

    # ------------------------------------------------------------------ solves
    @staticmethod
    def options(rtol=DEFAULT_RTOL, atol=DEFAULT_ATOL, restart=DEFAULT_RESTART,
                max_iterations=DEFAULT_MAX_ITERATIONS, mixed_precision=False, preconditioner="none") -> _Options:
        """mixed_precision: False (FP64), True or "stored" (inner solves on the stored c64
        coupling) or "matrix_free" (inner solves on the matrix-free c64 operator)."""
        mp = {False: 0, True: 1, "stored": 1, "matrix_free": 2}[mixed_precision]
        return _Options(float(rtol), float(atol), int(restart), int(max_iterations), mp,
                        _PC[preconditioner])

    def solve(self, incident, *, rtol=DEFAULT_RTOL, atol=DEFAULT_ATOL, restart=DEFAULT_RESTART,
              max_iterations=DEFAULT_MAX_ITERATIONS, mixed_precision=False, preconditioner="none",
              adjoint=False, convergence="tuples"):
        """Solve (I - T C) X = T B (cluster/iterative.rs:264-284).  Returns
        (X, [(iterations, residual_norm, rhs_norm), ...])."""
        cm, vector = _columns(incident, self.n)
        cols = cm.shape[0]
        out = torch.empty_like(cm)
        conv = (_Conv * cols)()
        opts = self.options(rtol, atol, restart, max_iterations, mixed_precision, preconditioner)
        check(lib().tg_cluster_solve(self._h, cm.data_ptr(), cols, ctypes.byref(opts), int(adjoint),
                                     out.data_ptr(), conv))
        reports = [Convergence(c.iterations, c.residual_norm, c.rhs_norm, c.refinements) for c in conv]
        if convergence == "tuples":
            reports = [r.as_tuple() for r in reports]
        return _back(out, vector), reports

    def record(self, incident, **options):
        """Solve and keep what the pullback needs (cluster/iterative.rs:288-300)."""
        x, conv = self.solve(incident, **options)
        return x, IterativeContext(self, incident, x, options), conv

    def cross_sections(self, incident, x, flux: float = 0.5):
        """(scattering, extinction) of one illumination
        (python/treams_rs/_framework_tmatrix.py:231-289)."""
        a, _ = _columns(incident, self.n)
        xc, _ = _columns(x, self.n)
        sca, ext = ctypes.c_double(), ctypes.c_double()
        check(lib().tg_cluster_cross_sections(self._h, a.data_ptr(), xc.data_ptr(), float(flux), ctypes.byref(sca),
                                              ctypes.byref(ext)))
        return sca.value, ext.value

    def overlap_pullback(self, u, v) -> np.ndarray:
        """Position gradient of Re<sum_c u_c v_c^H, R> (regular overlap pullback)."""
        uc, _ = _columns(u, self.n)
        vc, _ = _columns(v, self.n)
        g = np.zeros(3 * self.N)
        check(lib().tg_cluster_overlap_pullback(self._h, uc.data_ptr(), vc.data_ptr(), uc.shape[0], g.ctypes.data))
        return g.reshape(self.N, 3)

    def timing(self, enable: bool = True):
        check(lib().tg_cluster_timing_enable(self._h, int(enable)))

    def timing_get(self, kernel: str):
        ms, count = ctypes.c_double(), ctypes.c_int64()
        check(lib().tg_cluster_timing_get(self._h, kernel.encode(), ctypes.byref(ms), ctypes.byref(count)))
        return ms.value, count.value

    def timing_reset(self):
        check(lib().tg_cluster_timing_reset(self._h))


class IterativeContext:
    """Residual of a recorded solve (cluster/iterative.rs:306-489)."""

    def __init__(self, cluster: IterativeSphereCluster, incident, x, options):
        self.cluster = cluster
        self.incident = torch.as_tensor(incident).to(cluster._dev, torch.complex128)
        self.x = x
        self.options = dict(options)

    def pushforward(self, k0=0.0, radii=None, epsilon=None, positions=None, incident=None):
        """Directional derivative of the converged solution along tangents of k0, the
        radii, permittivities, positions and incident columns (None: zero), by one
        tangent GMRES solve (cluster/iterative.rs:362-447; python/treams_rs/
        iterative.py:152-180).  Returns (dX, [(iterations, residual, rhs_norm)])."""
        c = self.cluster
        a, vector = _columns(self.incident, c.n)
        xc, _ = _columns(self.x, c.n)
        dk, tr, te, tp = _tangents(c, k0, radii, epsilon, positions)
        ti = _incident_tangent(incident, a, c.n)
        cols = a.shape[0]
        out = torch.empty_like(a)
        conv = (_Conv * cols)()
        opts = c.options(**{k: v for k, v in self.options.items() if k in (
            "rtol", "atol", "restart", "max_iterations", "mixed_precision", "preconditioner")})
        check(lib().tg_cluster_pushforward(c._h, a.data_ptr(), xc.data_ptr(), cols, ctypes.byref(opts), dk,
                                           tr.ctypes.data, te.ctypes.data, tp.ctypes.data,
                                           None if ti is None else ti.data_ptr(), out.data_ptr(), conv))
        reports = [(v.iterations, v.residual_norm, v.rhs_norm) for v in conv]
        return _back(out, vector), reports

    def _opts(self):
        return self.cluster.options(**{k: v for k, v in self.options.items() if k in (
            "rtol", "atol", "restart", "max_iterations", "mixed_precision", "preconditioner")})

    def pushforward_local(self, k0=0.0, positions=None, local=None, incident=None):
        """dX with the local T-matrices as inputs: tangents of k0 (coupling only), the
        positions, the (N, M, M) local blocks and the incident columns
        (cluster/particles.rs:206-220, cluster/interaction.rs:464-478)."""
        c = self.cluster
        dx, conv = _pushforward_local(lib().tg_cluster_pushforward_local, c._h, c, self.incident, self.x,
                                      ctypes.byref(self._opts()), k0, positions, local, incident)
        return dx, [(v.iterations, v.residual_norm, v.rhs_norm) for v in conv]

    def pullback_local(self, cotangent):
        """(k0, positions[N, 3], local blocks (N, M, M), incident, convergence) with the
        local T-matrices as inputs (cluster/particles.rs:231-250)."""
        c = self.cluster
        gk, gp, gl, gi, conv = _pullback_local(lib().tg_cluster_pullback_local, c._h, c, self.incident, self.x,
                                               cotangent, ctypes.byref(self._opts()))
        return gk, gp, gl, gi, [(v.iterations, v.residual_norm, v.rhs_norm) for v in conv]

    def pullback(self, cotangent):
        """Gradients from the cotangent of X: returns
        (k0, radii[N], eps[N], positions[N, 3], incident, convergence)."""
        c = self.cluster
        a, vector = _columns(self.incident, c.n)
        xc, _ = _columns(self.x, c.n)
        g, _ = _columns(cotangent, c.n)
        cols = a.shape[0]
        inc = torch.empty_like(a)
        gk = ctypes.c_double()
        gr = np.zeros(c.N)
        ge = np.zeros(c.N, np.complex128)
        gp = np.zeros(3 * c.N)
        conv = (_Conv * cols)()
        opts = c.options(**{k: v for k, v in self.options.items() if k in (
            "rtol", "atol", "restart", "max_iterations", "mixed_precision", "preconditioner")})
        check(lib().tg_cluster_pullback(c._h, a.data_ptr(), xc.data_ptr(), g.data_ptr(), cols, ctypes.byref(opts),
                                        ctypes.byref(gk), gr.ctypes.data, ge.ctypes.data, gp.ctypes.data,
                                        inc.data_ptr(), conv))
        reports = [(v.iterations, v.residual_norm, v.rhs_norm) for v in conv]
        return gk.value, gr, ge, gp.reshape(c.N, 3), _back(inc, vector), reports


def _tangents(cluster, k0, radii, epsilon, positions):
    """Host arrays of the physical tangents (zeros where None)."""
    N = cluster.N
    tr = np.zeros(N) if radii is None else np.ascontiguousarray(np.broadcast_to(np.asarray(radii, np.float64), (N,)))
    te = (np.zeros(N, np.complex128) if epsilon is None else
          np.ascontiguousarray(np.broadcast_to(np.asarray(epsilon, np.complex128), (N,))))
    tp = (np.zeros((N, 3)) if positions is None else
          np.ascontiguousarray(np.asarray(positions, np.float64).reshape(N, 3)))
    return float(k0 or 0.0), tr, te, tp


def _incident_tangent(incident, like: torch.Tensor, n: int):
    if incident is None:
        return None
    t, _ = _columns(incident, n)
    if t.shape != like.shape:
        raise ValueError("incident tangent shape does not match the recorded incident fields")
    return t


def _factor_info(h) -> dict:
    p, v, f = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    check(lib().tg_factor_info(h, ctypes.byref(p), ctypes.byref(v), ctypes.byref(f)))
    return {"precision": "c64" if p.value == _lib.TG_STORE_C64 else "c128",
            "pivoting": "column" if v.value else "block", "fallback": bool(f.value)}


def _local_blocks_in(local, N: int, M: int, dev):
    """(N, M, M) local tangent blocks as column-major device memory (None stays None)."""
    if local is None:
        return None
    t = torch.as_tensor(local).to(dev, torch.complex128).reshape(N, M, M)
    return t.transpose(1, 2).contiguous()


def _pushforward_local(fn, handle, c, incident, x, cols_opts, k0, positions, local, incident_tangent):
    a, vector = _columns(incident, c.n)
    xc, _ = _columns(x, c.n)
    ti = _incident_tangent(incident_tangent, a, c.n)
    tp = (np.zeros((c.N, 3)) if positions is None else
          np.ascontiguousarray(np.asarray(positions, np.float64).reshape(c.N, 3)))
    tl = _local_blocks_in(local, c.N, c.modes_per_sphere, c._dev)
    cols = a.shape[0]
    out = torch.empty_like(a)
    conv = (_Conv * cols)()
    check(fn(handle, a.data_ptr(), xc.data_ptr(), cols, cols_opts, float(k0 or 0.0), tp.ctypes.data,
             None if tl is None else tl.data_ptr(), None if ti is None else ti.data_ptr(), out.data_ptr(), conv))
    return _back(out, vector), conv


def _pullback_local(fn, handle, c, incident, x, cotangent, opts):
    a, vector = _columns(incident, c.n)
    xc, _ = _columns(x, c.n)
    g, _ = _columns(cotangent, c.n)
    cols = a.shape[0]
    inc = torch.empty_like(a)
    M = c.modes_per_sphere
    gl = torch.empty((c.N, M, M), dtype=torch.complex128, device=c._dev)  # column-major blocks
    gk = ctypes.c_double()
    gp = np.zeros(3 * c.N)
    conv = (_Conv * cols)()
    check(fn(handle, a.data_ptr(), xc.data_ptr(), g.data_ptr(), cols, opts, ctypes.byref(gk), gp.ctypes.data,
             gl.data_ptr(), inc.data_ptr(), conv))
    return gk.value, gp.reshape(c.N, 3), gl.transpose(1, 2), _back(inc, vector), conv


class DenseFactor:
    """LU factors of a caller's dense n x n matrix on the device (Lu::new,
    linalg/mod.rs:398-428), with the cluster factor's kernels, equilibration and
    fallbacks; solves are refined in FP64 against the matrix to ``rtol``."""

    def __init__(self, matrix, precision: str = "c128", rtol: float = 1e-12):
        a = torch.as_tensor(matrix).to(_device(), torch.complex128)
        if a.ndim != 2 or a.shape[0] != a.shape[1]:
            raise ValueError("require a square matrix")
        self.n = int(a.shape[0])
        self.rtol = float(rtol)
        p = {"c64": _lib.TG_STORE_C64, "c128": _lib.TG_STORE_C128}[precision]
        acol = a.t().contiguous()  # column-major
        h = ctypes.c_void_p()
        check(lib().tg_dense_factor_create(self.n, acol.data_ptr(), p, ctypes.c_void_p(_stream_ptr()),
                                           ctypes.byref(h)))
        self._h = h
        self._destroy = lib().tg_factor_destroy
        self._dev = a.device

    def close(self) -> None:
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._destroy(h)
        self._h = None

    def __del__(self):
        self.close()

    @property
    def info(self) -> dict:
        return _factor_info(self._h)

    def solve(self, b, adjoint: bool = False):
        """A^{-1} b (A^{-H} b); returns (x, [Convergence])."""
        if self._h is None:
            raise ValueError("the factor was closed")
        cm, vector = _columns(b, self.n)
        cols = cm.shape[0]
        out = torch.empty_like(cm)
        conv = (_Conv * cols)()
        check(lib().tg_dense_factor_solve(self._h, cm.data_ptr(), cols, self.rtol, int(adjoint), out.data_ptr(),
                                          conv))
        return _back(out, vector), [Convergence(0, c.residual_norm, c.rhs_norm, c.refinements) for c in conv]


class SphereClusterFactor:
    """LU factors of (I - T C) on the device for any number of illuminations: mirrors
    sphere_cluster_factor / InteractionFactor (cluster/spheres.rs:85-94,
    cluster/interaction.rs:180-240).  ``precision`` "c128" or "c64" (complex FP32
    factors on MFMA); every solve is refined with the FP64 residual to ``rtol``."""

    def __init__(self, cluster: IterativeSphereCluster, precision: str = "c128", rtol: float = 1e-12):
        p = {"c64": _lib.TG_STORE_C64, "c128": _lib.TG_STORE_C128}[precision]
        self.cluster, self.precision, self.rtol = cluster, precision, float(rtol)
        h = ctypes.c_void_p()
        check(lib().tg_factor_create(cluster._h, p, ctypes.byref(h)))
        self._h = h
        self._destroy = lib().tg_factor_destroy

    def close(self) -> None:
        """Free the factors (n^2 x 8 or 16 bytes of device memory)."""
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._destroy(h)
        self._h = None

    def __del__(self):
        self.close()

    @property
    def dimension(self) -> int:
        return self.cluster.n

    def _solve(self, b, adjoint: bool):
        if self._h is None:
            raise ValueError("the factor was closed")
        cm, vector = _columns(b, self.cluster.n)
        cols = cm.shape[0]
        out = torch.empty_like(cm)
        conv = (_Conv * cols)()
        check(lib().tg_factor_solve(self._h, cm.data_ptr(), cols, self.rtol, int(adjoint), out.data_ptr(), conv))
        reports = [Convergence(0, c.residual_norm, c.rhs_norm, c.refinements) for c in conv]
        return _back(out, vector), reports

    def solve(self, incident):
        """X = (I - T C)^{-1} T A for the supplied columns (cluster/interaction.rs:213-218).
        Returns (X, [Convergence])."""
        return self._solve(incident, False)

    def solve_adjoint(self, rhs):
        """(I - T C)^{-H} rhs (linalg/mod.rs:436-438)."""
        return self._solve(rhs, True)

    def pullback_incident(self, cotangent):
        """T^H (I - T C)^{-H} G (cluster/interaction.rs:220-226)."""
        y, _ = self._solve(cotangent, True)
        return self.cluster.local(y, adjoint=True)

    def record(self, incident):
        """Solve and keep what the pullback needs (cluster/interaction.rs:228-239)."""
        x, conv = self.solve(incident)
        return x, FactorContext(self, incident, x), conv

    @property
    def info(self) -> dict:
        """The factor's precision ("c64"/"c128", after any fallback), pivoting ("block":
        inside 64-row diagonal blocks, "column": over whole columns) and whether a
        stalled or singular solve refactored it (fallback)."""
        return _factor_info(self._h)

    @property
    def equilibrated(self) -> bool:
        """Whether the factors are those of the row/column equilibrated operator
        (linalg/mod.rs:137-236)."""
        return bool(lib().tg_factor_equilibrated(self._h))

    def global_tmatrix(self, lmax: int | None = None):
        """The cluster's T-matrix in one spherical basis at the origin with degrees up
        to ``lmax`` (default: the spheres'),
        ``Cluster(...).solve().expand(SphericalBasis.default(lmax))``
        (python/treams_rs/_tmatrix.py:243-257): Mg x Mg, sphere mode order.  Returns
        (T, [Convergence] of the Mg solves)."""
        if self._h is None:
            raise ValueError("the factor was closed")
        lg = self.cluster.lmax if lmax is None else int(lmax)
        M = int(lib().tg_modes_per_sphere(lg))
        out = torch.empty((M, M), dtype=torch.complex128, device=self.cluster._dev)
        conv = (_Conv * M)()
        check(lib().tg_factor_global_tmatrix(self._h, lg, self.rtol, out.data_ptr(), conv))
        reports = [Convergence(0, c.residual_norm, c.rhs_norm, c.refinements) for c in conv]
        # column-major memory: out holds T_g transposed
        return out.t().contiguous(), reports


class FactorContext:
    """Pullback of a factored solve with respect to k0, radii, permittivities,
    positions and the incident field, the adjoint solve reusing the LU factors
    (cluster/spheres.rs:321-378)."""

    def __init__(self, factor: SphereClusterFactor, incident, x):
        self.factor = factor
        self.incident = torch.as_tensor(incident).to(factor.cluster._dev, torch.complex128)
        self.x = x

    def pushforward(self, k0=0.0, radii=None, epsilon=None, positions=None, incident=None):
        """Directional derivative of the solution along tangents of k0, the radii,
        permittivities, positions and incident columns (None: zero), the tangent
        solve reusing the LU factors (cluster/spheres.rs:263-300,
        cluster/interaction.rs:464-478).  Returns (dX, [Convergence])."""
        f, c = self.factor, self.factor.cluster
        a, vector = _columns(self.incident, c.n)
        xc, _ = _columns(self.x, c.n)
        dk, tr, te, tp = _tangents(c, k0, radii, epsilon, positions)
        ti = _incident_tangent(incident, a, c.n)
        cols = a.shape[0]
        out = torch.empty_like(a)
        conv = (_Conv * cols)()
        check(lib().tg_factor_pushforward(f._h, a.data_ptr(), xc.data_ptr(), cols, f.rtol, dk, tr.ctypes.data,
                                          te.ctypes.data, tp.ctypes.data, None if ti is None else ti.data_ptr(),
                                          out.data_ptr(), conv))
        reports = [Convergence(0, v.residual_norm, v.rhs_norm, v.refinements) for v in conv]
        return _back(out, vector), reports

    def pushforward_local(self, k0=0.0, positions=None, local=None, incident=None):
        """dX with the local T-matrices as inputs (tangents of k0 through the coupling,
        positions, (N, M, M) local blocks, incident columns), the tangent solve by the
        factors (cluster/particles.rs:206-220, cluster/interaction.rs:464-478)."""
        f, c = self.factor, self.factor.cluster
        dx, conv = _pushforward_local(lib().tg_factor_pushforward_local, f._h, c, self.incident, self.x,
                                      ctypes.c_double(f.rtol), k0, positions, local, incident)
        return dx, [Convergence(0, v.residual_norm, v.rhs_norm, v.refinements) for v in conv]

    def pullback_local(self, cotangent):
        """(k0, positions[N, 3], local blocks (N, M, M), incident, convergence) with the
        local T-matrices as inputs (cluster/particles.rs:231-250,
        cluster/interaction.rs:533-558)."""
        f, c = self.factor, self.factor.cluster
        gk, gp, gl, gi, conv = _pullback_local(lib().tg_factor_pullback_local, f._h, c, self.incident, self.x,
                                               cotangent, ctypes.c_double(f.rtol))
        return gk, gp, gl, gi, [Convergence(0, v.residual_norm, v.rhs_norm, v.refinements) for v in conv]

    def pullback(self, cotangent):
        """Returns (k0, radii[N], eps[N], positions[N, 3], incident, convergence)."""
        f, c = self.factor, self.factor.cluster
        a, vector = _columns(self.incident, c.n)
        xc, _ = _columns(self.x, c.n)
        g, _ = _columns(cotangent, c.n)
        cols = a.shape[0]
        inc = torch.empty_like(a)
        gk = ctypes.c_double()
        gr = np.zeros(c.N)
        ge = np.zeros(c.N, np.complex128)
        gp = np.zeros(3 * c.N)
        conv = (_Conv * cols)()
        check(lib().tg_factor_pullback(f._h, a.data_ptr(), xc.data_ptr(), g.data_ptr(), cols, f.rtol,
                                       ctypes.byref(gk), gr.ctypes.data, ge.ctypes.data, gp.ctypes.data,
                                       inc.data_ptr(), conv))
        reports = [(v.refinements, v.residual_norm, v.rhs_norm) for v in conv]
        return gk.value, gr, ge, gp.reshape(c.N, 3), _back(inc, vector), reports


def mode_index(l: int, m: int, pol: int) -> int:
    """Index of mode (l, m, pol) in one particle's block of the reference mode order
    (l = 1.., m = -l..l, polarisation 1 before 0; sw/mod.rs:114-125)."""
    return 2 * (l * l - 1 + m + l) + (1 - pol)


def particle_cluster(local, positions, k0: float, *, lmax=None, modes=None, precision: str = "c128",
                     rtol: float = 1e-12):
    """Coupled T-matrix (I - T C)^{-1} T of particles with arbitrary local T-matrices,
    over the concatenated particle bases (python/treams_rs/diff.py:1473-1523,
    cluster/particles.rs:65-85; spherical waves, helicity basis, one real wavenumber).

    local: one (n_i, n_i) T-matrix per particle; its basis is either the full basis
    up to lmax[i] (reference mode order) or the mode list modes[i] of (l, m, pol).
    The particles are embedded in the basis up to the largest degree with zero rows and
    columns for modes outside their own basis (those solve to zero, so the result is
    the reduced system's); the dense LU of the embedded system (refined in FP64 to
    rtol) gives the columns of the own modes."""
    pos = np.asarray(positions, np.float64).reshape(-1, 3)
    N = pos.shape[0]
    if len(local) != N or N == 0:
        raise ValueError("one local matrix and position required per particle")
    if (lmax is None) == (modes is None):
        raise ValueError("give either lmax (one per particle) or modes (one list per particle)")
    if modes is None:
        lm = np.broadcast_to(np.asarray(lmax, int), (N,))
        modes = []
        for li in lm:
            modes.append([(l, m, p) for l in range(1, int(li) + 1) for m in range(-l, l + 1) for p in (1, 0)])
    L = max(l for ms in modes for (l, _, _) in ms)
    M = int(lib().tg_modes_per_sphere(L))
    full = np.zeros((N, M, M), np.complex128)
    keep = []
    for i, (t, ms) in enumerate(zip(local, modes)):
        t = np.asarray(t, np.complex128)
        idx = np.array([mode_index(l, m, p) for (l, m, p) in ms], int)
        if t.shape != (len(idx), len(idx)) or len(set(idx.tolist())) != len(idx):
            raise ValueError("local matrix shape must match its basis, modes distinct")
        full[i][np.ix_(idx, idx)] = t
        keep.extend((i * M + idx).tolist())
    keep = np.asarray(keep, int)
    cl = IterativeSphereCluster.from_tmatrices(L, k0, pos, full)
    f = SphereClusterFactor(cl, precision, rtol)
    try:
        inc = torch.zeros((cl.n, keep.size), dtype=torch.complex128, device=cl._dev)
        inc[torch.as_tensor(keep, device=cl._dev), torch.arange(keep.size, device=cl._dev)] = 1.0
        x, _ = f.solve(inc)
        return x[torch.as_tensor(keep, device=cl._dev)]
    finally:
        f.close()


def sphere_cluster_factor(lmax: int, k0: float, radii, epsilon, positions, precision: str = "c128",
                          rtol: float = 1e-12) -> SphereClusterFactor:
    """Assemble and factor a sphere cluster (cluster/spheres.rs:85-94)."""
    return SphereClusterFactor(IterativeSphereCluster(lmax, k0, radii, epsilon, positions), precision, rtol)


def scattering_gradient(cluster, direction, pol, **options):
    """sigma_sca of one plane wave and its gradients with respect to the radii and the
    positions: through the solve (cluster pullback), the regular overlap R and the
    incident position phases e^{i k.r_j}
    (python/treams_rs/_framework_tmatrix.py:231-289 differentiated).  ``cluster`` is an
    IterativeSphereCluster (GMRES with ``options``) or a SphereClusterFactor (LU)."""
    if isinstance(cluster, SphereClusterFactor):
        solver, c = cluster, cluster.cluster
        a = c.plane_wave(direction, pol)
        x, ctx, conv = solver.record(a)
    else:
        c = cluster
        a = c.plane_wave(direction, pol)
        x, ctx, conv = c.record(a, **options)
    kvec = c.k0 * np.asarray(direction, np.float64) / np.linalg.norm(direction)
    w = 1.0 / c.k0 ** 2
    rx = c.overlap(x)
    rhx = c.overlap(x, adjoint=True)
    sca = float(torch.vdot(x, rx).real) * w
    G = (rx + rhx) * w
    gk, gr, ge, gp, abar, adjoint_conv = ctx.pullback(G)
    gp = gp + c.overlap_pullback(x * w, x)
    ar = a.reshape(c.N, -1)
    ab = abar.reshape(c.N, -1)
    phase = torch.sum(torch.conj(ab) * 1j * ar, dim=1).real.cpu().numpy()
    gp = gp + phase[:, None] * kvec[None, :]
    return sca, gr, gp, dict(k0=gk, epsilon=ge, convergence=conv)


def expansion_field(lmax: int, k0: float, coefficients, points, regular: bool = False,
                    center=(0.0, 0.0, 0.0)) -> torch.Tensor:
    """Electric field of one expansion in helicity vector spherical waves about
    ``center`` (degrees 1 .. lmax, 2 lmax (lmax + 2) modes in the reference order;
    M or M x P coefficients), outgoing or ``regular``, at points (Q x 3): (Q, 3) or
    (Q, 3, P) complex128 (fields.rs:947 for a single-origin SphericalBasis: the
    scattered or incident field of a global T-matrix, sw/expansion.rs:116)."""
    M = int(lib().tg_modes_per_sphere(int(lmax)))
    cm, vector = _columns(coefficients, M)
    dev = _device()
    pts = torch.as_tensor(points, dtype=torch.float64).reshape(-1, 3).to(dev).contiguous()
    q = pts.shape[0]
    c = np.ascontiguousarray(np.asarray(center, np.float64).reshape(3))
    out = torch.empty((cm.shape[0], 3 * q), dtype=torch.complex128, device=dev)
    check(lib().tg_expansion_field(int(lmax), float(k0), c.ctypes.data, cm.data_ptr(), cm.shape[0], q,
                                   pts.data_ptr(), int(regular), out.data_ptr(), ctypes.c_void_p(_stream_ptr())))
    out = out.reshape(cm.shape[0], q, 3).permute(1, 2, 0)
    return out[..., 0].contiguous() if vector else out.contiguous()


def plane_wave_field(k0: float, direction, pol, points) -> torch.Tensor:
    """Field (Q, 3) of a unit helicity plane wave at points (Q x 3): (N + (2 pol - 1) M)
    / sqrt(2) exp(i k.r) (pw/field.rs:267, pw/polarization.rs:138-203)."""
    if isinstance(pol, str):
        pol = {"positive_helicity": 1, "negative_helicity": 0}[pol]
    d = np.ascontiguousarray(np.asarray(direction, np.float64).reshape(3))
    pts = torch.as_tensor(points, dtype=torch.float64).reshape(-1, 3).to(_device()).contiguous()
    out = torch.empty(3 * pts.shape[0], dtype=torch.complex128, device=_device())
    check(lib().tg_plane_wave_field(float(k0), d.ctypes.data, int(pol), pts.shape[0], pts.data_ptr(),
                                    out.data_ptr(), _stream_ptr()))
    return out.reshape(-1, 3)


def plan_terms(lmax: int):
    """Host translation plan of the library (tg_plan_terms): starts, nsame, w, hidx."""
    H = lmax * (lmax + 2)
    count = ctypes.c_int64()
    check(lib().tg_plan_terms(lmax, None, None, None, None, ctypes.byref(count)))
    starts = np.zeros(H * H + 1, np.int32)
    nsame = np.zeros(H * H, np.uint8)
    w = np.zeros(count.value, np.float64)
    hidx = np.zeros(count.value, np.uint16)
    check(lib().tg_plan_terms(lmax, starts.ctypes.data, nsame.ctypes.data, w.ctypes.data, hidx.ctypes.data,
                              ctypes.byref(count)))
    return starts, nsame, w, hidx


def rotation_matrix(lmax: int, phi: float, theta: float = 0.0, psi: float = 0.0) -> np.ndarray:
    """Rotation matrix of z-y-z Euler angles in one spherical basis of degrees
    1 .. lmax (reference mode order): D^l_(m' m)(phi, theta, psi) between equal
    degrees and polarisations (rotation.rs:458-488, diff.rotation)."""
    M = int(lib().tg_modes_per_sphere(int(lmax)))
    out = np.zeros((M, M), np.complex128)
    check(lib().tg_rotation_matrix(int(lmax), float(phi), float(theta), float(psi), out.ctypes.data))
    return out


def rotate_tmatrix(tmatrix, lmax: int, phi: float, theta: float = 0.0, psi: float = 0.0):
    """A single-origin T-matrix rotated by z-y-z Euler angles, R T R^H
    (TMatrix.rotate, python/treams_rs/_framework_tmatrix.py:356-364); the product
    runs on the tensor's device."""
    t = torch.as_tensor(tmatrix)
    r = torch.as_tensor(rotation_matrix(lmax, phi, theta, psi)).to(t.device)
    return r @ t.to(torch.complex128) @ r.conj().T.contiguous()


def wigner3j(j1, j2, j3, m1, m2, m3) -> float:
    return lib().tg_wigner3j(j1, j2, j3, m1, m2, m3)


def translation_blocks_rect(lmax_dst: int, lmax_src: int, k: float, displacements,
                            regular: bool = False) -> torch.Tensor:
    """(count, Md, Ms) translation blocks from a basis of degree lmax_src to one of
    degree lmax_dst (sw/plan.rs:71-170,282-289), [destination, source]."""
    d = torch.as_tensor(displacements, dtype=torch.float64).reshape(-1, 3).to(_device()).contiguous()
    md, ms = int(lib().tg_modes_per_sphere(lmax_dst)), int(lib().tg_modes_per_sphere(lmax_src))
    out = torch.empty((d.shape[0], ms, md), dtype=torch.complex128, device=_device())
    check(lib().tg_translation_blocks_rect(int(lmax_dst), int(lmax_src), float(k), d.shape[0], d.data_ptr(),
                                           int(regular), out.data_ptr(), _stream_ptr()))
    return out.transpose(1, 2).contiguous()


def translation_blocks_tangent(lmax: int, k: float, displacements, tangents, k_tangent: float = 0.0,
                               regular: bool = False) -> torch.Tensor:
    """Directional derivatives of the M x M translation blocks along displacement
    tangents and a wavenumber tangent (TranslationPlan::pushforward, sw/plan.rs:552-595);
    returns (count, M, M) indexed [destination, source]."""
    d = torch.as_tensor(np.ascontiguousarray(np.asarray(displacements, np.float64).reshape(-1, 3)),
                        device=_device())
    t = torch.as_tensor(np.ascontiguousarray(np.asarray(tangents, np.float64).reshape(-1, 3)), device=_device())
    if t.shape != d.shape:
        raise ValueError("one displacement tangent per displacement")
    M = lib().tg_modes_per_sphere(lmax)
    out = torch.empty((d.shape[0], M, M), dtype=torch.complex128, device=d.device)
    check(lib().tg_translation_blocks_tangent(lmax, float(k), d.shape[0], d.data_ptr(), t.data_ptr(),
                                              float(k_tangent), int(regular), out.data_ptr(),
                                              ctypes.c_void_p(_stream_ptr())))
    return out.transpose(1, 2)


def translation_blocks(lmax: int, k: float, displacements, regular: bool = False) -> torch.Tensor:
    """M x M translation blocks (sw/plan.rs:282) for a list of displacements; returns
    (count, M, M) indexed [destination, source]."""
    d = torch.as_tensor(np.ascontiguousarray(np.asarray(displacements, np.float64).reshape(-1, 3)),
                        device=_device())
    M = lib().tg_modes_per_sphere(lmax)
    out = torch.empty((d.shape[0], M, M), dtype=torch.complex128, device=d.device)
    check(lib().tg_translation_blocks(lmax, float(k), d.shape[0], d.data_ptr(), int(regular), out.data_ptr(),
                                      ctypes.c_void_p(_stream_ptr())))
    return out.transpose(1, 2)
