"""The C++ host driver (the call sequence of a Rust host over the C ABI) on the
8-sphere lmax-3 cluster: LU and GMRES cross sections and the LU-adjoint cross-section
gradient against the FP64 oracle."""
import json
import os
import subprocess

import numpy as np
import pytest

from __graft_entry__ import load_package

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "treams-rs_amd", "cluster_driver")


def test_host_driver_against_oracle(oracle):
    pkg = load_package()
    assert os.path.exists(DRIVER), "cluster_driver is not built (treams-rs_amd/build.py)"
    lmax, k0 = 3, 1.3
    radii, eps, pos = pkg.random_cluster(8, 2.0, 0.1, (0.2, 0.3), 4 + 0.1j, seed=0)
    lines = "".join(" ".join(f"{float(v):.17g}" for v in (r, e.real, e.imag, *p)) + "\n"
                    for r, e, p in zip(radii, eps, pos))
    out = subprocess.run([DRIVER, str(lmax), repr(k0)], input=lines, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    res = json.loads(out.stdout)
    ref = oracle.Cluster(lmax, k0, radii, eps, pos)
    assert res["n"] == ref.n
    a = ref.plane_wave([0, 0, 1], 1)
    sca, ext = ref.cross_sections(a)
    assert abs(res["lu"]["scattering"] - sca) <= 1e-10 * abs(sca)
    assert abs(res["lu"]["extinction"] - ext) <= 1e-10 * abs(ext)
    assert res["lu"]["relative_residual"] <= 1e-12
    assert abs(res["gmres"]["scattering"] - sca) <= 1e-9 * abs(sca)
    s2, gr, gp = ref.scattering_gradient([0, 0, 1], 1)
    assert abs(s2 - sca) <= 1e-12 * abs(sca)
    assert np.abs(np.array(res["grad_radii"]) - gr).max() <= 1e-9 * np.abs(gr).max()
    assert np.abs(np.array(res["grad_positions"]).reshape(-1, 3) - gp).max() <= 1e-9 * np.abs(gp).max()
