"""Guards of the boundary contract on the GPU.

* Determinism: the reference's "same bits at any thread count" becomes "repeat runs
  bit-identical on one GPU" (fixed-order device reductions, no atomics in any sum):
  solves (FP64, mixed precision with the learned preconditioner, LU) and pullbacks
  repeated on one cluster give identical bits.
* Out of memory: an allocation the device cannot serve raises MemoryError (the
  reference's Error::OutOfMemory, treams-py context.rs:69-74) and leaves the library
  usable.
"""
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


def _cluster(pkg, N=40, lmax=4, seed=3):
    radii, eps, pos = pkg.random_cluster(N, 5.0, 0.1, (0.2, 0.3), None, seed=seed)
    return pkg.IterativeSphereCluster(lmax, 1.3, radii, eps, pos)


def test_repeat_solves_and_pullbacks_are_bit_identical(pkg):
    import treams_rs_amd.learned as learned

    cl = _cluster(pkg)
    learned.attach(cl, learned.MLP.load(learned.DEFAULT_WEIGHTS))
    B = torch.stack([cl.plane_wave([0, 0, 1], 1), cl.plane_wave([1, 0, 0], 0)], dim=1)
    for opts in (dict(rtol=1e-10), dict(rtol=1e-8, mixed_precision="matrix_free", preconditioner="learned"),
                 dict(rtol=1e-10, preconditioner="block_jacobi")):
        runs = []
        for rep in range(2):
            cl.reset()
            x, ctx, conv = cl.record(B, **opts)
            G = x.conj()
            gk, gr, ge, gp, gi, _ = ctx.pullback(G)
            runs.append((x.clone(), gk, gr.copy(), ge.copy(), gp.copy(), gi.clone()))
        a, b = runs
        assert torch.equal(a[0], b[0]), opts
        assert a[1] == b[1] and np.array_equal(a[2], b[2]) and np.array_equal(a[3], b[3]), opts
        assert np.array_equal(a[4], b[4]) and torch.equal(a[5], b[5]), opts
    for prec in ("c128", "c64"):
        xs = []
        for rep in range(2):
            f = pkg.SphereClusterFactor(cl, prec)
            x, _ = f.solve(B)
            xs.append(x.clone())
            f.close()
        assert torch.equal(xs[0], xs[1]), prec


def test_out_of_memory_raises_memory_error(pkg):
    # 4096 spheres at lmax 6: the stored c128 coupling would take 2 (4096 * 48)^2 * 16 B,
    # about 1.24 TB, far beyond the device's 288 GB
    g = np.stack(np.meshgrid(np.arange(16), np.arange(16), np.arange(16), indexing="ij"), -1).reshape(-1, 3)
    cl = pkg.IterativeSphereCluster(6, 1.3, np.full(4096, 0.3), np.full(4096, 2.5 + 0.0j), g * 1.0)
    with pytest.raises(MemoryError):
        cl.assemble("c128")
    del cl
    # the library still works afterwards
    small = _cluster(pkg, N=6, lmax=2)
    x, conv = small.solve(small.plane_wave([0, 0, 1], 1), rtol=1e-10)
    assert conv[0][1] <= 1e-10 * conv[0][2]
