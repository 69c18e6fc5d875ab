"""tg_translation_blocks_noalloc (plan staged in the library's fixed device array)
against tg_translation_blocks (plan in an allocated buffer), through the C ABI.

Both run the same kernel on the same plan values, so the blocks must agree bit for
bit.  Two calls in a row, the second with a smaller plan and fewer blocks, check that
the shared array's leftover contents from the first call do not leak into the second.
"""
import ctypes

import numpy as np
import pytest
import torch

from __graft_entry__ import load_package

pytestmark = pytest.mark.gpu


def _blocks(lib, fn, lmax, k, disp, regular):
    M = lib.tg_modes_per_sphere(lmax)
    d = torch.as_tensor(disp, dtype=torch.float64, device="cuda").contiguous()
    out = torch.full((d.shape[0], M, M), complex(np.nan, np.nan), dtype=torch.complex128, device="cuda")
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    st = fn(lmax, float(k), d.shape[0], d.data_ptr(), int(regular), out.data_ptr(), stream)
    return st, out.cpu().numpy()


def test_noalloc_blocks_equal_allocating_blocks():
    pkg = load_package()
    assert torch.cuda.is_available()
    lib = pkg.lib()
    rng = np.random.default_rng(11)
    cases = [(6, 1.3, rng.normal(size=(37, 3)) * 2.0, False),   # the larger plan first
             (3, 0.9, rng.normal(size=(5, 3)), True),             # then a smaller one
             (6, 1.3, rng.normal(size=(3, 3)), True)]
    cases[1][2][0] = 0.0  # a regular block at the origin (identity)
    for lmax, k, disp, regular in cases:
        print(f"noalloc: lmax={lmax} count={len(disp)} regular={regular}", flush=True)
        st_ref, ref = _blocks(lib, lib.tg_translation_blocks, lmax, k, disp, regular)
        st, got = _blocks(lib, lib.tg_translation_blocks_noalloc, lmax, k, disp, regular)
        assert st_ref == 0 and st == 0, lib.tg_last_error()
        assert np.all(np.isfinite(got))
        assert np.array_equal(got, ref)
    # a request the fixed array cannot serve is refused before any device work
    st, _ = _blocks(lib, lib.tg_translation_blocks_noalloc, 13, 1.0, np.zeros((1, 3)), False)
    assert st == 1 and b"lmax" in lib.tg_last_error()
