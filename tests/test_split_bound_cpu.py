"""The upper bound behind the split-FP16 tables' row shifts, checked on the CPU.

The split tables scale each row (harmonic (p, q), destination sphere i) by a power of
two taken from an upper bound of the row's entries |h_p(k d_ij) P_p^q(cos theta_ij)|
instead of their maximum:

    |h_p(k d_ij)| <= |h_p(k d_min(i))|          (|h_p(x)| decreases in x)
    |P_p^q(x)|    <= sqrt((p + q)! / (p - q)!)  (addition theorem, |x| <= 1)

with the reflection factor (p - q)! / (p + q)! for negative orders.  scipy's spherical
Bessel functions and associated Legendre functions give the entries here; a bound that
fails for any pair would let an FP16 hi half overflow.
"""
import math

import numpy as np
import pytest
from scipy import special


def _hankel_abs(p, x):
    return np.hypot(special.spherical_jn(p, x), special.spherical_yn(p, x))


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_row_bound_holds_for_random_clusters(seed):
    rng = np.random.default_rng(seed)
    N, L2, k = 24, 12, 1.3
    pos = rng.uniform(0, 4.0, size=(N, 3))
    for i in range(N):
        d = pos[i] - np.delete(pos, i, axis=0)
        r = np.linalg.norm(d, axis=1)
        ct = d[:, 2] / r
        for p in range(L2 + 1):
            hmin = _hankel_abs(p, k * r.min())
            assert np.all(_hankel_abs(p, k * r) <= hmin * (1 + 1e-12))
            for q in range(p + 1):
                fr = math.factorial(p + q) / math.factorial(p - q)
                entries = _hankel_abs(p, k * r) * np.abs(special.lpmv(q, p, ct))
                assert np.all(entries <= hmin * math.sqrt(fr) * (1 + 1e-10)), (i, p, q)
                # negative order: the reflection factor times the same magnitude
                assert np.all(entries / fr <= hmin / math.sqrt(fr) * (1 + 1e-10)), (i, p, -q)


def test_legendre_bound_is_attained_within_a_few_bits():
    """The harmonic part of the bound is tight to a factor below 2^2 for every (p, q)
    up to degree 24, so the shift loses at most a couple of the 17 full-precision
    binades."""
    x = np.linspace(-1, 1, 20001)
    for p in range(25):
        for q in range(p + 1):
            fr = math.factorial(p + q) / math.factorial(p - q)
            peak = np.abs(special.lpmv(q, p, x)).max()
            assert peak <= math.sqrt(fr) * (1 + 1e-12)
            assert peak >= math.sqrt(fr) / 4, (p, q)
