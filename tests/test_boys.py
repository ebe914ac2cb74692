"""FUNFM pinned by mathematics, independently of the reference: closed forms,
the recurrence identity and the incomplete gamma function."""
import numpy as np
import pytest
from scipy.special import gamma, gammainc


def boys_exact(m, t):
    t = np.asarray(t, dtype=float)
    return gamma(m + 0.5) * gammainc(m + 0.5, t) / (2.0 * t ** (m + 0.5))


def test_branch1_is_the_t0_limit(orc):
    f = orc.funfm(np.zeros(3), 16, 1)
    assert np.allclose(f[:, 0], 1.0 / (2 * np.arange(17) + 1), rtol=0, atol=0)


@pytest.mark.parametrize("itype,lo,hi,rtol", [(2, 1e-3, 17.0, 2e-14), (3, 17.0, 30.0, 1e-12)])
def test_branches_2_3_match_incomplete_gamma(orc, itype, lo, hi, rtol):
    t = np.linspace(lo, hi, 37)
    f = orc.funfm(t, 16, itype)
    for m in range(17):
        ex = boys_exact(m, t)
        assert np.max(np.abs(f[m] - ex) / ex) < rtol * (1 + m), m


def test_downward_recursion_identity(orc):
    # (2m+1) F_m = 2T F_{m+1} + exp(-T)
    t = np.linspace(0.5, 16.5, 9)
    f = orc.funfm(t, 10, 2)
    for m in range(10):
        lhs = (2 * m + 1) * f[m]
        rhs = 2 * t * f[m + 1] + np.exp(-t)
        assert np.allclose(lhs, rhs, rtol=1e-14)


def test_branch4_is_the_pure_asymptotic_series(orc):
    # the reference drops every exp(-T) term above T = 30 (bertha-scf.f:32590-32605):
    # F_m = (2m-1)!! sqrt(pi) / (2^(m+1) T^(m+1/2)); its error vs the exact
    # function grows with m (about 3e-3 relative at m=16, T=30)
    t = np.array([30.5, 45.0, 120.0])
    f = orc.funfm(t, 16, 4)
    for m in range(17):
        dfact = np.prod(np.arange(2 * m - 1, 0, -2)) if m > 0 else 1.0
        asym = dfact * np.sqrt(np.pi) / (2 ** (m + 1) * t ** (m + 0.5))
        assert np.allclose(f[m], asym, rtol=1e-13)
    err = abs(f[16, 0] - boys_exact(16, t[0])) / boys_exact(16, t[0])
    assert 1e-4 < err < 1e-2
