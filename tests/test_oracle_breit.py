"""Breit restatement in the oracle against the reference's own routines
(compiled from bertha-scf.f into oracle/_ref): EILS tables (EILSGN/EQLS),
single BIIFAST batches (Breit and Gaunt) and whole BRTFAST builds, on an
atom (ATOMIC), a diatomic on z (DIATOM) and an off-axis two-centre system
(NOSYMM).  One reference set-up per system (each takes seconds); the
f-shell system is pinned by the committed golden vectors instead."""
import numpy as np
import pytest

from conftest import small_systems


def _kq(k):
    return k // 2 if k % 2 == 0 else -(k // 2) - 1


@pytest.mark.parametrize("name", ["ne", "diat", "off"])
def test_breit_matches_reference(bdf, orc, name):
    if not orc.reference_available():
        pytest.skip("oracle/_ref/libbertha_ref.so not built (reference sources absent)")
    s0 = small_systems(bdf)[name]
    s = bdf.systems.System(s0.name, s0.xyz, s0.shells, s0.shape, hmlt="DHFB")
    R, O = orc.Reference(s), orc.Oracle(s)
    # the reference's fresh-process order: a stage-1 Coulomb build first
    d = bdf.systems.random_density(O.ndim, seed=7, real=(s.shape != "NOSYMM"))
    R.clmfast(d, 1)
    br, xr = R.brtfast(d, 3)
    bo, xo = O.brtfast(d, 3)
    assert np.abs(xr).max() > 1e-3
    assert np.abs(br - bo).max() < 1e-13
    assert np.abs(xr - xo).max() < 1e-13
    # EILS tables
    for a in range(1, s.ncnt + 1):
        for b in range(1, s.ncnt + 1):
            for ka in range(1, 2 * s.lmax(a - 1) + 2):
                for kb in range(1, 2 * s.lmax(b - 1) + 2):
                    for ma in range(1, abs(_kq(ka)) + 1):
                        for mb in range(1, abs(_kq(kb)) + 1):
                            er = R.etable(2, a, b, ka, kb, ma, mb)
                            eo = O.etable(2, a, b, ka, kb, ma, mb)
                            scale = max(np.abs(er).max(), 1e-300)
                            assert np.abs(er - eo).max() <= 1e-14 * scale
    # single BIIFAST batches, Breit and Gaunt
    rng = np.random.default_rng(3)
    for gaunt in (False, True):
        for _ in range(10):
            icnt = [int(x) for x in rng.integers(1, s.ncnt + 1, size=4)]
            kidx, midx = [], []
            for c in icnt:
                k = int(rng.integers(1, 2 * s.lmax(c - 1) + 2))
                kidx.append(k)
                midx.append(int(rng.integers(1, abs(_kq(k)) + 1)))
            i = int(rng.integers(1, s.nfnc_of(icnt[0] - 1, kidx[0] // 2) + 1))
            j = int(rng.integers(1, s.nfnc_of(icnt[1] - 1, kidx[1] // 2) + 1))
            rr = R.biifast(icnt, kidx, midx, i, j, gaunt)
            ro = O.biifast(icnt, kidx, midx, i, j, gaunt)
            assert np.abs(rr - ro).max() <= 1e-13 * max(1.0, np.abs(rr).max())


def test_brtfast_threads_agree(bdf, orc):
    s0 = small_systems(bdf)["tri"]
    s = bdf.systems.System(s0.name, s0.xyz, s0.shells, s0.shape, hmlt="DHFB")
    O = orc.Oracle(s)
    d = bdf.systems.random_density(O.ndim, seed=2)
    b1, x1 = O.brtfast(d, 3)
    b4, x4 = O.brtfast(d, 3, nthreads=4)
    assert np.abs(x1).max() > 1e-3
    assert np.abs(x1 - x4).max() < 1e-13 and np.abs(b1 - b4).max() < 1e-13
