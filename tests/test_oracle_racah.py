"""The oracle's COULOMB1 (one-centre Coulomb by Racah algebra, bertha-scf.f
:21540) against the reference's own CLMFAST + COULOMB1 sequence with RACAH1
on (HFSCF 5401-5421), compiled from bertha-scf.f by oracle/Makefile, and
against committed vectors written from it; plus the reference's stale beta
integrals, pinned against the McMurchie-Davidson one-centre integrals."""
import os

import numpy as np
import pytest

from conftest import sd_axial_density, small_systems

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "racah_vectors.npz")


@pytest.fixture(scope="module")
def ref_ok(orc):
    if not orc.reference_available():
        pytest.skip("oracle/_ref/libbertha_ref.so not built (reference sources absent)")
    return True


@pytest.mark.parametrize("name", ["ne", "datom", "fatom", "diat", "off", "fdiat"])
def test_coulomb1_matches_reference(bdf, orc, ref_ok, name):
    """Atoms up to f (CLMMT1Z), a diatomic on z (CLMMT1Z with two centres) and
    off-axis systems (CLMMT1G, complex density)."""
    s = small_systems(bdf)[name]
    d = bdf.systems.random_density(s.ndim, seed=11, real=(s.shape != "NOSYMM"))
    go, xo = orc.Oracle(s, racah1=True).clmfast_racah(d, 3)
    gr, xr = orc.Reference(s, racah1=True).clmfast_racah(d, 3)
    assert np.abs(go - gr).max() < 1e-12
    assert np.abs(xo - xr).max() < 1e-12
    # lower triangles are the conjugate upper ones, exactly (COULOMB1's end)
    for m in (go, xo):
        il = np.tril_indices(s.ndim, -1)
        assert np.array_equal(m[il], np.conj(m.T[il]))


def test_coulomb1_matches_golden_vectors(bdf, orc):
    """Reference outputs committed by tests/golden/make_racah_golden.py."""
    z = np.load(GOLD)
    sy = small_systems(bdf)
    for name in ("datom", "fdiat", "off"):
        s = sy[name]
        d = bdf.systems.random_density(s.ndim, seed=int(z[f"{name}_seed"]),
                                       real=(s.shape != "NOSYMM"))
        go, xo = orc.Oracle(s, racah1=True).clmfast_racah(d, 3)
        r, c = z[f"{name}_rows"], z[f"{name}_cols"]
        assert np.abs(go[r, c] - z[f"{name}_gdir"]).max() < 1e-12, name
        assert np.abs(xo[r, c] - z[f"{name}_gxch"]).max() < 1e-12, name


def test_reference_reads_stale_beta_integrals(bdf, orc, ref_ok):
    """The (pp|sd) block's first KQN combination fails IKL <= IKR, so
    BTACLM1 never runs for it and RKCLM1 reads the beta integrals the
    previous block left in COMMON /BETA/.  On a density that couples s and d
    of equal m only, the exact one-centre integrals (MD, RACAH1 off) and the
    reference's COULOMB1 then differ by O(0.1) in the p-p block of GDIR;
    the oracle reproduces the reference."""
    s = small_systems(bdf)["datom"]
    O = orc.Oracle(s, racah1=True)
    d = sd_axial_density(O.labels(), s.ndim)
    g_md, x_md = orc.Oracle(s).clmfast(d, 3)
    go, xo = O.clmfast_racah(d, 3)
    gr, xr = orc.Reference(s, racah1=True).clmfast_racah(d, 3)
    assert np.abs(go - gr).max() < 1e-12 and np.abs(xo - xr).max() < 1e-12
    assert np.abs(go - g_md).max() > 0.1
    assert np.abs(xo - x_md).max() < 1e-12  # no exchange through the (pp|sd) block here
