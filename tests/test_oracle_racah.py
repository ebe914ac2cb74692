"""The oracle's COULOMB1 and BREIT1 (one-centre Coulomb and Breit by Racah
algebra, bertha-scf.f:21540, :26014) against the reference's own CLMFAST +
COULOMB1 and BRTFAST + BREIT1 sequences with RACAH1 on (HFSCF 5401-5453),
compiled from bertha-scf.f by oracle/Makefile, and against committed vectors
written from it; plus the reference's stale beta integrals, pinned against
the McMurchie-Davidson one-centre integrals."""
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


@pytest.mark.parametrize("name", ["ne", "datom", "fatom", "diat", "off", "fdiat"])
def test_breit1_matches_reference(bdf, orc, ref_ok, name):
    b = small_systems(bdf)[name]
    s = bdf.systems.System(b.name, b.xyz, b.shells, b.shape, hmlt="DHFB")
    d = bdf.systems.random_density(s.ndim, seed=12, real=(s.shape != "NOSYMM"))
    bo, xo = orc.Oracle(s, racah1=True).brtfast_racah(d, 3)
    br, xr = orc.Reference(s, racah1=True).brtfast_racah(d, 3)
    assert np.abs(bo - br).max() == 0.0  # closed shells: BDIR untouched
    assert np.abs(xo - xr).max() < 1e-12
    il = np.tril_indices(s.ndim, -1)
    assert np.array_equal(xo[il], np.conj(xo.T[il]))
    # the one-centre part is not small
    b0, x0 = orc.Oracle(s, racah1=True).brtfast(d, 3)
    assert np.abs(xo - x0).max() > 1e-2


def test_breit1_matches_golden_vectors(bdf, orc):
    z = np.load(GOLD)
    sy = small_systems(bdf)
    for name in ("datom", "fdiat", "off"):
        b = sy[name]
        s = bdf.systems.System(b.name, b.xyz, b.shells, b.shape, hmlt="DHFB")
        d = bdf.systems.random_density(s.ndim, seed=int(z[f"{name}_seed"]),
                                       real=(s.shape != "NOSYMM"))
        bo, xo = orc.Oracle(s, racah1=True).brtfast_racah(d, 3)
        r, c = z[f"{name}_rows"], z[f"{name}_cols"]
        assert np.abs(xo[r, c] - z[f"{name}_bxch"]).max() < 1e-12, name


def test_racah_equals_md_on_spherical_density(bdf, orc):
    """On the converged Ne density (a closed-shell atom: spherical), the
    one-centre Racah builds equal the McMurchie-Davidson one-centre integrals
    of CLMFAST and BRTFAST with RACAH1 off, for Coulomb and Breit alike."""
    import importlib

    from scf_backends import OracleBackend
    scf = importlib.import_module(bdf.__name__ + ".scf")
    s = bdf.systems.ne_atom()
    d = scf.run_scf(s, 10, OracleBackend(bdf, orc, s)).density.numpy()
    g_md, x_md = orc.Oracle(s).clmfast(d, 3)
    g_rc, x_rc = orc.Oracle(s, racah1=True).clmfast_racah(d, 3)
    assert np.abs(g_md - g_rc).max() < 1e-11 and np.abs(x_md - x_rc).max() < 1e-11
    sb = bdf.systems.System(s.name, s.xyz, s.shells, s.shape, hmlt="DHFB")
    _, b_md = orc.Oracle(sb).brtfast(d, 3)
    _, b_rc = orc.Oracle(sb, racah1=True).brtfast_racah(d, 3)
    assert np.abs(b_md - b_rc).max() < 1e-11
    assert abs(np.real(np.sum(d * (b_md - b_rc)))) < 1e-12
