"""HIP one-centre Racah builds (bdf_coulomb1 and bdf_breit1, the reference's
COULOMB1 and BREIT1 per centre, bertha-scf.f:21540 and :26014, called at
5415-5421 and 5447-5453) against the oracle, which is pinned to the
reference's CLMFAST + COULOMB1 and BRTFAST + BREIT1 sequences."""
import numpy as np
import pytest

from conftest import sd_axial_density, small_systems

pytestmark = pytest.mark.gpu

FOCK_ABS = 1e-10


@pytest.mark.parametrize("name", ["ne", "datom", "fatom", "diat", "off", "fdiat", "tri", "ring4"])
def test_coulomb1_matches_oracle(bdf, orc, name):
    s = small_systems(bdf)[name]
    print(f"[coulomb1] {name}", flush=True)
    d = bdf.systems.random_density(s.ndim, seed=13, real=(s.shape != "NOSYMM"))
    O = orc.Oracle(s, racah1=True)
    with bdf.Engine(s, racah1=True) as eng:
        for ilev in (3, 1):
            g, x = eng.clmfast_racah(d, ilev)
            go, xo = O.clmfast_racah(d, ilev)
            assert np.abs(g - go).max() < FOCK_ABS, (name, ilev, np.abs(g - go).max())
            assert np.abs(x - xo).max() < FOCK_ABS, (name, ilev, np.abs(x - xo).max())
            il = np.tril_indices(s.ndim, -1)
            for m in (g, x):
                assert np.array_equal(m[il], np.conj(m.T[il]))


def test_coulomb1_device_matches_host(bdf, orc):
    import torch
    s = small_systems(bdf)["off"]
    n = s.ndim
    d = bdf.systems.random_density(n, seed=14)
    with bdf.Engine(s, racah1=True) as eng:
        g, x = eng.clmfast_racah(d, 3)
        dt = torch.tensor(np.ascontiguousarray(d.T), device="cuda")
        gt = torch.zeros((n, n), dtype=torch.complex128, device="cuda")
        xt = torch.zeros_like(gt)
        eng.clmfast_device(dt.data_ptr(), gt.data_ptr(), xt.data_ptr(), 3)
        eng.coulomb1_device(dt.data_ptr(), gt.data_ptr(), xt.data_ptr())
        torch.cuda.synchronize()
    assert np.abs(gt.cpu().numpy().T - g).max() < 1e-12
    assert np.abs(xt.cpu().numpy().T - x).max() < 1e-12


def test_exact_beta_matches_md_on_axial_density(bdf, orc):
    """exact_beta = 1 evaluates the (pp|sd) block's own beta integrals: on a
    density coupling s and d of equal m, the one-centre Coulomb matrices then
    equal the McMurchie-Davidson ones (RACAH1 off); exact_beta = 0 keeps the
    reference's stale integrals (pinned by the oracle)."""
    s = small_systems(bdf)["datom"]
    d = sd_axial_density(orc.Oracle(s).labels(), s.ndim)
    with bdf.Engine(s) as md:
        g_md, x_md = md.clmfast(d, 3)
    with bdf.Engine(s, racah1=True) as eng:
        g_ex, x_ex = eng.clmfast_racah(d, 3, exact_beta=True)
        g_rf, x_rf = eng.clmfast_racah(d, 3)
    go, xo = orc.Oracle(s, racah1=True).clmfast_racah(d, 3)
    assert np.abs(g_ex - g_md).max() < FOCK_ABS and np.abs(x_ex - x_md).max() < FOCK_ABS
    assert np.abs(g_rf - go).max() < FOCK_ABS and np.abs(x_rf - xo).max() < FOCK_ABS
    assert np.abs(g_rf - g_md).max() > 0.1


@pytest.mark.parametrize("name", ["ne", "water", "tlf", "ring"])
def test_workload_racah_reads_only_written_betas(bdf, name):
    """Every beta integral the workloads' stale blocks read was written by an
    earlier block of the same build, so the reference's values do not depend
    on what an earlier build or process left in COMMON /BETA/."""
    kw = {} if name == "ne" else {"seed": 0}
    s = bdf.systems.config_system(name, **kw)
    with bdf.Engine(s, racah1=True) as eng:
        terms, items, uncovered = eng.racah_info()
    print(f"[racah plan] {name}: {terms:.3e} terms, {items} items", flush=True)
    assert items > 0 and uncovered == 0


def _dhfb(bdf, s):
    return bdf.systems.System(s.name, s.xyz, s.shells, s.shape, hmlt="DHFB")


@pytest.mark.parametrize("name", ["ne", "datom", "fatom", "diat", "off", "fdiat", "tri"])
def test_breit1_matches_oracle(bdf, orc, name):
    s = _dhfb(bdf, small_systems(bdf)[name])
    print(f"[breit1] {name}", flush=True)
    d = bdf.systems.random_density(s.ndim, seed=15, real=(s.shape != "NOSYMM"))
    bo, xo = orc.Oracle(s, racah1=True).brtfast_racah(d, 3)
    with bdf.Engine(s, racah1=True) as eng:
        b, x = eng.breit_racah(d, 3)
        assert eng.racah_info(breit=True)[2] == 0
    assert np.abs(b).max() == 0.0
    assert np.abs(x - xo).max() < FOCK_ABS, (name, np.abs(x - xo).max())
    il = np.tril_indices(s.ndim, -1)
    assert np.array_equal(x[il], np.conj(x.T[il]))


def test_breit1_exact_beta_on_axial_density(bdf, orc):
    """exact_beta for BREIT1 on the s-d density: recorded against the MD
    one-centre Breit build (the oracle pins exact_beta = 0)."""
    s = _dhfb(bdf, small_systems(bdf)["datom"])
    d = sd_axial_density(orc.Oracle(s).labels(), s.ndim)
    with bdf.Engine(s) as md:
        _, x_md = md.breit(d, 3)
    with bdf.Engine(s, racah1=True) as eng:
        _, x_ex = eng.breit_racah(d, 3, exact_beta=True)
        _, x_rf = eng.breit_racah(d, 3)
    _, xo = orc.Oracle(s, racah1=True).brtfast_racah(d, 3)
    print(f"[breit1 sd] exact vs MD {np.abs(x_ex - x_md).max():.3e}, reference-faithful vs MD "
          f"{np.abs(x_rf - x_md).max():.3e}", flush=True)
    assert np.abs(x_rf - xo).max() < FOCK_ABS


def test_tlf_breit1_reads_only_written_betas(bdf):
    s = bdf.systems.config_system("tlf", seed=0, hmlt="DHFB")
    with bdf.Engine(s, racah1=True) as eng:
        terms, items, uncovered = eng.racah_info(breit=True)
    print(f"[racah plan] tlf breit: {terms:.3e} terms, {items} items", flush=True)
    assert items > 0 and uncovered == 0
