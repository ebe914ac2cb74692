"""HIP Breit path against the oracle (pinned to the reference's BRTFAST,
BIIFAST and EILS tables): LS tables, single BIIFAST batches and whole
BRTFAST builds, including the three-centre system."""
import numpy as np
import pytest

from conftest import small_systems

pytestmark = pytest.mark.gpu

ERI_ABS, ERI_REL = 1e-12, 1e-13
FOCK_ABS = 1e-10
NAMES = ["ne", "diat", "off", "tri", "fsh"]


def _kq(k):
    return k // 2 if k % 2 == 0 else -(k // 2) - 1


@pytest.fixture(scope="module")
def systems(bdf):
    return {k: bdf.systems.System(v.name, v.xyz, v.shells, v.shape, hmlt="DHFB")
            for k, v in small_systems(bdf).items()}


@pytest.mark.parametrize("name", ["ne", "off", "fsh"])
def test_ls_tables_match_oracle(bdf, orc, systems, name):
    s = systems[name]
    print(f"[ls tables] {name}", flush=True)
    O = orc.Oracle(s)
    with bdf.Engine(s) as eng:
        for a in range(1, s.ncnt + 1):
            for b in range(1, s.ncnt + 1):
                for ka in range(1, 2 * s.lmax(a - 1) + 2):
                    for kb in range(1, 2 * s.lmax(b - 1) + 2):
                        for ma in range(1, abs(_kq(ka)) + 1):
                            for mb in range(1, abs(_kq(kb)) + 1):
                                eg = eng.etable(2, a, b, ka, kb, ma, mb)
                                eo = O.etable(2, a, b, ka, kb, ma, mb)
                                assert eg.shape == eo.shape
                                scale = max(np.abs(eo).max(), 1e-300)
                                assert np.abs(eg - eo).max() <= 1e-13 * scale


@pytest.mark.parametrize("name", NAMES)
def test_bii_batches_match_oracle(bdf, orc, systems, name):
    s = systems[name]
    print(f"[biifast] {name}", flush=True)
    O = orc.Oracle(s)
    rng = np.random.default_rng(23)
    with bdf.Engine(s) as eng:
        for _ in range(25):
            icnt = [int(x) for x in rng.integers(1, s.ncnt + 1, size=4)]
            kidx, midx = [], []
            for c in icnt:
                k = int(rng.integers(1, 2 * s.lmax(c - 1) + 2))
                kidx.append(k)
                midx.append(int(rng.integers(1, abs(_kq(k)) + 1)))
            i = int(rng.integers(1, s.nfnc_of(icnt[0] - 1, kidx[0] // 2) + 1))
            j = int(rng.integers(1, s.nfnc_of(icnt[1] - 1, kidx[1] // 2) + 1))
            rg = eng.biifast(icnt, kidx, midx, i, j)
            ro = O.biifast(icnt, kidx, midx, i, j)
            tol = ERI_ABS + ERI_REL * np.abs(ro).max()
            assert np.abs(rg - ro).max() <= tol, (icnt, kidx, midx, i, j)


@pytest.mark.parametrize("name", NAMES)
def test_breit_matches_oracle(bdf, orc, systems, name):
    s = systems[name]
    print(f"[breit] {name}", flush=True)
    d = bdf.systems.random_density(s.ndim, seed=5, real=(s.shape != "NOSYMM"))
    bo, xo = orc.Oracle(s).brtfast(d, 3)
    with bdf.Engine(s) as eng:
        b, x = eng.breit(d, 3)
        b1, x1 = eng.breit(d, 1)
    assert np.abs(xo).max() > 1e-3
    assert np.abs(b - bo).max() < FOCK_ABS
    assert np.abs(x - xo).max() < FOCK_ABS, np.abs(x - xo).max()
    bo1, xo1 = orc.Oracle(s).brtfast(d, 1)
    assert np.abs(x1 - xo1).max() < FOCK_ABS


def test_gaunt_matches_oracle(bdf, orc, systems):
    s = systems["off"]
    d = bdf.systems.random_density(s.ndim, seed=6)
    bo, xo = orc.Oracle(s).brtfast(d, 3, gaunt=True)
    with bdf.Engine(s, gaunt=True) as eng:
        b, x = eng.breit(d, 3)
    assert np.abs(x - xo).max() < FOCK_ABS


@pytest.mark.parametrize("group,per_cu", [(1, 16), (1, 12), (1, 8), (2, 4), (4, 4)])
def test_breit_groups_match_oracle(bdf, orc, systems, group, per_cu):
    """Every AB-combination group size and waves-per-SIMD instance of the
    vector-ALU Breit kernel gives the oracle's build (Breit and Gaunt), on
    the complex three-centre system and the real f-shell system, where
    classes hold many combinations."""
    for name, gaunt in (("tri", False), ("fsh", False), ("off", True)):
        s = systems[name]
        print(f"[breit group] {group} per CU {per_cu} {name} gaunt={gaunt}", flush=True)
        d = bdf.systems.random_density(s.ndim, seed=7, real=(s.shape != "NOSYMM"))
        bo, xo = orc.Oracle(s).brtfast(d, 3, gaunt=gaunt)
        with bdf.Engine(s, tuned=False, gaunt=gaunt) as eng:
            eng.set_launch(breit_group=group, breit_per_cu=per_cu, breit_kernel=1)
            b, x = eng.breit(d, 3)
            # a second build on the same scratch (the exchange blocks are
            # summed by atomics, so only to rounding)
            b2, x2 = eng.breit(d, 3)
        assert np.abs(x - xo).max() < FOCK_ABS, (name, group)
        assert np.abs(b - bo).max() < FOCK_ABS, (name, group)
        assert np.abs(x2 - x).max() < 1e-13 and np.abs(b2 - b).max() < 1e-13


def test_breit_group_rejects_bad_sizes(bdf, systems):
    with bdf.Engine(systems["ne"]) as eng:
        for bad in (3, 5, -1):
            with pytest.raises(bdf.BerthaError):
                eng.set_launch(breit_group=bad)
