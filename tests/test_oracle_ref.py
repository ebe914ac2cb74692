"""The oracle against the reference's own routines (compiled from
bertha-scf.f by oracle/Makefile).  Skipped where the reference build is
absent (e.g. on a machine without the reference sources)."""
import numpy as np
import pytest

from conftest import small_systems


@pytest.fixture(scope="module")
def ref_ok(orc):
    if not orc.reference_available():
        pytest.skip("oracle/_ref/libbertha_ref.so not built (reference sources absent)")
    return True


def _blocks(sys_):
    for a in range(1, sys_.ncnt + 1):
        for b in range(1, sys_.ncnt + 1):
            for ka in range(1, 2 * sys_.lmax(a - 1) + 2):
                for kb in range(1, 2 * sys_.lmax(b - 1) + 2):
                    kqa = ka // 2 if ka % 2 == 0 else -(ka // 2) - 1
                    kqb = kb // 2 if kb % 2 == 0 else -(kb // 2) - 1
                    for ma in range(1, abs(kqa) + 1):
                        for mb in range(1, abs(kqb) + 1):
                            yield a, b, ka, kb, ma, mb


@pytest.mark.parametrize("name", ["ne", "diat", "off"])
def test_etables_match_reference(bdf, orc, ref_ok, name):
    s = small_systems(bdf)[name]
    R = orc.Reference(s)
    O = orc.Oracle(s)
    for itt in (1, 4):
        for blk in _blocks(s):
            er = R.etable(itt, *blk)
            eo = O.etable(itt, *blk)
            assert er.shape == eo.shape
            scale = max(np.abs(er).max(), 1e-300)
            assert np.abs(er - eo).max() <= 1e-13 * scale, (itt, blk)


@pytest.mark.parametrize("name", ["ne", "diat", "off"])
def test_clmfast_matches_reference(bdf, orc, ref_ok, name):
    s = small_systems(bdf)[name]
    R = orc.Reference(s)
    O = orc.Oracle(s)
    d = bdf.systems.random_density(R.ndim, seed=3, real=(s.shape != "NOSYMM"))
    # stage 1 first: the reference's SS sparsity lists are then still those
    # of a fresh process (see the stage-1 note in bdf_oracle.c)
    for ilev in (1, 3, 2):
        gr, xr = R.clmfast(d, ilev)
        go, xo = O.clmfast(d, ilev)
        assert np.abs(gr - go).max() < 1e-12, ilev
        assert np.abs(xr - xo).max() < 1e-12, ilev


def test_clmfast_racah1_skips_one_centre(bdf, orc, ref_ok):
    s = small_systems(bdf)["diat"]
    R = orc.Reference(s, racah1=True)
    O = orc.Oracle(s, racah1=True)
    d = bdf.systems.random_density(R.ndim, seed=5, real=True)
    gr, xr = R.clmfast(d, 3)
    go, xo = O.clmfast(d, 3)
    assert np.abs(gr - go).max() < 1e-12
    assert np.abs(xr - xo).max() < 1e-12


@pytest.mark.parametrize("name", ["diat", "off"])
def test_clmfast_intsym_off_matches_reference(bdf, orc, ref_ok, name):
    """INTSYM off: every quartet digested by batches 1 and 5 and no Hermitian
    completion (:12174-12177); equal to the INTSYM-on build."""
    s = small_systems(bdf)[name]
    R = orc.Reference(s, intsym=False)
    O = orc.Oracle(s, intsym=False)
    d = bdf.systems.random_density(R.ndim, seed=12, real=(s.shape != "NOSYMM"))
    gr, xr = R.clmfast(d, 3)
    go, xo = O.clmfast(d, 3)
    assert np.abs(gr - go).max() < 1e-12
    assert np.abs(xr - xo).max() < 1e-12
    g1, x1 = orc.Oracle(s).clmfast(d, 3)
    assert np.abs(g1 - go).max() < 1e-12 and np.abs(x1 - xo).max() < 1e-12


@pytest.mark.parametrize("name", ["ne", "off"])
def test_erifast_batches_match_reference(bdf, orc, ref_ok, name):
    s = small_systems(bdf)[name]
    R = orc.Reference(s)
    O = orc.Oracle(s)
    rng = np.random.default_rng(11)
    nc = s.ncnt
    count = 0
    for _ in range(60):
        icnt = list(rng.integers(1, nc + 1, size=4))
        kidx, midx = [], []
        for c in icnt:
            k = int(rng.integers(1, 2 * s.lmax(c - 1) + 2))
            kq = k // 2 if k % 2 == 0 else -(k // 2) - 1
            kidx.append(k)
            midx.append(int(rng.integers(1, abs(kq) + 1)))
        la, lb = kidx[0] // 2, kidx[1] // 2
        ibas = int(rng.integers(1, s.nfnc_of(icnt[0] - 1, la) + 1))
        jbas = int(rng.integers(1, s.nfnc_of(icnt[1] - 1, lb) + 1))
        itn = [int(rng.choice([1, 4])), int(rng.choice([1, 4]))]
        rr = R.erifast(icnt, kidx, midx, ibas, jbas, itn)
        ro = O.erifast(icnt, kidx, midx, ibas, jbas, itn)
        # the real (linear/atomic) path leaves RR 2,3,5,8,9,12,14,15 unset
        use = [0, 3, 5, 6, 9, 10, 12, 15] if s.shape != "NOSYMM" else list(range(16))
        assert np.abs(rr[use] - ro[use]).max() <= 1e-12 + 1e-13 * np.abs(rr[use]).max()
        count += 1
    assert count == 60


def test_boys_and_rmake_match_reference(orc, ref_ok, bdf):
    s = small_systems(bdf)["ne"]
    R = orc.Reference(s)
    rng = np.random.default_rng(2)
    for itype, lo, hi in ((1, 0.0, 1e-11), (2, 1e-6, 17.0), (3, 17.0, 30.0), (4, 30.0, 400.0)):
        t = rng.uniform(lo, hi, size=40)
        for lam in (0, 4, 16):
            fr = R.funfm(t, lam, itype)
            fo = orc.funfm(t, lam, itype)
            assert np.allclose(fr, fo, rtol=2e-15, atol=0.0)
    qp = rng.uniform(-2, 2, size=(30, 3))
    aph = rng.uniform(0.05, 50.0, size=30)
    for lam in (0, 3, 8, 13):
        rr = R.rmake(qp, aph, lam)
        ro = orc.rmake(qp, aph, lam)
        assert np.allclose(rr, ro, rtol=1e-13, atol=1e-300 + 1e-14 * np.abs(rr).max())


def test_labels_match_reference(bdf, orc, ref_ok):
    for name in ("ne", "diat", "off"):
        s = small_systems(bdf)[name]
        R = orc.Reference(s)
        O = orc.Oracle(s)
        for a, b in zip(R.labels(), O.labels()):
            assert np.array_equal(a, b)
