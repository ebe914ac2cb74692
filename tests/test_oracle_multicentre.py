"""The oracle's three- and four-centre quartets, which the reference build (two
centres at most) cannot produce, pinned by what the mathematics fixes:

* relabelling the centres permutes the rows and columns of GDIR/GXCH (the
  FOCKIND order is centre-major inside the large and small blocks,
  bertha-scf.f:2605-2645) and changes nothing else -- a quartet dropped,
  doubled or given the wrong permutation phase in the centre-quadruple loop
  would show up, since relabelling changes which quartets are canonical;
* CLMFAST with INTSYM on (one canonical quartet per permutation class,
  digested by all eleven batches) equals CLMFAST with INTSYM off (every
  quartet computed and digested by batches 1 and 5 only, :11555-11561);
* translating the molecule changes nothing.
"""
import numpy as np
import pytest

from conftest import small_systems


def _blocks(system):
    sizes = [sum(2 * (2 * l + 1) * len(e) for l, e in enumerate(sh)) for sh in system.shells]
    offs = np.concatenate([[0], np.cumsum(sizes)])
    return sizes, offs, int(offs[-1])


def _relabel(bdf, s, order):
    """System with centre j = old centre order[j], and the map from new matrix
    rows to old ones."""
    t = bdf.systems.System(s.name, s.xyz[order], [s.shells[c] for c in order], s.shape)
    sizes, offs, nskp = _blocks(s)
    rows = []
    for half in (0, nskp):
        for c in order:
            rows.append(np.arange(offs[c], offs[c] + sizes[c]) + half)
    return t, np.concatenate(rows)


@pytest.mark.parametrize("name,order", [("tri", [2, 0, 1]), ("quad", [2, 0, 3, 1]),
                                        ("ring4", [3, 1, 0, 2])])
def test_relabelled_centres_permute_the_matrices(bdf, orc, name, order):
    s = small_systems(bdf)[name]
    t, p = _relabel(bdf, s, order)
    d = bdf.systems.random_density(bdf.systems.System(s.name, s.xyz, s.shells, s.shape).ndim,
                                   seed=21)
    g, x = orc.Oracle(s).clmfast(d, 3)
    gt, xt = orc.Oracle(t).clmfast(np.asfortranarray(d[np.ix_(p, p)]), 3)
    assert np.abs(gt - g[np.ix_(p, p)]).max() < 1e-12
    assert np.abs(xt - x[np.ix_(p, p)]).max() < 1e-12


@pytest.mark.parametrize("name", ["tri", "quad", "ring4"])
def test_intsym_on_and_off_agree(orc, bdf, name):
    s = small_systems(bdf)[name]
    d = bdf.systems.random_density(orc.Oracle(s).ndim, seed=22)
    for ilev in (1, 3):
        g1, x1 = orc.Oracle(s, intsym=True).clmfast(d, ilev)
        g0, x0 = orc.Oracle(s, intsym=False).clmfast(d, ilev)
        assert np.abs(g1 - g0).max() < 1e-12, ilev
        assert np.abs(x1 - x0).max() < 1e-12, ilev


def test_translated_four_centres_agree(orc, bdf):
    s = small_systems(bdf)["quad"]
    t = bdf.systems.System(s.name, s.xyz + np.array([-1.7, 0.4, 2.9]), s.shells, s.shape)
    d = bdf.systems.random_density(orc.Oracle(s).ndim, seed=23)
    g, x = orc.Oracle(s).clmfast(d, 3)
    gt, xt = orc.Oracle(t).clmfast(d, 3)
    assert np.abs(gt - g).max() < 1e-11 and np.abs(xt - x).max() < 1e-11


@pytest.mark.parametrize("name", ["diat", "tri", "quad"])
def test_sampled_oracle_build_is_the_build_at_stride_one(bdf, orc, name):
    """bench.py's CPU baseline times orc_clmfast_sample (classes as work
    items, every stride-th primitive pair): at stride 1 it is the oracle's
    CLMFAST build, and a larger stride reports the fraction of ERIFAST calls
    it made."""
    from conftest import small_systems
    s = small_systems(bdf)[name]
    O = orc.Oracle(s)
    d = bdf.systems.random_density(s.ndim, seed=4, real=(s.shape != "NOSYMM"))
    g, x = O.clmfast(d, 3)
    frac, g1, x1 = O.clmfast_sample(d, 3, nthreads=4, stride=1)
    assert frac == 1.0
    assert np.abs(g - g1).max() < 1e-13 and np.abs(x - x1).max() < 1e-13
    frac3, _, _ = O.clmfast_sample(d, 3, nthreads=4, stride=3)
    assert 1.0 / 3.0 - 1e-12 <= frac3 < 0.6


@pytest.mark.parametrize("name", ["diat", "off"])
def test_sampled_oracle_breit_is_the_build_at_stride_one(bdf, orc, name):
    from conftest import small_systems
    s0 = small_systems(bdf)[name]
    s = bdf.systems.System(s0.name, s0.xyz, s0.shells, s0.shape, hmlt="DHFB")
    O = orc.Oracle(s)
    d = bdf.systems.random_density(s.ndim, seed=4, real=(s.shape != "NOSYMM"))
    b, x = O.brtfast(d, 3)
    frac, b1, x1 = O.brtfast_sample(d, 3, nthreads=4, stride=1)
    assert frac == 1.0
    assert np.abs(b - b1).max() < 1e-13 and np.abs(x - x1).max() < 1e-13
    frac3, _, _ = O.brtfast_sample(d, 3, nthreads=4, stride=3)
    assert 1.0 / 3.0 - 1e-12 <= frac3 < 0.6
