"""Schwarz tables and screening in the oracle.

GDSC (CLMDIAG, bertha-scf.f:38126) is pinned to the reference's own integral
code: every entry is sqrt|RR| of the reference's ERIFAST (:12269) for the
self-pair quartet (A B|A B) at kl = ij.  (The reference's CLMDIAG itself
cannot produce it in a fresh process: ERI loops N = 1..MAXN with MAXN of
/ISCR/ never set before the first SCHWARZ call, so GDSC stays zero.)

The (ab|ba) bound table is pinned by the Cauchy-Schwarz inequality of the
Coulomb metric on random ERIFAST batches, which it must satisfy and saturate.

The SCHWARZ masks (:37889) are pinned to the reference's own SCHWARZ,
called through oracle/ref/ref_driver.f with the pinned GDSC injected as its
bound-table argument (the reference's CLMDIAG cannot fill it, above): for
random block quartets, primitive pairs, densities and SENS from 1e-12 to 1,
IMTX, ISCR and MAXN equal the oracle's masks bit for bit.  Monotonic growth
of the dropped set with SENS is checked on whole builds as well.
"""
import numpy as np
import pytest

from conftest import small_systems


@pytest.fixture(scope="module")
def systems(bdf):
    return small_systems(bdf)


def _block_start(labels, c, kq, mj):
    icn, kqn, mqn = labels
    return int(np.where((icn == c) & (kqn == kq) & (mqn == mj))[0][0])


def _random_block(rng, s, n):
    out = []
    for _ in range(n):
        c = int(rng.integers(1, s.ncnt + 1))
        k = int(rng.integers(1, 2 * s.lmax(c - 1) + 2))
        kq = k // 2 if k % 2 == 0 else -(k // 2) - 1
        out.append((c, k, kq, int(rng.integers(1, abs(kq) + 1))))
    return out


@pytest.mark.parametrize("name", ["ne", "diat", "off", "fsh"])
def test_gdsc_matches_reference_erifast(orc, systems, name):
    if not orc.reference_available():
        pytest.skip("oracle/_ref not built")
    s = systems[name]
    R = orc.Reference(s)
    O = orc.Oracle(s)
    gd = O.clmdiag()
    lab = O.labels()
    rng = np.random.default_rng(5)
    worst = 0.0
    for _ in range(16):
        (a, ka, kqa, ma), (b, kb, kqb, mb) = _random_block(rng, s, 2)
        t1, t2 = int(rng.choice([1, 4])), int(rng.choice([1, 4]))
        na, nb = s.nfnc_of(a - 1, ka // 2), s.nfnc_of(b - 1, kb // 2)
        i, j = int(rng.integers(1, na + 1)), int(rng.integers(1, nb + 1))
        rr = R.erifast([a, b, a, b], [ka, kb, ka, kb], [ma, mb, ma, mb], i, j, [t1, t2], 3)
        m = (i - 1) * nb + (j - 1)
        r0 = [_block_start(lab, a, kqa, -(2 * ma - 1)), _block_start(lab, a, kqa, 2 * ma - 1)]
        c0 = [_block_start(lab, b, kqb, -(2 * mb - 1)), _block_start(lab, b, kqb, 2 * mb - 1)]
        ra = [x + (O.nskp if t1 == 4 else 0) + i - 1 for x in r0]
        cb = [x + (O.nskp if t2 == 4 else 0) + j - 1 for x in c0]
        for (sa, sb), n in {(0, 0): 0, (0, 1): 5, (1, 0): 10, (1, 1): 15}.items():
            ref = abs(rr[n, m])
            worst = max(worst, abs(gd[ra[sa], cb[sb]] ** 2 - ref))
    # compared as integrals (GDSC^2): ERIFAST's 1e-10 E(CD) screening leaves
    # differences far below the Fock tolerance
    assert worst < 1e-11


@pytest.mark.parametrize("name", ["ne", "diat", "off", "tri", "fsh"])
def test_pair_bound_is_cauchy_schwarz(orc, systems, name):
    s = systems[name]
    O = orc.Oracle(s)
    Q = O.pairbound()
    lab = O.labels()
    rng = np.random.default_rng(1)
    worst = 0.0
    for _ in range(25):
        blk = _random_block(rng, s, 4)
        icnt = [b[0] for b in blk]
        kidx = [b[1] for b in blk]
        midx = [b[3] for b in blk]
        nf = [s.nfnc_of(b[0] - 1, b[1] // 2) for b in blk]
        i, j = int(rng.integers(1, nf[0] + 1)), int(rng.integers(1, nf[1] + 1))
        itn = [int(rng.choice([1, 4])), int(rng.choice([1, 4]))]
        rr = O.erifast(icnt, kidx, midx, i, j, itn, 3)
        off = []
        for k, (c, _, kq, mq) in enumerate(blk):
            base = O.nskp if itn[k // 2] == 4 else 0
            off.append([_block_start(lab, c, kq, -(2 * mq - 1)) + base,
                        _block_start(lab, c, kq, 2 * mq - 1) + base])
        for n in range(16):
            sa, sb, sc, sd = (n >> 3) & 1, (n >> 2) & 1, (n >> 1) & 1, n & 1
            qab = Q[off[0][sa] + i - 1, off[1][sb] + j - 1]
            qcd = Q[np.ix_(np.array(off[2][sc]) + np.arange(nf[2]),
                           np.array(off[3][sd]) + np.arange(nf[3]))].reshape(-1)
            r = np.abs(rr[n])
            ok = r > 1e-14
            if ok.any():
                worst = max(worst, float((r[ok] / (qab * qcd[ok])).max()))
    assert worst <= 1.0 + 1e-10
    assert worst > 0.9  # the bound is reached (quartets of the (ab|ba) form)


@pytest.mark.parametrize("name", ["diat", "tri"])
def test_schwarz_mask_limits(bdf, orc, systems, name):
    s = systems[name]
    O = orc.Oracle(s)
    d = bdf.systems.random_density(O.ndim, seed=3, real=(s.shape != "NOSYMM"))
    g0, x0 = O.clmfast(d, 3)
    # screening off and screening on with nothing dropped agree exactly;
    # GDSC entries that vanish (real parts of (-+|-+) on the ISYM path) make
    # XLRG = 0, so some kl drop even at the smallest SENS
    g1, x1 = O.clmfast(d, 3, schwarz=False)
    assert np.array_equal(g0, g1) and np.array_equal(x0, x1)
    counts = []
    for sens in (1e-300, 1e-10, 1e-6, 1e-2):
        O.clmfast(d, 3, schwarz=True, sens=sens)
        counts.append(O.schwarz_stats()[1])
    assert counts == sorted(counts) and counts[-1] > counts[0]


@pytest.mark.parametrize("name", ["diat", "tri", "quad", "ring4"])
def test_rigorous_masks_drop_nothing_above_sens(bdf, orc, name):
    """Mode 2 bounds every group by Q(A,B) Q(C,D) >= |(ab|cd)|: no dropped
    digestion term exceeds SENS, while the reference's GDSC masks drop terms
    many orders above it (GDSC is no bound for complex spinor quartets);
    and the rigorous build stays within (dropped terms) x SENS of the
    unscreened one."""
    s = small_systems(bdf)[name]
    O = orc.Oracle(s)
    d = bdf.systems.random_density(O.ndim, seed=3, real=(s.shape != "NOSYMM"))
    g0, x0 = O.clmfast(d, 3)
    counts = []
    for sens in (1e-12, 1e-6, 1e-3):
        g2, x2 = O.clmfast(d, 3, schwarz="rigorous", sens=sens)
        drop_max, drop_count = O.schwarz_stats()
        assert drop_max <= sens
        assert max(np.abs(g2 - g0).max(), np.abs(x2 - x0).max()) <= drop_count * sens
        counts.append(drop_count)
    assert counts == sorted(counts) and counts[-1] > 0
    O.clmfast(d, 3, schwarz=True, sens=1e-12)
    assert O.schwarz_stats()[0] > 1e-6


@pytest.mark.parametrize("name", ["ne", "diat", "off", "fsh"])
def test_schwarz_masks_match_reference(bdf, orc, systems, name):
    if not orc.reference_available():
        pytest.skip("oracle/_ref not built")
    s = systems[name]
    R = orc.Reference(s)
    O = orc.Oracle(s)
    g = O.clmdiag()  # pinned to the reference's ERIFAST (test above)
    lab = O.labels()
    rng = np.random.default_rng(11)
    real = s.shape != "NOSYMM"
    d = bdf.systems.random_density(O.ndim, seed=8, real=real)
    # elements spread over decades, so that every group of the decision
    # tree decides somewhere in the SENS sweep
    d = d * np.exp(rng.uniform(-12.0, 0.0, size=d.shape))
    d = np.asfortranarray(0.5 * (d + d.conj().T))
    seen_partial = seen_drop = seen_keep = 0
    for _ in range(24):
        blk = _random_block(rng, s, 4)
        itn = [int(rng.choice([1, 4])), int(rng.choice([1, 4]))]
        nlab = []
        for k, (c, _, kq, mq) in enumerate(blk):
            base = O.nskp if itn[k // 2] == 4 else 0
            nlab += [_block_start(lab, c, kq, -(2 * mq - 1)) + base,
                     _block_start(lab, c, kq, 2 * mq - 1) + base]
        nf = [s.nfnc_of(b[0] - 1, b[1] // 2) for b in blk]
        i, j = int(rng.integers(1, nf[0] + 1)), int(rng.integers(1, nf[1] + 1))
        for sens in (1e-12, 1e-9, 1e-6, 1e-4, 1e-2, 1.0):
            imtx, iscr, maxn = R.schwarz_masks(g, d, nlab, i, j, nf[2], nf[3], sens)
            msk = O.schwarz_probe(g, d, nlab, i, j, nf[2], nf[3], sens)
            ref_bits = (imtx * (1 << np.arange(11))[:, None]).sum(axis=0)
            # a dropped kl (ISCR = 0) has every IMTX bit cleared, which is
            # the oracle's msk = 0
            assert np.array_equal(iscr, (msk != 0).astype(np.int32)), (name, sens)
            assert np.array_equal(np.where(iscr == 1, ref_bits, 0), msk), (name, sens)
            assert maxn == int((msk != 0).sum())
            seen_drop += int((msk == 0).sum())
            seen_keep += int((msk == 0x7ff).sum())
            seen_partial += int(((msk != 0) & (msk != 0x7ff)).sum())
    # the comparison reached every outcome of the decision tree
    assert seen_drop > 0 and seen_keep > 0 and seen_partial > 0, (seen_drop, seen_keep,
                                                                   seen_partial)
