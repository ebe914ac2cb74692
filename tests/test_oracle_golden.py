"""The oracle against vectors written by the reference's own routines
(tests/golden/reference_vectors.npz, from tests/golden/make_golden.py).
Runs everywhere: it needs neither the reference sources nor a GPU."""
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "reference_vectors.npz")


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD, allow_pickle=False)


@pytest.fixture(scope="module")
def meta(gold):
    return json.loads(bytes(gold["meta"]).decode())


def _system(bdf, m):
    sh = [[np.array(e) for e in s] for s in m["shells"]]
    return bdf.systems.System("golden", np.array(m["xyz"]), sh, shape=m["shape"])


def test_funfm_matches_reference_vectors(orc, gold):
    for itype in (1, 2, 3, 4):
        t = gold[f"funfm{itype}_t"]
        assert np.array_equal(orc.funfm(t, 12, itype), gold[f"funfm{itype}_f"]), itype


def test_rmake_matches_reference_vectors(orc, gold):
    r = orc.rmake(gold["rmake_qp"], gold["rmake_aph"], 10)
    ref = gold["rmake_r"]
    assert np.abs(r - ref).max() <= 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("name", ["atom", "diat", "off"])
def test_etables_match_reference_vectors(bdf, orc, gold, meta, name):
    O = orc.Oracle(_system(bdf, meta[name]))
    for key in meta[name]["etables"]:
        ref = gold[f"{name}_etab_{'_'.join(map(str, key))}"]
        e = O.etable(*key)
        assert np.abs(e - ref).max() <= 1e-13 * max(np.abs(ref).max(), 1e-300), key


@pytest.mark.parametrize("name", ["atom", "diat", "off"])
def test_erifast_matches_reference_vectors(bdf, orc, gold, meta, name):
    s = _system(bdf, meta[name])
    O = orc.Oracle(s)
    use = [0, 3, 5, 6, 9, 10, 12, 15] if s.shape != "NOSYMM" else list(range(16))
    for i, (icnt, kidx, midx, ibas, jbas, itn) in enumerate(meta[name]["erifast"]):
        ref = gold[f"{name}_rr{i}"]
        rr = O.erifast(icnt, kidx, midx, ibas, jbas, itn)
        assert np.abs(rr[use] - ref[use]).max() <= 1e-12 + 1e-13 * np.abs(ref[use]).max()


@pytest.mark.parametrize("name", ["atom", "diat", "off"])
def test_clmfast_matches_reference_vectors(bdf, orc, gold, meta, name):
    O = orc.Oracle(_system(bdf, meta[name]))
    d = gold[f"{name}_dent"]
    for ilev in (1, 2, 3):
        g, x = O.clmfast(d, ilev)
        assert np.abs(g - gold[f"{name}_gdir{ilev}"]).max() < 1e-12, ilev
        assert np.abs(x - gold[f"{name}_gxch{ilev}"]).max() < 1e-12, ilev


BREIT = os.path.join(HERE, "golden", "breit_vectors.npz")


@pytest.fixture(scope="module")
def bgold():
    return np.load(BREIT, allow_pickle=False)


@pytest.fixture(scope="module")
def bmeta(bgold):
    return json.loads(bytes(bgold["meta"]).decode())


@pytest.mark.parametrize("name", ["atom", "diat", "off", "fsh"])
def test_breit_matches_reference_vectors(bdf, orc, bgold, bmeta, name):
    m = bmeta[name]
    sh = [[np.array(e) for e in s] for s in m["shells"]]
    s = bdf.systems.System("golden", np.array(m["xyz"]), sh, shape=m["shape"], hmlt="DHFB")
    O = orc.Oracle(s)
    for key in m["eils"]:
        ref = bgold[f"{name}_eils_{'_'.join(map(str, key))}"]
        assert np.abs(O.etable(*key) - ref).max() <= 1e-14 * max(np.abs(ref).max(), 1e-300)
    for k, (icnt, kidx, midx, i, j, gaunt) in enumerate(m["biifast"]):
        ref = bgold[f"{name}_bii{k}"]
        rr = O.biifast(icnt, kidx, midx, i, j, gaunt)
        assert np.abs(rr - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max()), k
    b, x = O.brtfast(bgold[f"{name}_dent"], 3)
    assert np.abs(x - bgold[f"{name}_bxch"]).max() < 1e-13
