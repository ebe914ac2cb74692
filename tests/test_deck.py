"""Input decks in the reference's CARDIN format (bertha-scf.f:150-611) read by
bertha-program_amd/deck.py.  The decks here are written for the tests in the
reference's record layout; the CPU tests check the parsed system against the
benchmark builders and the reference's rules, the GPU test solves one."""
import importlib

import numpy as np
import pytest

NE_DECK = """'OUTPUT STRING'
Ne
'CALCULATION TREE: HFSCF, MBPTN, MCSCF, EXPVL, PLOTS'
HFSCF
'HAMILTONIAN: NORL, BARE, DHFR, DHFP, DHFB OR DHFQ'
DHFR
'CONTINUING CALCULATION: RESUME (TRUE), NEW START (FALSE)'
.FALSE.
'NUMBER OF ATOMIC CENTRES'
1
'CENTER 1'
'CARTESIAN COORDINATES'
0.0000,0.0000,0.0000
'NUCLEUS: MODEL, CHARGE, MASS'
POINT,10,19.992
'ATOMIC BASIS: TYPE AND LMAX'
EVENTEMPR,1
'ELECTRON CONFIGURATION AND # ELECTRONS'
AUFBAU,10
'BASIS SET EXPONENTS'
0.08D0,2.6D0,12,1
0.08D0,2.6D0,8,1
'NUMBER OF CLOSED- AND OPEN-SHELL ORBITALS AND OPEN SHELL ELECTRONS'
10,0,0
'LEVEL SHIFT PARAMETER FOR EACH INTEGRAL INCLUSION LEVEL'
1.0,1.0,1.0
'STARTING STAGE OF INTEGRAL INCLUSION LEVEL (1-3)'
1
"""

HF_DECK = """'OUTPUT STRING'
HF
'CALCULATION TREE: HFSCF, MBPTN, MCSCF, EXPVL, PLOTS'
HFSCF
'HAMILTONIAN: NORL, BARE, DHFR, DHFP, DHFB OR DHFQ'
DHFB
'CONTINUING CALCULATION: RESUME (TRUE), NEW START (FALSE)'
.FALSE.
'NUMBER OF ATOMIC CENTRES'
2
'CENTER 1'
'CARTESIAN COORDINATES'
0.3,-0.2,0.1
'NUCLEUS: MODEL, CHARGE, MASS'
FERMI,9,18.998
'ATOMIC BASIS: TYPE AND LMAX'
GEOMETRIC,1
'ELECTRON CONFIGURATION AND # ELECTRONS'
MANUAL,9
'MANUAL CONFIGURATION'
2,2,2
1,5
'BASIS SET EXPONENTS'
0.1,3.0,6
0.2,2.5,3
'CENTER 2'
'CARTESIAN COORDINATES'
1.3,0.4,1.5
'NUCLEUS: MODEL, CHARGE, MASS'
POINT,1,1.008
'ATOMIC BASIS: TYPE AND LMAX'
OPTIMISED,0
'ELECTRON CONFIGURATION AND # ELECTRONS'
AUFBAU,1
'BASIS SET EXPONENTS'
3
1.2D+01
2.0D+00
3.0D-01
'NUMBER OF CLOSED- AND OPEN-SHELL ORBITALS AND OPEN SHELL ELECTRONS'
10,0,0
'LEVEL SHIFT PARAMETER FOR EACH INTEGRAL INCLUSION LEVEL'
1.0,1.0,0.5
'STARTING STAGE OF INTEGRAL INCLUSION LEVEL (1-3)'
1
"""


@pytest.fixture(scope="module")
def deck(bdf):
    return importlib.import_module(bdf.__name__ + ".deck")


def test_atom_deck_is_the_ne_workload(bdf, deck):
    d = deck.parse_deck(NE_DECK)
    ref = bdf.systems.ne_atom()
    assert d.hmlt == "DHFR" and d.tree == "HFSCF" and not d.readin
    assert d.nocc == 10 and d.shifts == (1.0, 1.0, 1.0)
    assert d.ilev == 3  # one centre: CARDIN starts at stage 3 (:534-536)
    assert d.system.shape == "ATOMIC" and d.system.ndim == ref.ndim == 144
    for a, b in zip(d.system.shells[0], ref.shells[0]):
        assert np.allclose(a, b, rtol=1e-15, atol=0)
    assert list(d.system.charges) == [10.0] and d.system.nucleus == "POINT"


def test_diatomic_deck_geometry_basis_and_nuclei(bdf, deck):
    d = deck.parse_deck(HF_DECK)
    s = d.system
    assert d.hmlt == "DHFB" and s.hmlt == "DHFB"
    assert d.ilev == 1 and d.shifts == (1.0, 1.0, 0.5) and d.nocc == 10
    assert d.configurations == ["MANUAL", "AUFBAU"]
    # collinear: laid on z with the bond length kept (NUCGEOM's DIATOM)
    assert s.shape == "DIATOM"
    r = np.linalg.norm(np.array([1.3, 0.4, 1.5]) - np.array([0.3, -0.2, 0.1]))
    assert np.allclose(s.xyz, [[0, 0, 0], [0, 0, r]])
    # GEOMETRIC: alpha * beta^k; OPTIMISED: the listed exponents
    assert np.allclose(s.shells[0][0], 0.1 * 3.0 ** np.arange(6))
    assert np.allclose(s.shells[0][1], 0.2 * 2.5 ** np.arange(3))
    assert np.allclose(s.shells[1][0], [12.0, 2.0, 0.3])
    # FERMI runs as a Gaussian nucleus without a nuclear data file
    assert list(s.nucleus) == ["GAUSS", "POINT"]
    xi = bdf.systems.nuclear_exponents(s)
    assert xi[0] > 0 and xi[1] == 0.0
    r_fm = 0.836 * 18.998 ** (1.0 / 3.0) + 0.57
    assert abs(xi[0] - 1.5 / (r_fm / bdf.systems.BOHR_FM) ** 2) < 1e-12 * xi[0]


def test_eventempr_start_index(bdf, deck):
    text = NE_DECK.replace("0.08D0,2.6D0,8,1", "0.08D0,2.6D0,8,3")
    d = deck.parse_deck(text)
    assert np.allclose(d.system.shells[0][1], 0.08 * 2.6 ** np.arange(2, 10))


@pytest.mark.parametrize("change,message", [
    (("DHFR\n", "NORL\n"), "outside this engine"),
    (("10,0,0", "8,2,2"), "open shells"),
    (("HFSCF\n", "OTHER\n"), "calculation tree"),
    (("0.08D0,2.6D0,8,1\n", ""), "expected a number"),
])
def test_rejected_decks(deck, change, message):
    with pytest.raises(deck.DeckError, match=message):
        deck.parse_deck(NE_DECK.replace(*change, 1))


def test_truncated_deck(deck):
    with pytest.raises(deck.DeckError, match="deck ends early"):
        deck.parse_deck(NE_DECK[:NE_DECK.index("'LEVEL SHIFT")])


@pytest.mark.gpu
def test_deck_solve_matches_the_driver(bdf, deck):
    """The Ne deck through deck.solve gives the SCF energy of the same
    workload run through the driver directly."""
    scf = importlib.import_module(bdf.__name__ + ".scf")
    print("[deck] Ne", flush=True)
    d = deck.parse_deck(NE_DECK)
    res, breit = deck.solve(bdf, d)
    assert res.converged and breit is None
    s = bdf.systems.ne_atom()
    be = scf.EngineBackend(bdf, s)
    try:
        r0 = scf.run_scf(s, 10, be)
    finally:
        be.close()
    assert abs(res.energy - r0.energy) < 1e-8
    assert abs(res.energy - (-128.6586819)) < 1e-6
    # DHFP: the same SCF plus one Breit build of the converged density
    res_p, breit_p = deck.solve(bdf, deck.parse_deck(NE_DECK.replace("DHFR\n", "DHFP\n", 1)))
    assert abs(res_p.energy - res.energy) < 1e-8
    # the oracle's Breit build of its converged Ne density gives 0.0166147 Eh
    assert breit_p is not None and abs(breit_p - 0.0166147) < 1e-6


@pytest.mark.gpu
def test_deck_resumes_from_its_wavefunction_file(bdf, deck, tmp_path):
    """The solve writes <name>_<HMLT>.wfn; the same deck with RESUME (READIN)
    starts from it and converges at once to the same energy."""
    import os
    scf = importlib.import_module(bdf.__name__ + ".scf")
    print("[deck] Ne write and resume", flush=True)
    d = deck.parse_deck(NE_DECK)
    res, _ = deck.solve(bdf, d, wfn_dir=str(tmp_path))
    path = deck.wfn_path(d, str(tmp_path))
    assert res.converged and os.path.basename(path) == "Ne_DHFR.wfn" and os.path.exists(path)
    w, c = scf.read_wfn(path, d.system.ndim)
    assert np.array_equal(w, res.eigenvalues.cpu().numpy())
    assert np.array_equal(c, res.coefficients.cpu().numpy())
    resume = deck.parse_deck(NE_DECK.replace(".FALSE.", ".TRUE.", 1))
    assert resume.readin
    r2, _ = deck.solve(bdf, resume, wfn_dir=str(tmp_path))
    assert r2.converged and abs(r2.energy - res.energy) < 1e-8
    assert r2.iterations <= 4 < res.iterations
