"""The matrix-core Breit kernel (real path: atoms and axial molecules)
against the oracle's BRTFAST (pinned to the reference) and against the
vector-ALU Breit kernel, for the Breit and the Gaunt interaction at stages 1
and 3; and its refusal of systems it does not take."""
import numpy as np
import pytest

from conftest import small_systems

pytestmark = pytest.mark.gpu

FOCK_ABS = 1e-10
REAL = ["ne", "diat", "fdiat", "datom", "fatom"]


@pytest.fixture(scope="module")
def systems(bdf):
    return {k: bdf.systems.System(v.name, v.xyz, v.shells, v.shape, hmlt="DHFB")
            for k, v in small_systems(bdf).items()}


@pytest.mark.parametrize("name", REAL)
@pytest.mark.parametrize("gaunt", [False, True])
def test_matrix_core_breit_matches_oracle(bdf, orc, systems, name, gaunt):
    s = systems[name]
    print(f"[breit mfma] {name} gaunt={gaunt}", flush=True)
    d = bdf.systems.random_density(s.ndim, seed=11, real=True)
    O = orc.Oracle(s)
    with bdf.Engine(s, tuned=False, gaunt=gaunt) as eng:
        eng.set_launch(breit_kernel=2)
        b3, x3 = eng.breit(d, 3)
        b1, x1 = eng.breit(d, 1)
        eng.set_launch(breit_kernel=1)
        bv, xv = eng.breit(d, 3)
    bo, xo = O.brtfast(d, 3, gaunt=gaunt)
    assert np.abs(xo).max() > 1e-3
    assert np.abs(x3 - xo).max() < FOCK_ABS, np.abs(x3 - xo).max()
    assert np.abs(b3).max() == 0.0
    assert np.abs(x3 - xv).max() <= 1e-12 * max(np.abs(xv).max(), 1.0)
    _, xo1 = O.brtfast(d, 1, gaunt=gaunt)
    assert np.abs(x1 - xo1).max() < FOCK_ABS


def test_matrix_core_breit_is_the_default_on_the_real_path(bdf, systems):
    s = systems["fdiat"]
    d = bdf.systems.random_density(s.ndim, seed=12, real=True)
    with bdf.Engine(s, tuned=False) as eng:
        assert eng.breit_launch_info(3)[0] == 2
        _, x0 = eng.breit(d, 3)
        eng.set_launch(breit_kernel=2)
        _, x2 = eng.breit(d, 3)
        eng.set_launch(breit_kernel=1)
        assert eng.breit_launch_info(3)[0] == 1
    assert np.abs(x0 - x2).max() <= 1e-12 * np.abs(x2).max()


def test_matrix_core_breit_refuses_the_complex_path(bdf, systems):
    with bdf.Engine(systems["off"], tuned=False) as eng:
        d = bdf.systems.random_density(eng.ndim, seed=13)
        assert eng.breit_launch_info(3)[0] == 1
        eng.breit(d, 3)  # the vector-ALU kernel
        with pytest.raises(bdf.BerthaError):
            eng.set_launch(breit_kernel=2)
            eng.breit(d, 3)
