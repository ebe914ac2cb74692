import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


@pytest.fixture(scope="session")
def bdf():
    from __graft_entry__ import load_package
    return load_package()


@pytest.fixture(scope="session")
def orc():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    return pyoracle


def small_systems(bdf):
    """Systems small enough for the oracle to finish in about a second."""
    S = bdf.systems
    ev = S.even_tempered
    ne = S.ne_atom()
    diat = S.System("diatomic-z", np.array([[0, 0, 0.0], [0, 0, 2.1]]),
                    [[ev(0.1, 2.7, 6), ev(0.15, 2.6, 4), ev(0.3, 2.5, 2)],
                     [ev(0.05, 3.0, 4), ev(0.2, 2.5, 2)]], shape="DIATOM")
    off = S.System("two-centre-off-axis", np.array([[0.1, -0.2, 0.3], [1.2, 0.9, -1.1]]),
                   [[ev(0.1, 2.7, 5), ev(0.15, 2.6, 3), ev(0.3, 2.5, 2)],
                    [ev(0.05, 3.0, 4), ev(0.2, 2.5, 2)]], shape="NOSYMM")
    tri = S.System("three-centre", np.array([[0.0, 0.0, 0.0], [1.4, 0.3, 0.9], [-1.1, 0.8, 0.5]]),
                   [[ev(0.1, 2.8, 4), ev(0.2, 2.6, 2), ev(0.5, 2.5, 1)],
                    [ev(0.06, 3.0, 3), ev(0.3, 2.5, 1)],
                    [ev(0.06, 3.0, 3)]], shape="NOSYMM")
    fsh = S.System("f-shell-pair", np.array([[0.0, 0.0, 0.0], [0.0, 0.3, 1.9]]),
                   [[ev(0.2, 3.0, 2), ev(0.3, 3.0, 2), ev(0.4, 3.0, 1), ev(0.5, 3.0, 1)],
                    [ev(0.1, 3.0, 2), ev(0.3, 3.0, 1)]], shape="NOSYMM")
    # the same shells on the z axis: the real path, whose SS f-f combinations
    # (165 Hermite rows each) are the largest the Coulomb kernel groups
    fdiat = S.System("f-shell-diatomic", np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 1.9]]),
                     [[ev(0.2, 3.0, 2), ev(0.3, 3.0, 2), ev(0.4, 3.0, 1), ev(0.5, 3.0, 1)],
                      [ev(0.1, 3.0, 2), ev(0.3, 3.0, 1)]], shape="DIATOM")
    # four distinct centres: the only systems whose quartets (AB|CD) can put
    # every label on its own centre; "ring4" is planar (z = 0), like the ring
    # workload
    quad = S.System("four-centre", np.array([[0.0, 0.0, 0.0], [1.5, 0.2, 0.4], [-0.4, 1.6, -0.3],
                                             [0.9, -1.2, 1.3]]),
                    [[ev(0.1, 2.8, 3), ev(0.2, 2.6, 2)],
                     [ev(0.08, 3.0, 3), ev(0.3, 2.5, 1)],
                     [ev(0.06, 3.0, 2), ev(0.3, 2.5, 1)],
                     [ev(0.07, 3.0, 2)]], shape="NOSYMM")
    ring4 = S.System("planar-four", np.array([[1.3, 0.0, 0.0], [0.05, 1.35, 0.0], [-1.32, 0.02, 0.0],
                                              [0.0, -1.28, 0.0]]),
                     [[ev(0.1, 2.8, 3), ev(0.2, 2.6, 1), ev(0.5, 2.5, 1)],
                      [ev(0.08, 3.0, 2), ev(0.3, 2.5, 1)],
                      [ev(0.1, 2.8, 3), ev(0.2, 2.6, 1)],
                      [ev(0.08, 3.0, 2)]], shape="NOSYMM")
    # single atoms with d and f shells: the one-centre Racah path's LQN blocks
    # up to (ff|ff), including those that read another block's beta integrals
    datom = S.System("d-atom", np.zeros((1, 3)),
                     [[ev(0.1, 2.7, 6), ev(0.15, 2.6, 4), ev(0.3, 2.5, 3)]], shape="ATOMIC")
    fatom = S.System("f-atom", np.zeros((1, 3)),
                     [[ev(0.2, 3.0, 3), ev(0.3, 3.0, 2), ev(0.4, 3.0, 2), ev(0.5, 3.0, 2)]],
                     shape="ATOMIC")
    return {"ne": ne, "diat": diat, "off": off, "tri": tri, "fsh": fsh, "fdiat": fdiat,
            "quad": quad, "ring4": ring4, "datom": datom, "fatom": fatom}


def sd_axial_density(labels, ndim, seed=1):
    """A real symmetric density coupling only s and d functions of equal m on
    each component (the form an atom polarised along its axis has): its
    one-centre Coulomb contributions reach the (pp|sd) LQN block, whose beta
    integrals COULOMB1 reads from the block before it."""
    icn, kqn, mqn = labels
    ns = ndim // 2
    lq = [k if k > 0 else -k - 1 for k in kqn]
    rng = np.random.default_rng(seed)
    d = np.zeros((ndim, ndim))
    for comp in (0, ns):
        for i in range(ns):
            for j in range(ns):
                if lq[i] == 0 and lq[j] == 2 and mqn[i] == mqn[j] and icn[i] == icn[j]:
                    v = rng.normal()
                    d[comp + i, comp + j] = d[comp + j, comp + i] = v
    return d
