#!/usr/bin/env python3
"""Writes tests/golden/tlf_racah.npz: the oracle's one-centre Racah Coulomb
and Breit contributions (COULOMB1 and BREIT1 for both centres,
bertha-scf.f:21540, :26014, each followed by its lower-triangle overwrite)
for the heavy-diatomic workload (Tl
28s24p18d12f + F 14s9p4d on z, NDIM 1340) at the 2000 matrix elements
tests/golden/tlf_samples.npz samples, with E_J, E_K and E_B of the
contributions,
for two densities: the superposed atomic density stored in tlf_samples.npz
(spherical atoms) and a seeded random real density (which also reaches the
LQN blocks that read another block's beta integrals).

COULOMB1 starts from zero matrices here, so the values are the one-centre
part alone.  The reference build holds 26 exponents per shell, fewer than
Tl's 28 s functions, so the vectors are the oracle's (restated in
oracle/bdf_oracle.c, pinned to the compiled reference on smaller atoms and
molecules).  About a minute of one core.
Run:  python tests/golden/make_tlf_racah.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import pyoracle  # noqa: E402
from __graft_entry__ import load_package  # noqa: E402
from make_tlf_samples import density  # noqa: E402

RANDOM_SEED = 41


def main():
    import importlib
    pkg = load_package()
    scf = importlib.import_module(pkg.__name__ + ".scf")
    gold = np.load(os.path.join(HERE, "tlf_samples.npz"))
    s = pkg.systems.config_system("tlf", seed=0, hmlt="DHFR")
    n = s.ndim
    occs = [(gold[f"occ{c}"], gold[f"w{c}"]) for c in range(s.ncnt)]
    idx = gold["idx"]
    out = {"random_seed": np.array(RANDOM_SEED)}
    o = pyoracle.Oracle(s, racah1=True)
    ob = pyoracle.Oracle(pkg.systems.config_system("tlf", seed=0, hmlt="DHFB"), racah1=True)
    zero = np.zeros((n, n), dtype=np.complex128)
    for tag, d in (("atoms", density(scf, s, occs)),
                   ("random", pkg.systems.random_density(n, seed=RANDOM_SEED, real=True))):
        t = time.time()
        g, x = o.coulomb1(d, zero, zero)
        print(f"{tag}: oracle COULOMB1 {time.time() - t:.1f} s, max|GDIR| {np.abs(g).max():.3e}",
              flush=True)
        out[f"{tag}_g"] = g[idx[:, 0], idx[:, 1]]
        out[f"{tag}_x"] = x[idx[:, 0], idx[:, 1]]
        out[f"{tag}_ej"] = 0.5 * np.real(np.sum(d * g))
        out[f"{tag}_ek"] = 0.5 * np.real(np.sum(d * x))
        t = time.time()
        _, bx = ob.breit1(d, zero, zero)
        print(f"{tag}: oracle BREIT1 {time.time() - t:.1f} s, max|BXCH| {np.abs(bx).max():.3e}",
              flush=True)
        out[f"{tag}_bx"] = bx[idx[:, 0], idx[:, 1]]
        out[f"{tag}_eb"] = -0.5 * np.real(np.sum(d * bx))
    np.savez_compressed(os.path.join(HERE, "tlf_racah.npz"), **out)


if __name__ == "__main__":
    main()
