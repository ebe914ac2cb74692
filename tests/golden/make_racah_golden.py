#!/usr/bin/env python3
"""Writes tests/golden/racah_vectors.npz from the reference's own routines.

HFSCF's Coulomb and Breit builds with RACAH1 on (bertha-scf.f:5401-5453):
CLMFAST / BRTFAST without one-centre quartets, then COULOMB1 / BREIT1 for
every centre, run by the reference compiled from the unmodified sources by oracle/Makefile
(oracle/_ref/libbertha_ref.so, ref_clmfast_racah in oracle/ref/ref_driver.f)
on seeded random densities of three small systems of tests/conftest.py: an
atom with s, p and d shells (ATOMIC, CLMMT1Z), the f-shell diatomic on z
(DIATOM, CLMMT1Z) and the off-axis two-centre system (NOSYMM, CLMMT1G); the
Breit build on the same system as DHFB.  Of GDIR, GXCH and BXCH it keeps 600
elements: the 300 largest in magnitude and
300 drawn at random.  The file holds inputs (system name, density seed) and
outputs only.  Run:  python tests/golden/make_racah_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import pyoracle  # noqa: E402
from __graft_entry__ import load_package  # noqa: E402
from conftest import small_systems  # noqa: E402


def main():
    pkg = load_package()
    sy = small_systems(pkg)
    out = {}
    rng = np.random.default_rng(2024)
    for name, seed in (("datom", 31), ("fdiat", 32), ("off", 33)):
        s = sy[name]
        d = pkg.systems.random_density(s.ndim, seed=seed, real=(s.shape != "NOSYMM"))
        g, x = pyoracle.Reference(s, racah1=True).clmfast_racah(d, 3)
        sb = pkg.systems.System(s.name, s.xyz, s.shells, s.shape, hmlt="DHFB")
        _, bx = pyoracle.Reference(sb, racah1=True).brtfast_racah(d, 3)
        n = s.ndim
        mag = np.abs(g) + np.abs(x) + np.abs(bx)
        big = np.argsort(mag.ravel())[-300:]
        rnd = rng.choice(n * n, size=300, replace=False)
        flat = np.unique(np.concatenate([big, rnd]))
        r, c = np.unravel_index(flat, (n, n))
        out[f"{name}_seed"] = np.array(seed)
        out[f"{name}_rows"] = r.astype(np.int32)
        out[f"{name}_cols"] = c.astype(np.int32)
        out[f"{name}_gdir"] = g[r, c]
        out[f"{name}_gxch"] = x[r, c]
        out[f"{name}_bxch"] = bx[r, c]
        print(name, n, len(flat), flush=True)
    np.savez_compressed(os.path.join(HERE, "racah_vectors.npz"), **out)


if __name__ == "__main__":
    main()
