#!/usr/bin/env python3
"""Writes tests/golden/tlf_samples.npz: the oracle's Dirac-Coulomb builds of
the heavy-diatomic workload (Tl 28s24p18d12f + F 14s9p4d on the z axis,
NDIM 1340, Gaussian nuclei) at stages 1 and 3, and its stage-3 Breit build,
at 2000 sampled matrix elements, with E_J and E_K of each Coulomb build and
E_B of the Breit build.

The reference build holds two centres and 26 exponents per shell
(parameters.h), so Tl's 28 s functions are beyond it and nothing of the
reference runs here: the vectors are the oracle's (CLMFAST restated in
oracle/bdf_oracle.c, itself pinned to the compiled reference on smaller
systems).  The GPU test that reads them checks the engine at the workload's
full size: the real (DIATOM) path, f shells, primitive-pair counts up to
784 and the capacity-2 kernel instance.

GXCH is Hermitian only to about 1e-10 relative in the reference's
algorithm on this basis (the oracle shows the same on smaller Tl bases): the
fixture stores the 200 most asymmetric element pairs on both sides.

Density: Tl and F solved alone in their basis with the bare-nucleus
Hamiltonian of their Gaussian nuclei, Z electrons in the lowest
positive-energy Kramers pairs (Tl's odd electron shared by the last pair),
superposed on the molecule.  The occupied spinors and occupations are
stored, not the matrices.  About half an hour of 8 cores (the Breit build
takes two thirds of it).
Run:  python tests/golden/make_tlf_samples.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import torch  # noqa: E402

import pyoracle  # noqa: E402
from __graft_entry__ import load_package  # noqa: E402

NSAMPLE = 2000


def atom_occupied(pkg, scf, system, c):
    """Occupied spinors and weights of centre c's atom, solved alone."""
    z = float(system.charges[c])
    xi_all = pkg.systems.nuclear_exponents(system)
    xi = np.zeros(1) if xi_all is None else np.array([xi_all[c]])
    a = pkg.systems.System("atom", np.zeros((1, 3)), [system.shells[c]], shape="ATOMIC",
                           charges=[z])
    o = pyoracle.Oracle(a)
    ov, hn, hk = (torch.from_numpy(np.asarray(m)) for m in o.oneel(np.array([z]), xi))
    o.close()
    _, coef, _ = scf.solve_fock(hn + hk, scf.orthogonaliser(ov))
    nskp = a.ndim // 2
    nel = int(round(z))
    full, odd = nel // 2, nel % 2
    occ = coef[:, nskp:nskp + 2 * (full + odd)].numpy()
    w = np.ones(occ.shape[1])
    if odd:
        w[-2:] = 0.5
    return occ, w


def density(scf, system, occs):
    blocks = [torch.from_numpy((o.conj() * w) @ o.T) for o, w in occs]
    return scf.superposed_density(system, blocks).numpy()


def main():
    import importlib
    pkg = load_package()
    scf = importlib.import_module(pkg.__name__ + ".scf")
    s = pkg.systems.config_system("tlf", seed=0, hmlt="DHFR")
    occs = [atom_occupied(pkg, scf, s, c) for c in range(s.ncnt)]
    dens = density(scf, s, occs)
    n = s.ndim
    idx = np.random.default_rng(3).integers(0, n, size=(NSAMPLE, 2)).astype(np.int32)
    out = {"idx": idx}
    for c, (o, w) in enumerate(occs):
        out[f"occ{c}"] = o
        out[f"w{c}"] = w
    orc = pyoracle.Oracle(s)
    threads = int(os.environ.get("ORACLE_THREADS", min(16, os.cpu_count() or 1)))
    for ilev in (1, 3):
        t = time.time()
        # the whole build (stride 1), classes as work items
        frac, gd, gx = orc.clmfast_sample(dens, ilev, nthreads=threads, stride=1)
        assert frac == 1.0
        print(f"stage {ilev}: oracle build {time.time() - t:.0f} s", flush=True)
        out[f"s{ilev}_ej"] = 0.5 * np.real(np.sum(dens * gd))
        out[f"s{ilev}_ek"] = 0.5 * np.real(np.sum(dens * gx))
        out[f"s{ilev}_g"] = gd[idx[:, 0], idx[:, 1]]
        out[f"s{ilev}_x"] = gx[idx[:, 0], idx[:, 1]]
        # GXCH is not Hermitian to rounding here: EPSAB/EPSCD cut the tight
        # Tl functions' terms differently on the two sides of a diagonal
        # block.  The 200 most asymmetric pairs (i, j) and both their
        # elements pin that asymmetry.
        asym = np.abs(gx - gx.conj().T)
        flat = np.argsort(asym, axis=None)[::-1][:400]
        pairs = np.array(np.unravel_index(flat, asym.shape)).T
        pairs = pairs[pairs[:, 0] < pairs[:, 1]][:200].astype(np.int32)
        out[f"s{ilev}_asym_idx"] = pairs
        out[f"s{ilev}_asym_x"] = gx[pairs[:, 0], pairs[:, 1]]
        out[f"s{ilev}_asym_xt"] = gx[pairs[:, 1], pairs[:, 0]]
        print(f"stage {ilev}: max|GXCH - GXCH^H| {asym.max():.3e}, max|GXCH| "
              f"{np.abs(gx).max():.3e}, max|GDIR - GDIR^H| "
              f"{np.abs(gd - gd.conj().T).max():.3e}", flush=True)
    # the Breit build (BRTFAST) of the same density on a DHFB system
    sb = pkg.systems.config_system("tlf", seed=0, hmlt="DHFB")
    orb = pyoracle.Oracle(sb)
    t = time.time()
    frac, bd, bx = orb.brtfast_sample(dens, 3, nthreads=threads, stride=1)
    assert frac == 1.0
    print(f"breit: oracle build {time.time() - t:.0f} s", flush=True)
    out["b3_eb"] = 0.5 * np.real(np.sum(dens * (bd - bx)))
    out["b3_x"] = bx[idx[:, 0], idx[:, 1]]
    np.savez_compressed(os.path.join(HERE, "tlf_samples.npz"), **out)


if __name__ == "__main__":
    main()
