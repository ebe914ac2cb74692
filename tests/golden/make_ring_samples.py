#!/usr/bin/env python3
"""Writes tests/golden/ring_samples.npz: the oracle's stage-3 Dirac-Coulomb
build of the 12-centre ring workload (6 C, 6 H, NDIM 1368) at 2000 sampled
matrix elements, unscreened and with the SCHWARZ masks at SENS 1e-12.

The reference build holds two centres at most, so nothing of the reference
runs here: the vectors are the oracle's (bertha-scf.f's CLMFAST restated in
oracle/bdf_oracle.c), and the GPU test that reads them checks the engine at
the full size of the workload (3.0 M work items, 19 M ERIFAST calls), where
32-bit table offsets and plan sizes are exercised as nowhere else.

Density: atoms solved alone in their basis with the bare-nucleus
Hamiltonian (C: six lowest positive-energy spinors; H: one electron shared
by the lowest Kramers pair), superposed on the ring (a physical density:
Hermitian, positive, 42 electrons, elements below 1).  The occupied
spinors are stored, not the matrices.  About 45 minutes of 8 cores.
Run:  python tests/golden/make_ring_samples.py
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


def atom_spinors(pkg, scf, shells, z, nel):
    a = pkg.systems.System("atom", np.zeros((1, 3)), [shells], shape="ATOMIC", charges=[z])
    o = pyoracle.Oracle(a)
    ov, hn, hk = (torch.from_numpy(np.asarray(m)) for m in o.oneel(np.array([z]), np.zeros(1)))
    o.close()
    w, c, _ = scf.solve_fock(hn + hk, scf.orthogonaliser(ov))
    nskp = a.ndim // 2
    return c[:, nskp:nskp + max(nel, 2)].numpy()


def ring_density(pkg, scf, system, occ_c, occ_h):
    dc = torch.from_numpy(occ_c.conj() @ occ_c.T)
    dh = torch.from_numpy(0.5 * (occ_h.conj() @ occ_h.T))
    return scf.superposed_density(system, [dc if z == 6.0 else dh for z in system.charges]).numpy()


def main():
    import importlib
    pkg = load_package()
    scf = importlib.import_module(pkg.__name__ + ".scf")
    ring = pkg.systems.config_system("ring", seed=0)
    occ_c = atom_spinors(pkg, scf, ring.shells[0], 6.0, 6)
    occ_h = atom_spinors(pkg, scf, ring.shells[6], 1.0, 1)
    dens = ring_density(pkg, scf, ring, occ_c, occ_h)
    n = ring.ndim
    idx = np.random.default_rng(1).integers(0, n, size=(NSAMPLE, 2)).astype(np.int32)
    out = {"occ_c": occ_c, "occ_h": occ_h, "idx": idx}
    o = pyoracle.Oracle(ring)
    threads = min(16, os.cpu_count() or 1)
    for label, kw in (("fp64", {}), ("schwarz", {"schwarz": True})):
        t = time.time()
        gd, gx = o.clmfast(dens, 3, nthreads=threads, **kw)
        print(f"{label}: oracle build {time.time() - t:.0f} s", flush=True)
        out[label + "_ej"] = 0.5 * np.real(np.sum(dens * gd))
        out[label + "_ek"] = 0.5 * np.real(np.sum(dens * gx))
        out[label + "_g"] = gd[idx[:, 0], idx[:, 1]]
        out[label + "_x"] = gx[idx[:, 0], idx[:, 1]]
    np.savez_compressed(os.path.join(HERE, "ring_samples.npz"), **out)


if __name__ == "__main__":
    main()
