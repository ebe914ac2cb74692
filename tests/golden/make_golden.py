#!/usr/bin/env python3
"""Writes tests/golden/reference_vectors.npz from the reference's own routines.

The reference (bertha-scf.f) ships no tests or vectors for this path, so the
vectors are produced here by running its routines, compiled from the
unmodified sources by oracle/Makefile (oracle/_ref/libbertha_ref.so, driven
through oracle/ref/ref_driver.f), on small inputs:
  - FUNFM for each branch and RMAKE on seeded arguments;
  - E0LL/E0SS blocks, ERIFAST batches and CLMFAST (GDIR, GXCH) at stages 1-3
    for an atom (ATOMIC), a diatomic on z (DIATOM) and an off-axis
    two-centre system (NOSYMM).
The file holds inputs and outputs only.  Run:  python tests/golden/make_golden.py
With --breit it writes tests/golden/breit_vectors.npz instead: EILS blocks,
BIIFAST batches (Breit and Gaunt) and BRTFAST's BXCH (see breit()).
"""
import json
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


def systems(pkg):
    S = pkg.systems
    ev = S.even_tempered
    return {
        "atom": S.System("golden-atom", np.zeros((1, 3)), [[ev(0.1, 3.0, 6), ev(0.12, 3.0, 4)]],
                         shape="ATOMIC"),
        "diat": S.System("golden-diatomic", np.array([[0, 0, 0.0], [0, 0, 2.3]]),
                         [[ev(0.1, 3.0, 4), ev(0.2, 2.8, 3), ev(0.5, 2.5, 1)],
                          [ev(0.07, 3.2, 3), ev(0.25, 2.5, 1)]], shape="DIATOM"),
        "off": S.System("golden-off-axis", np.array([[0.2, -0.1, 0.3], [1.0, 1.1, -1.2]]),
                        [[ev(0.1, 3.0, 4), ev(0.2, 2.8, 2), ev(0.5, 2.5, 1)],
                         [ev(0.07, 3.2, 3), ev(0.25, 2.5, 1)]], shape="NOSYMM"),
    }


def _kq(k):
    return k // 2 if k % 2 == 0 else -(k // 2) - 1


def breit(pkg):
    """tests/golden/breit_vectors.npz: EILS blocks, BIIFAST batches (Breit
    and Gaunt) and BRTFAST's BXCH for DHFB versions of the golden systems
    plus an f-shell pair."""
    S = pkg.systems
    ev = S.even_tempered
    sy = systems(pkg)
    sy["fsh"] = S.System("golden-f-shell", np.array([[0.0, 0.0, 0.0], [0.0, 0.3, 1.9]]),
                         [[ev(0.2, 3.0, 2), ev(0.3, 3.0, 2), ev(0.4, 3.0, 1), ev(0.5, 3.0, 1)],
                          [ev(0.1, 3.0, 2), ev(0.3, 3.0, 1)]], shape="NOSYMM")
    out, meta = {}, {}
    rng = np.random.default_rng(2025)
    for name, s0 in sy.items():
        s = S.System(s0.name, s0.xyz, s0.shells, s0.shape, hmlt="DHFB")
        R = pyoracle.Reference(s)
        meta[name] = {"xyz": s.xyz.tolist(), "shells": [[e.tolist() for e in sh] for sh in s.shells],
                      "shape": s.shape}
        d = pkg.systems.random_density(R.ndim, seed=31, real=(s.shape != "NOSYMM"))
        out[f"{name}_dent"] = d
        R.clmfast(d, 1)  # fresh-process order
        b, x = R.brtfast(d, 3)
        out[f"{name}_bxch"] = x
        blocks, batches = [], []
        for _ in range(8):
            a, bb = [int(v) for v in rng.integers(1, s.ncnt + 1, size=2)]
            ka = int(rng.integers(1, 2 * s.lmax(a - 1) + 2))
            kb = int(rng.integers(1, 2 * s.lmax(bb - 1) + 2))
            ma = int(rng.integers(1, abs(_kq(ka)) + 1))
            mb = int(rng.integers(1, abs(_kq(kb)) + 1))
            key = (2, a, bb, ka, kb, ma, mb)
            out[f"{name}_eils_{'_'.join(map(str, key))}"] = R.etable(*key)
            blocks.append(key)
        for k in range(12):
            icnt = [int(v) for v in rng.integers(1, s.ncnt + 1, size=4)]
            kidx, midx = [], []
            for c in icnt:
                kk = int(rng.integers(1, 2 * s.lmax(c - 1) + 2))
                kidx.append(kk)
                midx.append(int(rng.integers(1, abs(_kq(kk)) + 1)))
            i = int(rng.integers(1, s.nfnc_of(icnt[0] - 1, kidx[0] // 2) + 1))
            j = int(rng.integers(1, s.nfnc_of(icnt[1] - 1, kidx[1] // 2) + 1))
            gaunt = k >= 8
            out[f"{name}_bii{k}"] = R.biifast(icnt, kidx, midx, i, j, gaunt)
            batches.append((icnt, kidx, midx, i, j, gaunt))
        meta[name]["eils"] = blocks
        meta[name]["biifast"] = batches
    out["meta"] = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
    path = os.path.join(HERE, "breit_vectors.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes")


def main():
    if not pyoracle.reference_available():
        sys.exit("oracle/_ref/libbertha_ref.so missing: run make -C oracle ref")
    pkg = load_package()
    if "--breit" in sys.argv:
        breit(pkg)
        return
    out = {}
    meta = {}
    rng = np.random.default_rng(2024)
    R0 = pyoracle.Reference(systems(pkg)["atom"])
    for itype, lo, hi in ((1, 0.0, 1e-11), (2, 1e-6, 17.0), (3, 17.0, 30.0), (4, 30.0, 300.0)):
        t = rng.uniform(lo, hi, size=16)
        out[f"funfm{itype}_t"] = t
        out[f"funfm{itype}_f"] = R0.funfm(t, 12, itype)
    qp = rng.uniform(-2, 2, size=(12, 3))
    aph = rng.uniform(0.05, 20.0, size=12)
    out["rmake_qp"], out["rmake_aph"] = qp, aph
    out["rmake_r"] = R0.rmake(qp, aph, 10)
    for name, s in systems(pkg).items():
        R = pyoracle.Reference(s)
        meta[name] = {"xyz": s.xyz.tolist(), "shells": [[e.tolist() for e in sh] for sh in s.shells],
                      "shape": s.shape}
        d = pkg.systems.random_density(R.ndim, seed=21, real=(s.shape != "NOSYMM"))
        out[f"{name}_dent"] = d
        for ilev in (1, 2, 3):  # stage 1 first: fresh-process sparsity lists
            g, x = R.clmfast(d, ilev)
            out[f"{name}_gdir{ilev}"], out[f"{name}_gxch{ilev}"] = g, x
        blocks = []
        for itt in (1, 4):
            for a in range(1, s.ncnt + 1):
                for b in range(1, s.ncnt + 1):
                    ka = int(rng.integers(1, 2 * s.lmax(a - 1) + 2))
                    kb = int(rng.integers(1, 2 * s.lmax(b - 1) + 2))
                    kqa = ka // 2 if ka % 2 == 0 else -(ka // 2) - 1
                    kqb = kb // 2 if kb % 2 == 0 else -(kb // 2) - 1
                    ma, mb = int(rng.integers(1, abs(kqa) + 1)), int(rng.integers(1, abs(kqb) + 1))
                    key = (itt, a, b, ka, kb, ma, mb)
                    blocks.append(key)
                    out[f"{name}_etab_{'_'.join(map(str, key))}"] = R.etable(*key)
        meta[name]["etables"] = blocks
        batches = []
        for _ in range(12):
            icnt = [int(x) for x in rng.integers(1, s.ncnt + 1, size=4)]
            kidx, midx = [], []
            for c in icnt:
                k = int(rng.integers(1, 2 * s.lmax(c - 1) + 2))
                kq = k // 2 if k % 2 == 0 else -(k // 2) - 1
                kidx.append(k)
                midx.append(int(rng.integers(1, abs(kq) + 1)))
            ibas = int(rng.integers(1, s.nfnc_of(icnt[0] - 1, kidx[0] // 2) + 1))
            jbas = int(rng.integers(1, s.nfnc_of(icnt[1] - 1, kidx[1] // 2) + 1))
            itn = [int(rng.choice([1, 4])), int(rng.choice([1, 4]))]
            key = (icnt, kidx, midx, ibas, jbas, itn)
            out[f"{name}_rr{len(batches)}"] = R.erifast(icnt, kidx, midx, ibas, jbas, itn)
            batches.append(key)
        meta[name]["erifast"] = batches
    out["meta"] = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
    path = os.path.join(HERE, "reference_vectors.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
