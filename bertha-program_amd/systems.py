"""Basis sets and geometries of the benchmark workloads.

The reference reads these from an input deck (CARDIN, bertha-scf.f:150-611);
even-tempered exponents follow its EVENTEMPR rule alpha_k = alpha0 * beta**k
(:390-419).  Geometries with a random part are drawn from a seeded generator
so that every run of a workload sees the same molecule.
"""
from __future__ import annotations

import dataclasses
from typing import List, Sequence

import numpy as np


@dataclasses.dataclass
class System:
    name: str
    xyz: np.ndarray                 # (ncnt, 3) bohr
    shells: List[List[np.ndarray]]  # shells[c][l] = exponents
    shape: str = "NOSYMM"           # ATOMIC | LINEAR | DIATOM | NOSYMM
    hmlt: str = "DHFR"              # DHFR (Dirac-Coulomb) | DHFB (+ Breit)
    charges: Sequence[float] = ()
    nucleus: object = "POINT"       # POINT | GAUSS, or one of them per centre (NMDL,
                                    # bertha-scf.f:1182-1204)
    masses: Sequence[float] = ()    # mass numbers A (Gaussian nuclei)

    @property
    def ncnt(self) -> int:
        return len(self.shells)

    def lmax(self, c: int) -> int:
        return len(self.shells[c]) - 1

    def nfnc_of(self, c: int, l: int) -> int:
        return len(self.shells[c][l])

    @property
    def ndim(self) -> int:
        return sum(4 * (2 * l + 1) * len(e) for sh in self.shells for l, e in enumerate(sh))

    def arrays(self, lstride: int = 4, mbs: int = 32):
        n = self.ncnt
        xyz = np.ascontiguousarray(self.xyz, dtype=np.float64).reshape(n, 3)
        lmax = np.array([self.lmax(c) for c in range(n)], dtype=np.int32)
        nfnc = np.zeros((n, lstride), dtype=np.int32)
        exps = np.zeros((n, lstride, mbs), dtype=np.float64)
        for c in range(n):
            for l, e in enumerate(self.shells[c]):
                nfnc[c, l] = len(e)
                exps[c, l, :len(e)] = e
        return xyz, lmax, nfnc, exps, lstride, mbs


# conversions and constants of PHYSPRM (bertha-scf.f:954-993)
BOHR_FM = 5.29177211e4   # CFM: femtometres per bohr
PROTON_RADIUS_FM = 0.842  # PRAD


def nuclear_exponents(system: "System"):
    """Gaussian-nucleus exponents xi = 3/(2 R_rms^2) per centre with the
    reference's RMS-radius rule (bertha-scf.f:1182-1204); None for point
    nuclei."""
    models = ([system.nucleus] * system.ncnt if isinstance(system.nucleus, str)
              else list(system.nucleus))
    if all(m.upper() == "POINT" for m in models):
        return None
    xi = []
    for z, a, m in zip(system.charges, system.masses, models):
        if m.upper() == "POINT":
            xi.append(0.0)  # xi <= 0: point nucleus
            continue
        if int(z) == 1:
            rfm = PROTON_RADIUS_FM
        elif z < 90.0:
            rfm = 0.836 * a ** (1.0 / 3.0) + 0.57
        else:
            rfm = 0.770 * a ** (1.0 / 3.0) + 0.98
        r = rfm / BOHR_FM
        xi.append(1.5 / (r * r))
    return np.array(xi)


def nuclear_repulsion(system: "System") -> float:
    """ENUC of point charges (bertha-scf.f:2549-2574)."""
    e = 0.0
    for i in range(system.ncnt):
        for j in range(i):
            e += system.charges[i] * system.charges[j] / np.linalg.norm(system.xyz[i] - system.xyz[j])
    return e


def even_tempered(alpha0: float, beta: float, n: int) -> np.ndarray:
    return alpha0 * beta ** np.arange(n, dtype=np.float64)


def ne_atom() -> System:
    """Ne, 12s8p, alpha_k = 0.08 * 2.6**k (NDIM 144)."""
    sh = [even_tempered(0.08, 2.6, 12), even_tempered(0.08, 2.6, 8)]
    return System("ne-12s8p", np.zeros((1, 3)), [sh], shape="ATOMIC", charges=[10.0])


def water(seed: int = 0) -> System:
    """H2O, O 14s9p4d (0.05 * 2.4**k), H 8s3p (0.04 * 2.8**k), seeded
    O-H ~ U(1.77, 1.85) bohr and HOH ~ U(102.5, 106.5) degrees (NDIM 380)."""
    rng = np.random.default_rng(seed)
    r1, r2 = rng.uniform(1.77, 1.85, size=2)
    ang = np.deg2rad(rng.uniform(102.5, 106.5))
    xyz = np.array([[0.0, 0.0, 0.0],
                    [r1 * np.sin(ang / 2), 0.0, r1 * np.cos(ang / 2)],
                    [-r2 * np.sin(ang / 2), 0.0, r2 * np.cos(ang / 2)]])
    o = [even_tempered(0.05, 2.4, 14), even_tempered(0.05, 2.4, 9), even_tempered(0.05, 2.4, 4)]
    h = [even_tempered(0.04, 2.8, 8), even_tempered(0.04, 2.8, 3)]
    return System(f"water-s{seed}", xyz, [o, h, [e.copy() for e in h]], shape="NOSYMM",
                  charges=[8.0, 1.0, 1.0])


def tlf(seed: int = 0, hmlt: str = "DHFB") -> System:
    """Tl-F, R ~ U(3.84, 4.04) bohr on z; Tl 28s24p18d12f, F 14s9p4d (NDIM 1340).
    The workload names the shell counts only; the even-tempered parameters
    here span Tl's 1s region (alpha up to ~1e7) and F's valence."""
    rng = np.random.default_rng(seed)
    r = rng.uniform(3.84, 4.04)
    tl = [even_tempered(0.02, 2.1, 28), even_tempered(0.02, 2.1, 24),
          even_tempered(0.04, 2.2, 18), even_tempered(0.08, 2.3, 12)]
    f = [even_tempered(0.06, 2.5, 14), even_tempered(0.06, 2.5, 9), even_tempered(0.2, 2.5, 4)]
    # Gaussian nuclei with A = 1.5 Z, as the workload states
    return System(f"tlf-s{seed}", np.array([[0.0, 0.0, 0.0], [0.0, 0.0, r]]), [tl, f],
                  shape="DIATOM", hmlt=hmlt, charges=[81.0, 9.0], nucleus="GAUSS",
                  masses=[1.5 * 81.0, 1.5 * 9.0])


def ring(seed: int = 0) -> System:
    """Planar C6H6-like ring: 6 C at r=2.64, 6 H at r=4.69 bohr, 60 degrees
    apart, seeded +-0.02 bohr jitter; C 12s6p3d (0.06 * 2.5**k), H 6s2p
    (0.05 * 2.8**k) (NDIM 1368)."""
    rng = np.random.default_rng(seed)
    xyz = []
    shells = []
    c = [even_tempered(0.06, 2.5, 12), even_tempered(0.06, 2.5, 6), even_tempered(0.06, 2.5, 3)]
    h = [even_tempered(0.05, 2.8, 6), even_tempered(0.05, 2.8, 2)]
    for k in range(6):
        t = np.pi / 3 * k
        xyz.append([2.64 * np.cos(t), 2.64 * np.sin(t), 0.0])
        shells.append([e.copy() for e in c])
    for k in range(6):
        t = np.pi / 3 * k
        xyz.append([4.69 * np.cos(t), 4.69 * np.sin(t), 0.0])
        shells.append([e.copy() for e in h])
    xyz = np.array(xyz) + rng.uniform(-0.02, 0.02, size=(12, 3)) * np.array([1.0, 1.0, 0.0])
    return System(f"ring-s{seed}", xyz, shells, shape="NOSYMM", charges=[6.0] * 6 + [1.0] * 6)


CONFIGS = {"ne": ne_atom, "water": water, "tlf": tlf, "ring": ring}


def config_system(name: str, **kw) -> System:
    return CONFIGS[name](**kw)


def random_density(ndim: int, seed: int = 0, real: bool = False, scale: float = 0.05) -> np.ndarray:
    """Hermitian test density (column-major complex)."""
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((ndim, ndim))
    if not real:
        a = a + 1j * rng.standard_normal((ndim, ndim))
    d = (a + a.conj().T) * (0.5 * scale)
    return np.asfortranarray(d.astype(np.complex128))
