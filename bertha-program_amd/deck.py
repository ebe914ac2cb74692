"""Reader of the reference's input decks (CARDIN, bertha-scf.f:150-611).

A deck is a sequence of Fortran list-directed records: a quoted title line
before every value line, numbers with D exponents, logicals as .TRUE./.FALSE.
parse_deck() turns one into a System (geometry, basis, nuclear model) and the
SCF settings the reference reads with it (closed-shell count NCLS, level
shifts SHLEV, starting stage ILEV, READIN), so that the engine's SCF driver
and serving loop can be fed the reference's own decks.

solve() runs a deck on the device and, given a directory, reads and writes
the deck's wavefunction file (<name>_<HMLT>.wfn, the reference's WFNFL) in
the reference's record layout (scf.read_wfn / scf.write_wfn).

Scope: the HFSCF tree's Dirac Hamiltonians DHFR (Coulomb), DHFB (Coulomb and
Breit) and DHFP (Coulomb SCF, Breit added once at the end); POINT nuclei and
Gaussian ones (GAUSS, and FERMI/UNIFM as the reference runs them without a
nuclear data file); OPTIMISED, GEOMETRIC and EVENTEMPR basis lists; closed
shells.  The rest of the reference's options (NORL/BARE/DHFQ, open shells)
raise DeckError.

Geometry: one centre is placed at the origin (ATOMIC); collinear centres are
laid on the z axis, the first at the origin (DIATOM or LINEAR); anything else
keeps its coordinates (NOSYMM, as NUCGEOM ends for every non-linear molecule,
bertha-scf.f:2440-2444).  NUCGEOM's own translation and rotation differ, which
leaves energies unchanged.
"""
from __future__ import annotations

import dataclasses
import os
import re
from typing import List

import numpy as np

from .systems import System


class DeckError(ValueError):
    pass


@dataclasses.dataclass
class Deck:
    name: str
    tree: str
    hmlt: str                 # as written: DHFR, DHFB or DHFP
    readin: bool
    system: System            # hmlt DHFR for DHFP decks
    nocc: int                 # NCLS
    shifts: tuple             # SHLEV(1..3)
    ilev: int                 # starting stage after CARDIN's rules (:527-541)
    configurations: List[str] = dataclasses.field(default_factory=list)  # CNFG per centre


_TOKEN = re.compile(r"'([^']*)'|\"([^\"]*)\"|([^,\s]+)")


def _tokens(line: str) -> List[str]:
    out = []
    for m in _TOKEN.finditer(line):
        out.append(m.group(1) if m.group(1) is not None else
                   m.group(2) if m.group(2) is not None else m.group(3))
    return out


class _Reader:
    """READ(5, *) semantics: each read starts a new record and takes as many
    values as it asks for, continuing onto following records if needed."""

    def __init__(self, text: str):
        self.lines = [l for l in text.splitlines() if l.strip()]
        self.pos = 0

    def read(self, n: int = 1) -> List[str]:
        vals: List[str] = []
        while len(vals) < n:
            if self.pos >= len(self.lines):
                raise DeckError("deck ends early")
            vals += _tokens(self.lines[self.pos])
            self.pos += 1
        return vals[:n]

    def title(self) -> None:
        self.read(1)

    def read_counted(self) -> List[str]:
        """A record whose first value n is followed by n more values."""
        if self.pos >= len(self.lines):
            raise DeckError("deck ends early")
        vals = _tokens(self.lines[self.pos])
        self.pos += 1
        n = _int(vals[0])
        while len(vals) < n + 1:
            if self.pos >= len(self.lines):
                raise DeckError("deck ends early")
            vals += _tokens(self.lines[self.pos])
            self.pos += 1
        return vals[1:n + 1]


def _num(s: str) -> float:
    try:
        return float(s.replace("D", "E").replace("d", "e"))
    except ValueError:
        raise DeckError(f"expected a number, read {s!r}") from None


def _int(s: str) -> int:
    return int(_num(s))


def _logical(s: str) -> bool:
    t = s.strip().upper().strip(".")
    if t in ("TRUE", "T"):
        return True
    if t in ("FALSE", "F"):
        return False
    raise DeckError(f"not a logical: {s!r}")


def _shape_and_coordinates(xyz: np.ndarray):
    n = len(xyz)
    if n == 1:
        return "ATOMIC", np.zeros((1, 3))
    d = xyz - xyz[0]
    axis = d[np.argmax(np.linalg.norm(d, axis=1))]
    length = np.linalg.norm(axis)
    if length < 1e-12:
        raise DeckError("coincident centres")
    u = axis / length
    off = np.linalg.norm(d - np.outer(d @ u, u), axis=1)
    if np.all(off < 1e-6):
        z = d @ u
        return ("DIATOM" if n == 2 else "LINEAR"), np.column_stack(
            [np.zeros(n), np.zeros(n), z])
    return "NOSYMM", np.array(xyz, dtype=float)


def parse_deck(text: str) -> Deck:
    r = _Reader(text)
    r.title()
    name = r.read(1)[0]
    r.title()
    tree = r.read(1)[0].upper()
    if tree not in ("HFSCF", "MBPTN", "EXPVL", "PLOTS"):
        raise DeckError(f"invalid calculation tree {tree!r}")
    r.title()
    hmlt = r.read(1)[0].upper()[:4]
    if hmlt not in ("NORL", "BARE", "DHFR", "DHFP", "DHFB", "DHFQ"):
        raise DeckError(f"unknown Hamiltonian {hmlt!r}")
    if hmlt not in ("DHFR", "DHFP", "DHFB"):
        raise DeckError(f"Hamiltonian {hmlt} is outside this engine (DHFR, DHFP, DHFB)")
    r.title()
    readin = _logical(r.read(1)[0])
    r.title()
    ncnt = _int(r.read(1)[0])
    xyz, shells, charges, masses, models, cnfg = [], [], [], [], [], []
    for _ in range(ncnt):
        r.title()  # 'CENTER n'
        r.title()  # 'CARTESIAN COORDINATES'
        xyz.append([_num(v) for v in r.read(3)])
        r.title()
        model, z, a = r.read(3)
        model = model.upper()
        if model not in ("POINT", "GAUSS", "FERMI", "UNIFM"):
            raise DeckError(f"illegal nuclear model {model!r}")
        # FERMI and UNIFM nuclei are Gaussian fits read from
        # output/<name>_nuclear.dat; without that file NUCCOUL falls back to
        # one Gaussian of the same RMS radius (bertha-scf.f:1196-1262), which
        # is what a deck alone defines
        models.append("POINT" if model == "POINT" else "GAUSS")
        charges.append(_num(z))
        masses.append(_num(a))
        r.title()
        btyp, lmax = r.read(2)
        btyp, lmax = btyp.upper(), _int(lmax)
        r.title()
        conf, _iq = r.read(2)
        conf = conf.upper()
        if conf not in ("AUFBAU", "MANUAL"):
            raise DeckError(f"illegal configuration option {conf!r}")
        cnfg.append(conf)
        if conf == "MANUAL":
            r.title()
            for _l in range(lmax + 1):
                r.read_counted()  # NLVL and the occupations of each level (:330-333)
        r.title()
        sh = []
        for _l in range(lmax + 1):
            if btyp == "OPTIMISED":
                nf = _int(r.read(1)[0])
                sh.append(np.array([_num(r.read(1)[0]) for _ in range(nf)]))
            elif btyp == "GEOMETRIC":
                al, be, nf = r.read(3)
                al, be, nf = _num(al), _num(be), _int(nf)
                sh.append(al * be ** np.arange(nf))
            elif btyp == "EVENTEMPR":
                al, be, nf, n1 = r.read(4)
                al, be, nf, n1 = _num(al), _num(be), _int(nf), _int(n1)
                sh.append(al * be ** np.arange(n1 - 1, n1 - 1 + nf))
            else:
                raise DeckError(f"illegal basis type {btyp!r}")
        shells.append(sh)
    r.title()
    ncls, nopn, _noelec = (_int(v) for v in r.read(3))
    if nopn != 0:
        raise DeckError("open shells are outside this engine (closed shells only)")
    r.title()
    shifts = tuple(_num(v) for v in r.read(3))
    r.title()
    ilev = _int(r.read(1)[0])
    if ncnt == 1 or readin:
        ilev = 3  # CARDIN :534-536
    shape, coords = _shape_and_coordinates(np.array(xyz, dtype=float))
    nucleus = "GAUSS" if any(m == "GAUSS" for m in models) else "POINT"
    if nucleus == "GAUSS" and any(m == "POINT" for m in models):
        nucleus = list(models)
    system = System(str(name), coords, shells, shape=shape,
                    hmlt="DHFB" if hmlt == "DHFB" else "DHFR", charges=charges,
                    nucleus=nucleus, masses=masses)
    return Deck(str(name), tree, hmlt, readin, system, ncls, shifts, ilev, cnfg)


def read_deck(path: str) -> Deck:
    with open(path) as f:
        return parse_deck(f.read())


def wfn_path(deck: Deck, directory: str) -> str:
    """WFNFL: <directory>/<name>_<HMLT>.wfn (the reference's
    'output/'//MOLCL//'_'//HMLT//'.wfn', bertha-scf.f:241)."""
    return os.path.join(directory, f"{deck.name}_{deck.hmlt}.wfn")


def solve(pkg, deck: Deck, max_iter: int = 100, verbose: bool = False, wfn_dir=None,
          **engine_kw):
    """HFSCF of a deck on the device (DHFR/DHFB), or a DHFR solve followed by
    one Breit build with the converged density (DHFP).  Returns
    (SCFResult, Breit energy added after the SCF or None).

    With wfn_dir, the wavefunction file of the deck is used as the reference
    uses it: a RESUME deck (READIN) whose file exists starts from the density
    of its spinors (stage 3 from the first iteration, bertha-scf.f:534-536,
    597-604), and the converged eigenvalues and spinors are written back.

    racah1=True (engine_kw) runs the reference's default one-centre path:
    CLMFAST and BRTFAST without one-centre quartets plus COULOMB1 and BREIT1
    (bdf_coulomb1, bdf_breit1) in every build."""
    import torch

    from . import scf as _scf
    dens0 = None
    path = wfn_path(deck, wfn_dir) if wfn_dir is not None else None
    if path is not None and deck.readin and os.path.exists(path):
        _, coef = _scf.read_wfn(path, deck.system.ndim)
        nskp = deck.system.ndim // 2
        dens0 = _scf.density_from_coefficients(torch.from_numpy(coef), nskp, deck.nocc)
    be = _scf.EngineBackend(pkg, deck.system, **engine_kw)
    try:
        res = _scf.run_scf(deck.system, deck.nocc, be, shifts=deck.shifts, ilev=deck.ilev,
                           max_iter=max_iter, verbose=verbose, dens0=dens0)
    finally:
        be.close()
    if path is not None:
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        _scf.write_wfn(path, res.eigenvalues, res.coefficients)
    if deck.hmlt != "DHFP":
        return res, None
    sb = dataclasses.replace(deck.system, hmlt="DHFB")
    bb = _scf.EngineBackend(pkg, sb, **engine_kw)
    try:
        bdir, bxch = bb.breit(res.density, 3)
        e = 0.5 * float(torch.sum(res.density * (bdir - bxch)).real)
    finally:
        bb.close()
    return res, e
