"""Closed-shell Dirac-Hartree-Fock SCF driver on the engine.

Restates the iteration of the reference's HFSCF (bertha-scf.f:5097-6680)
for closed shells, with every matrix resident on the device:
  - one-electron matrices once (OVRLP/ONEEL, :5317-5318);
  - per iteration the two-electron matrices of the current stage (CLMFAST
    :5401, BRTFAST :5441 for DHFB), Fock assembly F = HNUC + HKIN + GDIR -
    GXCH + BDIR - BXCH (:5516-5523), energies from the previous density
    (ENRGIES :3698), virtual level shift (LEVSHFT :3641), C-DIIS
    (DIIS :3883, with the workload's 6-deep history instead of MSV = 20),
    the generalised eigenproblem F C = S C e (ZHEGV :5557, here a Cholesky
    reduction and a Hermitian eigensolver on the device) and the density of
    the lowest positive-energy spinors (DENSTY :3580);
  - the stage logic: (LL|LL) -> (LL|SS) -> (SS|SS) at weighted energy
    changes below 1e-8 and 5e-11, convergence at 1e-12 with a density
    change below 1e-10 (:5167-5170, :6037-6108); atoms start at stage 3
    (CARDIN :527-529) and molecules at stage 2 (see run_scf).
One-centre quartets are evaluated by McMurchie-Davidson like all others
(the reference's RACAH1 = .FALSE. setting).

Matrices are torch tensors in natural (row, column) order; the engine's
column-major arrays are exchanged through transposed copies.
"""
from __future__ import annotations

import dataclasses
import re
from typing import List, Optional

import numpy as np
import torch

from .systems import System, nuclear_exponents, nuclear_repulsion

# stage thresholds (HFSCF :5167-5170)
ENRGLV1 = 1.0e-8
ENRGLV2 = 5.0e-11
ENRGTOL = 1.0e-12
DSTYTOL = 1.0e-10
ERRCOM = 1.0e-1  # DIIS extrapolates only below this error (DIIS :3947)


@dataclasses.dataclass
class Energies:
    """ENRGIES (bertha-scf.f:3698-3880), closed shell."""
    nuclear: float
    kinetic: float
    potential: float
    coulomb_direct: float
    coulomb_exchange: float
    breit_direct: float = 0.0
    breit_exchange: float = 0.0

    @property
    def total(self) -> float:
        return (self.nuclear + self.kinetic + self.potential +
                (self.coulomb_direct - self.coulomb_exchange) +
                (self.breit_direct - self.breit_exchange))


@dataclasses.dataclass
class SCFResult:
    energy: float
    energies: Energies
    iterations: int
    converged: bool
    history: List[float]
    stages: List[int]
    eigenvalues: torch.Tensor       # level shift taken off the virtuals, as written to WFNFL
    density: torch.Tensor
    coefficients: Optional[torch.Tensor] = None  # COEF(J, I): column I is spinor I


class EngineBackend:
    """Two-electron and one-electron matrices from the HIP engine; inputs and
    outputs stay on the device (bdf_*_device entry points)."""

    def __init__(self, pkg, system: System, device: str = "cuda:0", **engine_kw):
        # the engine's HIP stream orders its kernels; torch's work runs on
        # the calling thread's current stream and is joined by events
        self.system = system
        self.eng = pkg.Engine(system, **engine_kw)
        self.dev = torch.device(device)
        self.n = self.eng.ndim
        self.nskp = self.eng.nskp
        n = self.n
        self._d = torch.zeros((n, n), dtype=torch.complex128, device=self.dev)
        self._a = torch.zeros_like(self._d)
        self._b = torch.zeros_like(self._d)
        self._c = torch.zeros_like(self._d)
        self.stream = torch.cuda.ExternalStream(self.eng.stream(), device=self.dev)

    def _sync(self):
        # the engine runs on its own stream: order torch's work after it
        torch.cuda.current_stream(self.dev).wait_stream(self.stream)

    def one_electron(self, cv):
        xi = nuclear_exponents(self.system)
        torch.cuda.current_stream(self.dev).synchronize()
        self.eng.one_electron_device(self._a.data_ptr(), self._b.data_ptr(), self._c.data_ptr(),
                                     xi=xi, cv=cv)
        return self._a.T.clone(), self._b.T.clone(), self._c.T.clone()

    def coulomb(self, dens, ilev):
        self._d.copy_(dens.T)
        self.stream.wait_stream(torch.cuda.current_stream(self.dev))
        self.eng.clmfast_device(self._d.data_ptr(), self._a.data_ptr(), self._b.data_ptr(), ilev)
        if self.eng.racah1:
            # RACAH1 (bertha-scf.f:5415-5421): one-centre terms by COULOMB1
            self.eng.coulomb1_device(self._d.data_ptr(), self._a.data_ptr(), self._b.data_ptr())
        self._sync()
        return self._a.T.clone(), self._b.T.clone()

    def breit(self, dens, ilev):
        self._d.copy_(dens.T)
        self.stream.wait_stream(torch.cuda.current_stream(self.dev))
        self.eng.breit_device(self._d.data_ptr(), self._a.data_ptr(), self._b.data_ptr(), ilev)
        if self.eng.racah1:
            # RACAH1 (bertha-scf.f:5447-5453): one-centre terms by BREIT1
            self.eng.breit1_device(self._d.data_ptr(), self._a.data_ptr(), self._b.data_ptr())
        self._sync()
        return self._a.T.clone(), self._b.T.clone()

    def close(self):
        self.eng.close()


class DIIS:
    """C-DIIS of the reference (DIIS, bertha-scf.f:3883-4233): error
    F P S - S P F with P = conj(D); coefficients minimise |B c| subject to
    sum(c) = 1 (ZGGLSE); the Fock matrix is replaced by the combination once
    the error is below ERRCOM and more than two vectors are stored."""

    def __init__(self, depth: int = 6):
        self.msv = depth
        self.err = [None] * depth
        self.fck = [None] * depth
        self.bsav = np.zeros((depth, depth), dtype=np.complex128)
        self.last_errmax = float("inf")

    def reset(self):
        self.bsav[:] = 0.0

    def __call__(self, nmlev: int, fock, dens, ovlp):
        if nmlev == 1:
            self.reset()
        p = dens.conj()
        evec = fock @ p @ ovlp - ovlp @ p @ fock
        klc = nmlev % self.msv
        if klc == 0:
            klc = self.msv
        kss = min(nmlev, self.msv)
        k = klc - 1
        self.err[k] = evec
        self.fck[k] = fock
        # max|error| and the new row of B in one device-to-host transfer
        row = torch.stack([torch.sum(self.err[m].conj() * evec) for m in range(kss)])
        vals = torch.cat([evec.abs().max().reshape(1).to(row.dtype), row]).cpu().numpy()
        errmax = float(vals[0].real)
        self.last_errmax = errmax
        for m in range(kss):
            v = complex(vals[1 + m])
            self.bsav[m, k] = v
            self.bsav[k, m] = np.conj(v)
        if errmax > ERRCOM or kss <= 2:
            return fock
        b = self.bsav[:kss, :kss]
        # ZGGLSE's problem min |B c| subject to sum c = 1, by the null-space
        # method: c = c0 + N y with c0 = 1/kss and N an orthonormal basis of
        # {x : sum x = 0}, y the least-squares solution of B N y = -B c0.
        # Unlike the normal equations it does not square B's condition
        # number, and it is invariant to B's scale (B ~ |error|^2 -> 0 near
        # convergence).
        q, _ = np.linalg.qr(np.ones((kss, 1), dtype=np.complex128), mode="complete")
        nsp = q[:, 1:]
        c0 = np.full(kss, 1.0 / kss, dtype=np.complex128)
        y = np.linalg.lstsq(b @ nsp, -(b @ c0), rcond=None)[0]
        c = c0 + nsp @ y
        out = torch.zeros_like(fock)
        for m in range(kss):
            out = out + complex(c[m]) * self.fck[m]
        return out


def energies(system: System, dens, hkin, hnuc, gdir, gxch, bdir=None, bxch=None) -> Energies:
    # the reference's contraction sum_ij D_ij M_ij (ENRGIES :3745) of every
    # matrix, fetched from the device in one transfer
    mats = [hkin, hnuc, gdir, gxch] + ([bdir, bxch] if bdir is not None else [])
    t = torch.stack([torch.sum(dens * m).real for m in mats]).cpu().tolist()
    e = Energies(nuclear=nuclear_repulsion(system), kinetic=t[0], potential=t[1],
                 coulomb_direct=0.5 * t[2], coulomb_exchange=0.5 * t[3])
    if bdir is not None:
        e.breit_direct = 0.5 * t[4]
        e.breit_exchange = 0.5 * t[5]
    return e


def density(coef, nskp: int, nocc: int):
    """DENSTY (bertha-scf.f:3580), closed shell: D_ij = sum over the nocc
    lowest positive-energy spinors o of conj(C_io) C_jo (columns nskp+o)."""
    occ = coef[:, nskp:nskp + nocc]
    return occ.conj() @ occ.T


def level_shift(fock, ovlp, coef, nskp: int, nocc: int, shlv: float):
    """LEVSHFT (bertha-scf.f:3641): F + shlv sum_v (S C_v)(S C_v)^H over the
    virtual spinors v > nskp + nocc."""
    a = ovlp @ coef[:, nskp + nocc:]
    return fock + shlv * (a @ a.conj().T)


def orthogonaliser(ovlp):
    """L^-1 of ZHEGV's Cholesky factor S = L L^H, from the lower triangle of S
    (once per geometry: S does not change between iterations)."""
    s = torch.tril(ovlp) + torch.tril(ovlp, -1).conj().T
    L = torch.linalg.cholesky(s)
    return torch.linalg.solve_triangular(L, torch.eye(L.shape[0], dtype=L.dtype, device=L.device),
                                         upper=False)


# largest NDIM whose complex transformed Fock matrix is diagonalised on the
# host (NDIM 40-748: 1.59 -> 2.26 solves/s against MAGMA on the device)
HOST_EIGH_MAX = 768


def solve_fock(fock, li, real: bool = False):
    """ZHEGV(1,'V','L'): F C = S C e from the lower triangle of F, with
    li = orthogonaliser(S).  Returns (e, C, finite) where finite is a device
    flag telling whether every eigenvector is finite.

    real: the matrices of an atom or axial molecule (SHAPE ATOMIC, DIATOM,
    LINEAR), whose one- and two-electron matrices the real path builds with
    zero imaginary parts: the transformed F is real symmetric and is
    diagonalised on the device in real arithmetic (rocSOLVER through torch),
    with no copy to the host."""
    f = torch.tril(fock) + torch.tril(fock, -1).conj().T
    fp = li @ f @ li.conj().T
    if real and fp.is_cuda:
        w, v = torch.linalg.eigh(fp.real.contiguous())
        v = v.to(fp.dtype)
    elif fp.is_cuda and fp.shape[0] <= HOST_EIGH_MAX:
        # small matrices: the host's LAPACK, as the reference's ZHEGV; MAGMA's
        # hybrid solver has a fixed cost that dominates at these sizes
        w, v = torch.linalg.eigh(fp.cpu())
        w, v = w.to(fp.device), v.to(fp.device)
    elif fp.is_cuda:
        # the default device eigensolver returns non-finite eigenvectors for
        # these Kramers-degenerate complex matrices; MAGMA's does not
        prev = torch.backends.cuda.preferred_linalg_library()
        torch.backends.cuda.preferred_linalg_library("magma")
        try:
            w, v = torch.linalg.eigh(fp)
        finally:
            torch.backends.cuda.preferred_linalg_library(prev)
    else:
        w, v = torch.linalg.eigh(fp)
    return w, li.conj().T @ v, torch.isfinite(v).all()


def run_scf(system: System, nocc: int, backend, *, shifts=(1.0, 1.0, 1.0), ilev: int = 1,
            diis_depth: int = 6, max_iter: int = 100, cv: float = 137.035999084,
            damping: float = 0.0, dens0=None, verbose: bool = False) -> SCFResult:
    """HFSCF for a closed shell of nocc electrons (NCLS = NOCC, NOPN = 0).

    damping > 0 (not in the reference) mixes that fraction of the previous
    density into the new one until the DIIS error first falls to ERRCOM, and
    never after: it stops the two-state oscillations a bare-nucleus start can
    fall into (3d below 4s in Ca, heavy diatomics at stage 2).

    dens0 (optional) is a starting density, as the reference's READIN start
    (a wavefunction read from file, bertha-scf.f:534-536, 5528-5534): the
    solve then starts at stage 3 and level-shifts from the first iteration
    that has orbitals."""
    n, nskp = backend.n, backend.nskp
    dev = backend.dev
    breit = system.hmlt.upper() == "DHFB"
    if system.ncnt == 1 or dens0 is not None:
        ilev = 3
    elif ilev == 1:
        # With one-centre quartets by McMurchie-Davidson, the reference's
        # stage 1 builds every one-centre G from SS column flags that
        # E0SSCD0 has not yet set (CLMFAST :11742, G reused across IT2 at
        # :12678): all one-centre repulsion vanishes and the SCF collapses.
        # The reference avoids this only through RACAH1 (COULOMB1), so
        # molecules start at stage 2.
        ilev = 2
    shlv = shifts[ilev - 1]
    ovlp, hnuc, hkin = backend.one_electron(cv)
    h1 = hnuc + hkin
    li = orthogonaliser(ovlp)
    if dens0 is None:
        dens = torch.zeros((n, n), dtype=torch.complex128, device=dev)
    else:
        dens = dens0.to(device=dev, dtype=torch.complex128).clone()
    dtmp = dens.clone()
    coef = None
    diis = DIIS(diis_depth)
    nmlev = [0, 0, 0, 0]
    esav = [0.0]
    wedn = [0.0]
    history, stages = [], []
    enrglv2 = ENRGLV2
    e = None
    eig = None
    converged = False
    it = 0
    for it in range(1, max_iter + 1):
        if nocc > 1:
            gdir, gxch = backend.coulomb(dens, ilev)
        else:
            gdir = gxch = torch.zeros_like(dens)
        bdir = bxch = None
        fock = h1 + gdir - gxch
        if breit and nocc > 1:
            bdir, bxch = backend.breit(dens, ilev)
            fock = fock + bdir - bxch
        e = energies(system, dens, hkin, hnuc, gdir, gxch, bdir, bxch)
        etot = e.total
        shifted = 0.0
        if it != 1 and coef is not None:
            fock = level_shift(fock, ovlp, coef, nskp, nocc, shlv)
            shifted = shlv
        nmlev[ilev] += 1
        fock = diis(nmlev[ilev], fock, dens, ovlp)
        eig, coef, finite = solve_fock(fock, li, real=system.shape != "NOSYMM")
        dprev = dens
        dens = density(coef, nskp, nocc)
        if it > 1 and diis.last_errmax <= ERRCOM:
            damping = 0.0  # DIIS has engaged: no more damping in this solve
        if damping > 0.0 and it > 1:
            dens = (1.0 - damping) * dens + damping * dprev
        wedn.append(abs(esav[-1] - etot) / (abs(etot) + 1.0))
        esav.append(etot)
        chk = torch.stack([torch.sqrt(torch.sum((dens - dtmp).abs() ** 2)),
                           finite.to(torch.float64)]).cpu().tolist()
        if chk[1] == 0.0:
            raise RuntimeError("eigensolver returned non-finite eigenvectors")
        dnrm = chk[0] / float(n * n)
        dtmp = dens.clone()
        if dnrm <= 1.0e-9:
            enrglv2 = 1.0e2 * dnrm
        history.append(etot)
        stages.append(ilev)
        if verbose:
            print(f"iter {it:3d} stage {ilev} E = {etot:.12f} dE/E {wedn[-1]:.2e} dD {dnrm:.2e}",
                  flush=True)
        if nmlev[ilev] > 15 and len(wedn) > 6:
            w = wedn[-6:]
            # six rising energy changes; below ENRGLV1 they are the noise of a
            # solve stalled near convergence, not a divergence (the reference
            # tests at any size)
            if all(w[k + 1] > w[k] for k in range(5)) and w[-1] > ENRGLV1:
                raise RuntimeError("SCF diverging: six rising energy changes (HFSCF :6005-6026)")
        if nocc <= 1:
            converged = True
            break
        if ilev == 1:
            if wedn[-1] < ENRGLV1:
                ilev = 2
                shlv = shifts[ilev - 1]
        elif ilev == 2:
            if wedn[-1] < enrglv2:
                ilev = 3
                shlv = shifts[ilev - 1]
        elif wedn[-1] < ENRGTOL and dnrm < DSTYTOL:
            converged = True
            break
    if eig is not None and shifted:
        # the virtuals' level shift comes off before EIGN is written (:5570-5575)
        eig = eig.clone()
        eig[nskp + nocc:] -= shifted
    return SCFResult(energy=e.total, energies=e, iterations=it, converged=converged,
                     history=history, stages=stages, eigenvalues=eig, density=dens,
                     coefficients=coef)


def density_from_coefficients(coef, nskp: int, nocc: int):
    """DENSTY's closed-shell density of the aufbau occupation: the nocc
    spinors above the nskp negative-energy ones, D_ij = sum_o conj(C_io) C_jo."""
    occ = coef[:, nskp:nskp + nocc]
    return occ.conj() @ occ.T


def write_wfn(path: str, eigenvalues, coef) -> None:
    """The reference's wavefunction file (WFNFL, bertha-scf.f:5578-5583): one
    list-directed record per spinor I, EIGN(I) followed by COEF(J,I) for
    J = 1..NDIM as (re,im) pairs, 17 significant digits."""
    w = np.asarray(eigenvalues.detach().cpu() if hasattr(eigenvalues, "detach") else eigenvalues,
                   dtype=np.float64)
    c = np.asarray(coef.detach().cpu() if hasattr(coef, "detach") else coef, dtype=np.complex128)
    n = c.shape[0]
    if c.shape != (n, n) or w.shape != (n,):
        raise ValueError("write_wfn: eigenvalues (ndim,) and coefficients (ndim, ndim) expected")
    with open(path, "w") as f:
        for i in range(n):
            col = c[:, i]
            f.write(f" {w[i]:.17G} " + " ".join(f"({z.real:.17G},{z.imag:.17G})" for z in col)
                    + "\n")


# a list-directed value: a complex constant (re,im) or a real
_WFN_TOKEN = re.compile(r"\(\s*([^,()\s]+)\s*,\s*([^,()\s]+)\s*\)|([^,\s()]+)")


def read_wfn(path: str, ndim: int):
    """Read WFNFL as the reference does (bertha-scf.f:597-604): NDIM
    list-directed records READ(8,*) EIGN(I),(COEF(J,I),J=1,NDIM); a record may
    continue over several lines, values may be separated by blanks or commas
    and use D exponents.  Returns (eigenvalues (ndim,), coefficients
    (ndim, ndim)) as numpy arrays."""
    def num(t):
        try:
            return float(t.replace("D", "E").replace("d", "e"))
        except ValueError:
            raise ValueError(f"read_wfn: not a number: {t!r}") from None

    w = np.zeros(ndim)
    c = np.zeros((ndim, ndim), dtype=np.complex128)
    with open(path) as f:
        lines = iter(f.read().splitlines())
        for i in range(ndim):
            # a record continues over lines until it holds 1 + NDIM values; a
            # complex constant may itself be split over two lines (Fortran
            # runtimes wrap list-directed output)
            buf, vals = "", []
            while len(vals) < ndim + 1:
                try:
                    buf += " " + next(lines)
                except StopIteration:
                    raise ValueError(f"read_wfn: file ends inside record {i + 1}") from None
                if buf.count("(") > buf.count(")"):
                    continue
                vals = []
                for m in _WFN_TOKEN.finditer(buf):
                    if m.group(3) is not None:
                        vals.append(num(m.group(3)))
                    else:
                        vals.append(complex(num(m.group(1)), num(m.group(2))))
            if isinstance(vals[0], complex):
                raise ValueError(f"read_wfn: record {i + 1} starts with a complex value")
            w[i] = vals[0]
            c[:, i] = vals[1:ndim + 1]
    return w, c


def superposed_density(system: System, atom_densities) -> torch.Tensor:
    """Starting density of a molecule from densities of its atoms solved in
    the same basis (the atomic guess the reference's bertha-atomic.f feeds
    through a .wfn file): atom c's large and small blocks are placed at the
    molecule's centre-c blocks, whose FOCKIND order is centre-major
    (bertha-scf.f:2605-2645).  atom_densities[c] is that atom's ndim_c x ndim_c
    density in its own FOCKIND order."""
    sizes = [sum(2 * (2 * l + 1) * len(e) for l, e in enumerate(sh)) for sh in system.shells]
    nskp = sum(sizes)
    d0 = atom_densities[0]
    out = torch.zeros((2 * nskp, 2 * nskp), dtype=torch.complex128, device=d0.device)
    off = 0
    for c, dc in enumerate(atom_densities):
        m = sizes[c]
        if tuple(dc.shape) != (2 * m, 2 * m):
            raise ValueError(f"atom {c}: density {tuple(dc.shape)} does not match its basis")
        for r0, rr in ((0, off), (m, nskp + off)):
            for c0, cc in ((0, off), (m, nskp + off)):
                out[rr:rr + m, cc:cc + m] = dc[r0:r0 + m, c0:c0 + m]
        off += m
    return out
