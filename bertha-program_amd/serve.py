"""Serving loop: Dirac-Hartree-Fock solves of uneven size on one GPU.

Requests (an atom or a diatomic, with its basis and electron count) arrive
over time.  A request is admitted as soon as a slot is free and its device
footprint fits the HBM budget; each admitted solve runs the SCF driver on
its own engine, HIP stream and torch stream, so solves that are in flight
together overlap on the device (continuous batching).  Each solve's
density, Fock matrices and DIIS history stay on the device for its whole
life.  Latency is measured from arrival to the converged energy.
No counterpart in the reference, which runs one SCF per process.
"""
from __future__ import annotations

import collections
import dataclasses
import threading
import time
from typing import List, Optional, Sequence

import numpy as np
import torch

from . import scf as _scf
from .systems import System, even_tempered

# device bytes of the SCF state per basis function pair: S, HNUC, HKIN, F,
# D, G, X, B, temporaries and a 6-deep DIIS history of errors and Fock
# matrices, complex128
_SCF_MATRICES = 40


@dataclasses.dataclass
class Request:
    rid: int
    system: System
    nocc: int
    arrival: float = 0.0  # seconds after the start of the run


@dataclasses.dataclass
class Result:
    rid: int
    name: str
    ndim: int
    energy: float
    iterations: int
    converged: bool
    arrival: float
    start: float
    finish: float
    error: str = ""

    @property
    def latency(self) -> float:
        return self.finish - self.arrival


def _launch_shape(system: System, maxn: int):
    """(workgroups per CU, slab rows) of the engine's Coulomb build, as the
    engine chooses them: R rows N(16) = 969, Boys rows 17 and 4 * cap * N(8)
    G rows; the real path (DIATOM/ATOMIC) runs two waves per SIMD at group
    capacity 2, the complex path three waves at capacity 1 while its
    exchange accumulators 2 (12 n + 16 n) doubles fit 1/12 of the LDS
    (n <= 16 primitives per shell), two waves at capacity 2 up to n = 18 and
    at capacity 1 beyond.  A DHFB engine's Breit scratch holds 12 N(7) G
    rows at eight workgroups per CU."""
    base = 969 + 17
    if system.shape in ("DIATOM", "LINEAR", "ATOMIC"):
        per_cu, rows = 8, base + 4 * 2 * 165
    elif maxn <= 16:
        per_cu, rows = 12, base + 4 * 165
    elif maxn <= 18:
        per_cu, rows = 8, base + 4 * 2 * 165
    else:
        per_cu, rows = 8, base + 4 * 165
    if system.hmlt == "DHFB":
        per_cu, rows = 8, max(rows, base + 12 * 120)
    return per_cu, rows


def estimate_bytes(system: System, workgroups: int = 0, cus: int = 256) -> int:
    """Device footprint of one solve before its engine exists: E tables,
    per-workgroup scratch (for `workgroups` persistent workgroups, or the
    engine's default count on `cus` compute units) and the SCF matrices."""
    words = 0
    maxn = 1
    for a in range(system.ncnt):
        for b in range(system.ncnt):
            for la in range(system.lmax(a) + 1):
                for lb in range(system.lmax(b) + 1):
                    na, nb = system.nfnc_of(a, la), system.nfnc_of(b, lb)
                    # kappa/|m| combinations of two l shells: (2l+1)(2l'+1)
                    nc = (2 * la + 1) * (2 * lb + 1)
                    for extra in ((0, 2, 1, 1, 1) if system.hmlt == "DHFB" else (0, 2)):
                        L = la + lb + extra
                        words += nc * na * nb * (L + 1) * (L + 2) * (L + 3) // 6
        maxn = max([maxn] + [system.nfnc_of(a, l) for l in range(system.lmax(a) + 1)])
    # E, its folded E(CD) copy (double4 each)
    table_bytes = 2 * 32 * words
    ld = ((maxn * maxn + 63) // 64) * 64
    per_cu, rows = _launch_shape(system, maxn)
    wg = workgroups or per_cu * cus
    n = system.ndim
    return table_bytes + 8 * rows * ld * wg + _SCF_MATRICES * 16 * n * n


class Server:
    """Continuous-batching DHF server on one device."""

    def __init__(self, pkg, slots: int = 8, hbm_bytes: Optional[int] = None,
                 workgroups: int = 0, device: str = "cuda:0", max_iter: int = 100,
                 damping: float = 0.5, atomic_guess: bool = True, on_result=None):
        self.pkg = pkg
        # called with each Result as its solve finishes (from a worker thread)
        self.on_result = on_result
        # molecules start from the superposed densities of their atoms in the
        # same basis (closed-shell atoms only), cached across requests
        self.atomic_guess = atomic_guess
        self._atoms = {}
        self.slots = slots
        # persistent workgroups per solve's engine (0: the engine's default,
        # twelve per CU): every solve may use the whole device, and solves in
        # flight together share it through the hardware queues
        self.workgroups = workgroups
        self._cus = torch.cuda.get_device_properties(torch.device(device)).multi_processor_count
        self.device = torch.device(device)
        if hbm_bytes is None:
            free, _ = torch.cuda.mem_get_info(self.device)
            hbm_bytes = int(0.8 * free)
        self.budget = hbm_bytes
        self.max_iter = max_iter
        # density damping before DIIS engages (see scf.run_scf): bare-nucleus
        # starts of Ca- and Kr-containing requests oscillate without it
        self.damping = damping
        self._lock = threading.Condition()
        self._pending: collections.deque = collections.deque()
        self._used = 0
        self._active = 0
        self._closed = False
        self.results: List[Result] = []
        self.peak_active = 0

    def _take(self):
        """Next pending request that fits (arrival order, first fit)."""
        with self._lock:
            while True:
                for k, (req, need) in enumerate(self._pending):
                    if self._used + need <= self.budget or self._active == 0:
                        del self._pending[k]
                        self._used += need
                        self._active += 1
                        self.peak_active = max(self.peak_active, self._active)
                        return req, need
                if self._closed and not self._pending:
                    return None, 0
                self._lock.wait(0.01)

    @staticmethod
    def _atom_key(system: System, c: int):
        return (float(system.charges[c]),
                tuple(tuple(float(x) for x in e) for e in system.shells[c]))

    def _remember(self, system: System, c: int, dens):
        with self._lock:
            self._atoms[self._atom_key(system, c)] = dens.detach().clone()

    def _guess(self, system: System):
        """Superposed atomic densities for a molecule of closed-shell atoms
        (each atom solved once in its basis and cached); None otherwise."""
        if system.ncnt == 1 or any(int(z) not in CLOSED_SHELL_Z for z in system.charges):
            return None
        dens = []
        for c in range(system.ncnt):
            key = self._atom_key(system, c)
            with self._lock:
                d = self._atoms.get(key)
            if d is None:
                z = int(system.charges[c])
                atom = System(f"Z{z}", np.zeros((1, 3)), [system.shells[c]], shape="ATOMIC",
                              charges=[float(z)])
                be = _scf.EngineBackend(self.pkg, atom, device=str(self.device),
                                        workgroups=self.workgroups)
                try:
                    r = _scf.run_scf(atom, z, be, max_iter=self.max_iter, damping=self.damping)
                finally:
                    be.close()
                if not r.converged:
                    return None
                d = r.density
                self._remember(system, c, d)
            dens.append(d)
        return _scf.superposed_density(system, dens)

    def _need(self, system: System) -> int:
        """Admission footprint: the solve's engine plus, for a molecule
        started from atomic densities, the largest atom engine that _guess
        may create while the molecule's engine is alive."""
        need = estimate_bytes(system, self.workgroups, self._cus)
        if self.atomic_guess and system.ncnt > 1:
            atoms = [System("atom", np.zeros((1, 3)), [system.shells[c]], shape="ATOMIC",
                            charges=[float(system.charges[c])]) for c in range(system.ncnt)]
            need += max(estimate_bytes(a, self.workgroups, self._cus) for a in atoms)
        return need

    def _release(self, need: int):
        with self._lock:
            self._used -= need
            self._active -= 1
            self._lock.notify_all()

    def _worker(self, t0: float):
        stream = torch.cuda.Stream(device=self.device)
        while True:
            req, need = self._take()
            if req is None:
                return
            start = time.perf_counter() - t0
            energy, iters, ok, err = float("nan"), 0, False, ""
            try:
                with torch.cuda.stream(stream):
                    be = _scf.EngineBackend(self.pkg, req.system, device=str(self.device),
                                            workgroups=self.workgroups)
                    try:
                        dens0 = self._guess(req.system) if self.atomic_guess else None
                        r = _scf.run_scf(req.system, req.nocc, be, max_iter=self.max_iter,
                                         damping=self.damping, dens0=dens0)
                        energy, iters, ok = r.energy, r.iterations, r.converged
                        if ok and req.system.ncnt == 1 and self.atomic_guess:
                            self._remember(req.system, 0, r.density)
                    finally:
                        be.close()
            except RuntimeError as x:  # a solve that diverges is reported, not fatal
                err = str(x)
            finish = time.perf_counter() - t0
            res = Result(req.rid, req.system.name, req.system.ndim, energy, iters, ok,
                         req.arrival, start, finish, err)
            with self._lock:
                self.results.append(res)
            self._release(need)
            if self.on_result is not None:
                self.on_result(res)

    def run(self, requests: Sequence[Request]) -> List[Result]:
        """Serve requests at their arrival times; returns results by id."""
        self.results = []
        t0 = time.perf_counter()
        workers = [threading.Thread(target=self._worker, args=(t0,), daemon=True)
                   for _ in range(self.slots)]
        for w in workers:
            w.start()
        for req in sorted(requests, key=lambda r: r.arrival):
            delay = req.arrival - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            need = self._need(req.system)
            with self._lock:
                self._pending.append((req, need))
                self._lock.notify_all()
        with self._lock:
            self._closed = True
            self._lock.notify_all()
        for w in workers:
            w.join()
        return sorted(self.results, key=lambda r: r.rid)


# ---- the serving workload -------------------------------------------------
ATOMS = (2, 4, 10, 12, 18, 20, 36, 54)
# neutral atoms whose ground configuration is a closed shell
CLOSED_SHELL_Z = (2, 4, 10, 12, 18, 20, 30, 36, 38, 48, 54, 56, 70, 80, 86)


def atom_basis(z: int):
    """Even-tempered uncontracted basis sized by the nuclear charge (the
    workload leaves the recipe open): more and tighter functions for heavier
    atoms, d functions from Ar on."""
    ns = 10 + z // 4
    np_ = 0 if z <= 2 else 4 + z // 4
    nd = 0 if z < 18 else 2 + z // 12
    sh = [even_tempered(0.05, 2.5, ns)]
    if np_:
        sh.append(even_tempered(0.08, 2.5, np_))
    if nd:
        sh.append(even_tempered(0.15, 2.6, nd))
    return sh


def request_mix(n: int, seed: int = 0, rate: float = 20.0, atoms: Sequence[int] = ATOMS):
    """n requests, Poisson arrivals at `rate` per second; 60 % atoms, 40 %
    diatomics of the same elements with R ~ U(2.5, 8.0) bohr."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for rid in range(n):
        t += rng.exponential(1.0 / rate)
        if rng.uniform() < 0.6:
            z = int(rng.choice(atoms))
            s = System(f"Z{z}", np.zeros((1, 3)), [atom_basis(z)], shape="ATOMIC",
                       charges=[float(z)])
            nocc = z
        else:
            z1, z2 = (int(x) for x in rng.choice(atoms, size=2))
            r = float(rng.uniform(2.5, 8.0))
            s = System(f"Z{z1}-Z{z2}", np.array([[0.0, 0.0, 0.0], [0.0, 0.0, r]]),
                       [atom_basis(z1), atom_basis(z2)], shape="DIATOM",
                       charges=[float(z1), float(z2)])
            nocc = z1 + z2
        out.append(Request(rid, s, nocc, t))
    return out


def percentiles(results: Sequence[Result]):
    lat = np.array([r.latency for r in results])
    return float(np.percentile(lat, 50)), float(np.percentile(lat, 99))
