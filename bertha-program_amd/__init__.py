"""MI355X two-electron engine for four-component Dirac-Hartree-Fock.

Host-side mirror of the reference's two-electron interface (bertha-scf.f):
``Engine.__init__`` stands for EQSAVE (:32984), ``Engine.clmfast`` for
CLMFAST (:11489) and ``Engine.erifast`` for ERIFAST (:12269).  All arithmetic
runs in the HIP library ``libbertha_amd.so`` built from ``csrc/``; this module
only marshals arrays through the C ABI declared in ``include/bertha_amd.h``.
There is no CPU fallback: if the library is missing or no GPU is present the
calls raise.
"""
from __future__ import annotations

import ctypes as _C
import os as _os

import numpy as _np

from .systems import System, config_system  # noqa: F401

_HERE = _os.path.dirname(_os.path.abspath(__file__))
LIBPATH = _os.path.join(_HERE, "libbertha_amd.so")

BDF_OK = 0
BDF_EINVAL = -1
BDF_ENOMEM = -2
BDF_EDEVICE = -3
BDF_ELIMIT = -4
BDF_EUNSUPP = -5

SHAPES = {"NOSYMM": 0, "LINEAR": 1, "DIATOM": 1, "ATOMIC": 2}
CV_REFERENCE = 137.035999084  # speed of light of the reference (bertha-scf.f:974)
HMLTS = {"DHFR": 1, "DHFB": 2}


class BerthaError(RuntimeError):
    """A C-ABI call returned a negative status."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"bdf error {code}: {msg}")
        self.code = code


class _System(_C.Structure):
    _fields_ = [
        ("ncnt", _C.c_int),
        ("xyz", _C.POINTER(_C.c_double)),
        ("lmax", _C.POINTER(_C.c_int)),
        ("nfnc", _C.POINTER(_C.c_int)),
        ("exps", _C.POINTER(_C.c_double)),
        ("lstride", _C.c_int),
        ("mbs", _C.c_int),
        ("shape", _C.c_int),
        ("hmlt", _C.c_int),
    ]


class _Launch(_C.Structure):
    _fields_ = [
        ("graph", _C.c_int),
        ("waves", _C.c_int),
        ("cap", _C.c_int),
        ("per_cu", _C.c_int),
        ("breit_per_cu", _C.c_int),
        ("breit_group", _C.c_int),
        ("breit_kernel", _C.c_int),
    ]


class _Options(_C.Structure):
    _fields_ = [
        ("intsym", _C.c_int),
        ("racah1", _C.c_int),
        ("precision", _C.c_int),
        ("split_threshold", _C.c_double),
        ("device", _C.c_int),
        ("schwarz", _C.c_int),
        ("sens", _C.c_double),
        ("gaunt", _C.c_int),
        ("workgroups", _C.c_int),
    ]


_lib = None


def library() -> _C.CDLL:
    """Load libbertha_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not _os.path.exists(LIBPATH):
            raise RuntimeError(
                f"{LIBPATH} is missing: build it with `make -C bertha-program_amd` "
                "(or __graft_entry__.build())")
        # torch ships its own HIP runtime under the same soname as the
        # system's: whichever loads first serves the whole process, and torch
        # finds no GPU through the system's.  Load torch's first where torch
        # is installed, so that the engine and torch share it.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        lib = _C.CDLL(LIBPATH)
        P = _C.c_void_p
        lib.bdf_last_error.restype = _C.c_char_p
        lib.bdf_create.argtypes = [_C.POINTER(_System), _C.POINTER(_Options),
                                   _C.POINTER(P), _C.POINTER(_C.c_int), _C.POINTER(_C.c_int)]
        lib.bdf_coulomb.argtypes = [P, _C.c_int, P, P, _C.c_int, P, P, P, P]
        lib.bdf_coulomb_device.argtypes = [P, _C.c_int, P, P, P]
        lib.bdf_coulomb_device_static.argtypes = [P, _C.c_int, P, P, P]
        lib.bdf_coulomb1.argtypes = [P, _C.c_int, P, P, _C.c_int, P, P, P, P]
        lib.bdf_coulomb1_device.argtypes = [P, _C.c_int, P, P, P]
        lib.bdf_breit1.argtypes = [P, _C.c_int, P, P, _C.c_int, P, P, P, P]
        lib.bdf_breit1_device.argtypes = [P, _C.c_int, P, P, P]
        lib.bdf_racah_info.argtypes = [P, _C.c_int, _C.POINTER(_C.c_double),
                                       _C.POINTER(_C.c_long), _C.POINTER(_C.c_long)]
        lib.bdf_eri_batch.argtypes = [P, P, P, P, _C.c_int, _C.c_int, P, _C.c_int, P, _C.c_int]
        lib.bdf_etable.argtypes = [P, _C.c_int, _C.c_int, _C.c_int, _C.c_int, _C.c_int,
                                   _C.c_int, _C.c_int, P, _C.c_long, _C.POINTER(_C.c_long)]
        lib.bdf_last_stats.argtypes = [P, _C.POINTER(_C.c_double), _C.POINTER(_C.c_double)]
        lib.bdf_plan_info.argtypes = [P, _C.c_int, _C.POINTER(_C.c_double),
                                      _C.POINTER(_C.c_double), _C.POINTER(_C.c_long)]
        lib.bdf_schwarz_tables.argtypes = [P, P, P, _C.c_int]
        lib.bdf_breit.argtypes = [P, _C.c_int, P, P, _C.c_int, P, P, P, P]
        lib.bdf_one_electron.argtypes = [P, P, P, _C.c_double, P, P, P, _C.c_int]
        lib.bdf_one_electron_device.argtypes = [P, P, P, _C.c_double, P, P, P]
        lib.bdf_breit_device.argtypes = [P, _C.c_int, P, P, P]
        lib.bdf_bii_batch.argtypes = [P, P, P, P, _C.c_int, _C.c_int, P, _C.c_int]
        lib.bdf_breit_plan_info.argtypes = [P, _C.c_int, _C.POINTER(_C.c_double),
                                            _C.POINTER(_C.c_double), _C.POINTER(_C.c_long)]
        lib.bdf_memory_info.argtypes = [P, _C.POINTER(_C.c_size_t)]
        lib.bdf_set_launch.argtypes = [P, _C.POINTER(_Launch)]
        lib.bdf_launch_info.argtypes = [P, _C.c_int, _C.POINTER(_C.c_int), _C.POINTER(_C.c_int),
                                        _C.POINTER(_C.c_int), _C.POINTER(_C.c_size_t)]
        lib.bdf_breit_launch_info.argtypes = [P, _C.c_int, _C.POINTER(_C.c_int),
                                              _C.POINTER(_C.c_int)]
        lib.bdf_stream.argtypes = [P]
        lib.bdf_stream.restype = P
        lib.bdf_destroy.argtypes = [P]
        lib.bdf_destroy.restype = None
        _lib = lib
    return _lib


def _schwarz_mode(schwarz) -> int:
    """False/0/None: off; True/1/"reference": the reference's SCHWARZ masks
    (GDSC); "rigorous"/2: the masks with every group bounded by Q(a,b) Q(c,d)."""
    modes = {"reference": 1, "rigorous": 2}
    if isinstance(schwarz, str):
        if schwarz not in modes:
            raise ValueError(f"schwarz must be one of {sorted(modes)}, False or True, not {schwarz!r}")
        return modes[schwarz]
    if schwarz is None or schwarz in (0, 1, 2):  # bools are 0 and 1
        return int(schwarz or 0)
    raise ValueError(f"schwarz must be 0, 1 or 2, not {schwarz!r}")


def _check(rc: int) -> None:
    if rc != BDF_OK:
        raise BerthaError(rc, library().bdf_last_error().decode())


class Engine:
    """One system's device tables and Fock-build plans.

    Construction builds the E0LL/E0SS tables on the GPU (the reference's
    EQSAVE).  ``clmfast`` returns (GDIR, GXCH) for a density, matching the
    reference's CLMFAST outputs element by element.
    """

    def __init__(self, system: System, intsym: bool = True, racah1: bool = False,
                 device: int = 0, schwarz=False, sens: float = 0.0,
                 precision: str = "fp64", split_threshold: float = 0.0, gaunt: bool = False,
                 workgroups: int = 0, tuned: bool = True, graph: bool = False):
        lib = library()
        self.system = system
        self.racah1 = bool(racah1)
        xyz, lmax, nfnc, exps, lstride, mbs = system.arrays()
        self._keep = (xyz, lmax, nfnc, exps)
        s = _System(system.ncnt, xyz.ctypes.data_as(_C.POINTER(_C.c_double)),
                    lmax.ctypes.data_as(_C.POINTER(_C.c_int)),
                    nfnc.ctypes.data_as(_C.POINTER(_C.c_int)),
                    exps.ctypes.data_as(_C.POINTER(_C.c_double)), lstride, mbs,
                    SHAPES[system.shape], HMLTS[system.hmlt])
        prec = {"fp64": 0, "mixed": 1}[precision]
        o = _Options(1 if intsym else 0, 1 if racah1 else 0, prec, split_threshold, device,
                     _schwarz_mode(schwarz), sens, 1 if gaunt else 0, workgroups)
        h = _C.c_void_p()
        nd = _C.c_int()
        ns = _C.c_int()
        _check(lib.bdf_create(_C.byref(s), _C.byref(o), _C.byref(h), _C.byref(nd), _C.byref(ns)))
        self._h = h
        self.ndim = nd.value
        self.nskp = ns.value
        # launch settings (bdf_set_launch), with the autotuner's shapes for
        # this kind of system (tune.py) unless tuned is False
        self._launch = {"graph": False, "waves": 0, "cap": 0, "per_cu": 0, "breit_per_cu": 0,
                        "breit_group": 0, "breit_kernel": 0}
        self.tuning = None
        upd = {"graph": bool(graph)}
        if tuned:
            from . import tune as _tune
            entry = _tune.lookup(system)
            if entry:
                upd.update(waves=entry["waves"], cap=entry["cap"], per_cu=entry["per_cu"],
                           breit_per_cu=entry.get("breit_per_cu", 0),
                           breit_group=entry.get("breit_group", 0),
                           breit_kernel=entry.get("breit_kernel", 0))
                self.tuning = entry
        if any(upd.values()):
            self.set_launch(**upd)

    def close(self) -> None:
        if getattr(self, "_h", None):
            library().bdf_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def clmfast(self, dent: _np.ndarray, ilev: int = 3):
        """GDIR, GXCH for total density ``dent`` (ndim x ndim complex)."""
        n = self.ndim
        d = _np.asfortranarray(dent, dtype=_np.complex128)
        if d.shape != (n, n):
            raise ValueError(f"density must be {n}x{n}")
        g = _np.zeros((n, n), dtype=_np.complex128, order="F")
        x = _np.zeros((n, n), dtype=_np.complex128, order="F")
        _check(library().bdf_coulomb(self._h, ilev, d.ctypes.data, None, n,
                                     g.ctypes.data, x.ctypes.data, None, None))
        return g, x

    def clmfast_device(self, d_dent: int, d_gdir: int, d_gxch: int, ilev: int = 3) -> None:
        """Enqueue a build on device pointers (e.g. torch tensors' data_ptr())."""
        _check(library().bdf_coulomb_device(self._h, ilev, d_dent, d_gdir, d_gxch))

    def clmfast_device_static(self, d_dent: int, d_gdir: int, d_gxch: int, ilev: int = 3) -> None:
        """bdf_coulomb_device_static: the same build without allocating (the
        stage's plan must exist: plan_info(ilev) builds it)."""
        _check(library().bdf_coulomb_device_static(self._h, ilev, d_dent, d_gdir, d_gxch))

    def coulomb1(self, dent: _np.ndarray, gdir: _np.ndarray, gxch: _np.ndarray,
                 exact_beta: bool = False):
        """bdf_coulomb1: every centre's COULOMB1 on top of a CLMFAST result
        (gdir, gxch of an engine with racah1=True); returns the new matrices."""
        n = self.ndim
        d = _np.asfortranarray(dent, dtype=_np.complex128)
        g = _np.array(gdir, dtype=_np.complex128, order="F")
        x = _np.array(gxch, dtype=_np.complex128, order="F")
        if d.shape != (n, n) or g.shape != (n, n) or x.shape != (n, n):
            raise ValueError(f"matrices must be {n}x{n}")
        _check(library().bdf_coulomb1(self._h, 1 if exact_beta else 0, d.ctypes.data, None, n,
                                      g.ctypes.data, x.ctypes.data, None, None))
        return g, x

    def coulomb1_device(self, d_dent: int, d_gdir: int, d_gxch: int,
                        exact_beta: bool = False) -> None:
        _check(library().bdf_coulomb1_device(self._h, 1 if exact_beta else 0, d_dent, d_gdir,
                                             d_gxch))

    def clmfast_racah(self, dent: _np.ndarray, ilev: int = 3, exact_beta: bool = False):
        """HFSCF's Coulomb build with RACAH1 on: CLMFAST without one-centre
        quartets, then COULOMB1 for every centre (engine built with
        racah1=True)."""
        g, x = self.clmfast(dent, ilev)
        return self.coulomb1(dent, g, x, exact_beta)

    def breit1(self, dent: _np.ndarray, bdir: _np.ndarray, bxch: _np.ndarray,
               exact_beta: bool = False):
        """bdf_breit1: every centre's BREIT1 on top of a BRTFAST result
        (bdir, bxch of an engine with racah1=True); returns the matrices."""
        n = self.ndim
        d = _np.asfortranarray(dent, dtype=_np.complex128)
        b = _np.array(bdir, dtype=_np.complex128, order="F")
        x = _np.array(bxch, dtype=_np.complex128, order="F")
        if d.shape != (n, n) or b.shape != (n, n) or x.shape != (n, n):
            raise ValueError(f"matrices must be {n}x{n}")
        _check(library().bdf_breit1(self._h, 1 if exact_beta else 0, d.ctypes.data, None, n,
                                    b.ctypes.data, x.ctypes.data, None, None))
        return b, x

    def breit1_device(self, d_dent: int, d_bdir: int, d_bxch: int,
                      exact_beta: bool = False) -> None:
        _check(library().bdf_breit1_device(self._h, 1 if exact_beta else 0, d_dent, d_bdir,
                                           d_bxch))

    def breit_racah(self, dent: _np.ndarray, ilev: int = 3, exact_beta: bool = False):
        """HFSCF's Breit build with RACAH1 on: BRTFAST without one-centre
        quartets, then BREIT1 for every centre (engine built with
        racah1=True)."""
        b, x = self.breit(dent, ilev)
        return self.breit1(dent, b, x, exact_beta)

    def racah_info(self, breit: bool = False):
        """(digestion terms per build, work items, uncovered beta reads) of
        the one-centre Racah build of the Coulomb or Breit interaction."""
        t, i, u = _C.c_double(), _C.c_long(), _C.c_long()
        _check(library().bdf_racah_info(self._h, 1 if breit else 0, _C.byref(t), _C.byref(i),
                                        _C.byref(u)))
        return t.value, i.value, u.value

    def erifast(self, icnt, kidx, midx, ibas: int, jbas: int, itn, ilev: int = 3):
        """RR(N,1..16) of one ERIFAST batch, shape (16, maxcd) complex."""
        sysm = self.system
        lq = []
        for c, k in zip(icnt[2:], kidx[2:]):
            lq.append(k // 2)
        maxcd = sysm.nfnc_of(icnt[2] - 1, lq[0]) * sysm.nfnc_of(icnt[3] - 1, lq[1])
        rr = _np.zeros((16, maxcd), dtype=_np.complex128)
        a = [_np.asarray(v, dtype=_np.int32) for v in (icnt, kidx, midx, itn)]
        _check(library().bdf_eri_batch(self._h, a[0].ctypes.data, a[1].ctypes.data,
                                       a[2].ctypes.data, ibas, jbas, a[3].ctypes.data, ilev,
                                       rr.ctypes.data, maxcd))
        return rr

    def etable(self, itt: int, ica: int, icb: int, ka: int, kb: int, ma: int, mb: int):
        """One E0LL (itt=1) / E0SS (itt=4) block as [4][ntuv*maxab]; EILS
        (itt=2, DHFB engines) as [12][ntuv*maxab]."""
        nc = 12 if itt == 2 else 4
        cap = nc * 400000
        out = _np.zeros(cap)
        n = _C.c_long()
        _check(library().bdf_etable(self._h, itt, ica, icb, ka, kb, ma, mb, out.ctypes.data,
                                    cap, _C.byref(n)))
        return out[:nc * n.value].reshape(nc, n.value)

    def one_electron(self, znuc=None, xi=None, cv: float = CV_REFERENCE):
        """(OVLP, HNUC, HKIN) of OVRLP/ONEEL.  znuc defaults to the system's
        charges; xi None: point nuclei, else per-centre Gaussian exponents."""
        n = self.ndim
        z = _np.ascontiguousarray(self.system.charges if znuc is None else znuc, dtype=_np.float64)
        if z.size != self.system.ncnt:
            raise ValueError("one nuclear charge per centre is needed")
        x = None if xi is None else _np.ascontiguousarray(xi, dtype=_np.float64)
        out = [_np.zeros((n, n), dtype=_np.complex128, order="F") for _ in range(3)]
        _check(library().bdf_one_electron(self._h, z.ctypes.data,
                                          None if x is None else x.ctypes.data, cv,
                                          *[o.ctypes.data for o in out], n))
        return tuple(out)

    def one_electron_device(self, d_ovlp: int, d_hnuc: int, d_hkin: int, znuc=None, xi=None,
                            cv: float = CV_REFERENCE) -> None:
        z = _np.ascontiguousarray(self.system.charges if znuc is None else znuc, dtype=_np.float64)
        x = None if xi is None else _np.ascontiguousarray(xi, dtype=_np.float64)
        _check(library().bdf_one_electron_device(self._h, z.ctypes.data,
                                                 None if x is None else x.ctypes.data, cv,
                                                 d_ovlp, d_hnuc, d_hkin))

    def breit(self, dent: _np.ndarray, ilev: int = 3):
        """BDIR, BXCH (BRTFAST) for total density ``dent``."""
        n = self.ndim
        d = _np.asfortranarray(dent, dtype=_np.complex128)
        if d.shape != (n, n):
            raise ValueError(f"density must be {n}x{n}")
        b = _np.zeros((n, n), dtype=_np.complex128, order="F")
        x = _np.zeros((n, n), dtype=_np.complex128, order="F")
        _check(library().bdf_breit(self._h, ilev, d.ctypes.data, None, n, b.ctypes.data,
                                   x.ctypes.data, None, None))
        return b, x

    def breit_device(self, d_dent: int, d_bdir: int, d_bxch: int, ilev: int = 3) -> None:
        _check(library().bdf_breit_device(self._h, ilev, d_dent, d_bdir, d_bxch))

    def biifast(self, icnt, kidx, midx, ibas: int, jbas: int):
        """RR(N,1..16) of one BIIFAST batch, shape (16, maxcd) complex."""
        sysm = self.system
        maxcd = sysm.nfnc_of(icnt[2] - 1, kidx[2] // 2) * sysm.nfnc_of(icnt[3] - 1, kidx[3] // 2)
        rr = _np.zeros((16, maxcd), dtype=_np.complex128)
        a = [_np.asarray(v, dtype=_np.int32) for v in (icnt, kidx, midx)]
        _check(library().bdf_bii_batch(self._h, a[0].ctypes.data, a[1].ctypes.data,
                                       a[2].ctypes.data, ibas, jbas, rr.ctypes.data, maxcd))
        return rr

    def breit_plan_info(self, ilev: int = 3):
        f = _C.c_double()
        q = _C.c_double()
        n = _C.c_long()
        _check(library().bdf_breit_plan_info(self._h, ilev, _C.byref(f), _C.byref(q),
                                             _C.byref(n)))
        return f.value, q.value, n.value

    def schwarz_tables(self):
        """(GDSC of CLMDIAG, sqrt|(ab|ba)| bound table), ndim x ndim real."""
        n = self.ndim
        g = _np.zeros((n, n), order="F")
        q = _np.zeros((n, n), order="F")
        _check(library().bdf_schwarz_tables(self._h, g.ctypes.data, q.ctypes.data, n))
        return g, q

    def plan_info(self, ilev: int = 3):
        """(algorithmic flops, primitive quartets, work items) of one build."""
        f = _C.c_double()
        q = _C.c_double()
        n = _C.c_long()
        _check(library().bdf_plan_info(self._h, ilev, _C.byref(f), _C.byref(q), _C.byref(n)))
        return f.value, q.value, n.value

    def set_launch(self, **kw) -> None:
        """bdf_set_launch: graph (HIP-graph replay of builds), waves, cap,
        per_cu, breit_per_cu, breit_group (the quartet kernels' launch
        shapes, 0: the engine's choice), breit_kernel (1 vector-ALU, 2
        matrix-core, 0 the engine's choice); settings not named keep their
        current values."""
        bad = set(kw) - set(self._launch)
        if bad:
            raise TypeError(f"unknown launch settings {sorted(bad)}")
        new = dict(self._launch, **kw)
        lo = _Launch(1 if new["graph"] else 0, new["waves"], new["cap"], new["per_cu"],
                     new["breit_per_cu"], new["breit_group"], new["breit_kernel"])
        _check(library().bdf_set_launch(self._h, _C.byref(lo)))
        self._launch = new

    @property
    def launch(self) -> dict:
        return dict(self._launch)

    def launch_info(self, ilev: int = 3):
        """(waves per SIMD, group capacity, workgroups, scratch bytes) of the
        stage's Coulomb build."""
        w, c, g = _C.c_int(), _C.c_int(), _C.c_int()
        b = _C.c_size_t()
        _check(library().bdf_launch_info(self._h, ilev, _C.byref(w), _C.byref(c), _C.byref(g),
                                         _C.byref(b)))
        return w.value, c.value, g.value, b.value

    def breit_launch_info(self, ilev: int = 3):
        """(kernel, workgroups) of the stage's Breit build: kernel 1 is the
        vector-ALU breit_kernel, 2 the matrix-core breit_tc_kernel."""
        k, g = _C.c_int(), _C.c_int()
        _check(library().bdf_breit_launch_info(self._h, ilev, _C.byref(k), _C.byref(g)))
        return k.value, g.value

    def device_bytes(self) -> int:
        """Device memory held by the engine."""
        b = _C.c_size_t()
        _check(library().bdf_memory_info(self._h, _C.byref(b)))
        return b.value

    def stream(self) -> int:
        return int(library().bdf_stream(self._h) or 0)

    def stats(self):
        q = _C.c_double()
        ms = _C.c_double()
        _check(library().bdf_last_stats(self._h, _C.byref(q), _C.byref(ms)))
        return q.value, ms.value
