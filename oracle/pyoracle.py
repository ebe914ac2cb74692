"""ctypes front ends of the checker libraries (test infrastructure only).

``Oracle``    -- the C restatement in oracle/liboracle.so (bdf_oracle.c).
``Reference`` -- the reference's own routines, compiled from bertha-scf.f into
                 oracle/_ref/libbertha_ref.so by oracle/Makefile (only where
                 the reference sources exist; its compiled limits apply:
                 at most 2 centres, 26 exponents per shell).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use
this module; the engine never does.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libbertha_ref.so")

SHAPE = {"NOSYMM": 0, "LINEAR": 1, "DIATOM": 1, "ATOMIC": 2}

_olib = None


def _oracle_lib():
    global _olib
    if _olib is None:
        if not os.path.exists(ORACLE_LIB):
            raise RuntimeError(f"{ORACLE_LIB} missing: run make -C oracle liboracle.so")
        lib = C.CDLL(ORACLE_LIB)
        P = C.c_void_p
        lib.orc_create.restype = P
        lib.orc_create.argtypes = [C.c_int, P, P, P, C.c_int, P, C.c_int, C.c_int, C.c_int, C.c_int]
        lib.orc_destroy.argtypes = [P]
        lib.orc_ndim.argtypes = [P]
        lib.orc_nskp.argtypes = [P]
        lib.orc_labels.argtypes = [P, P, P, P]
        lib.orc_funfm.argtypes = [P, C.c_int, C.c_int, C.c_int, P]
        lib.orc_rmake.argtypes = [P, P, C.c_int, C.c_int, P]
        lib.orc_etable.restype = C.c_long
        lib.orc_etable.argtypes = [P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                   C.c_int, P, C.c_long]
        lib.orc_erifast.argtypes = [P, P, P, P, C.c_int, C.c_int, P, C.c_int, P, P]
        lib.orc_clmfast.argtypes = [P, C.c_int, P, P, P]
        lib.orc_clmfast_mt.argtypes = [P, C.c_int, P, P, P, C.c_int]
        lib.orc_coulomb1.argtypes = [P, P, P, P]
        lib.orc_breit1.argtypes = [P, P, P, P]
        lib.orc_brtfast_sample.restype = C.c_double
        lib.orc_brtfast_sample.argtypes = [P, C.c_int, P, P, P, C.c_int, C.c_int]
        lib.orc_clmfast_sample.restype = C.c_double
        lib.orc_clmfast_sample.argtypes = [P, C.c_int, P, P, P, C.c_int, C.c_int]
        lib.orc_clmdiag.argtypes = [P, P]
        lib.orc_pairbound.argtypes = [P, P]
        lib.orc_set_schwarz.argtypes = [P, C.c_int, C.c_double]
        lib.orc_schwarz_stats.argtypes = [P, P, P]
        lib.orc_schwarz_probe.argtypes = [P, P, P, P, C.c_int, C.c_int, C.c_int, C.c_int,
                                          C.c_double, P]
        lib.orc_brtfast.argtypes = [P, C.c_int, P, P, P]
        lib.orc_brtfast_mt.argtypes = [P, C.c_int, P, P, P, C.c_int]
        lib.orc_set_gaunt.argtypes = [P, C.c_int]
        lib.orc_biifast.argtypes = [P, P, P, P, C.c_int, C.c_int, P, P]
        lib.orc_oneel.argtypes = [P, P, P, C.c_double, P, P, P]
        _olib = lib
    return _olib


def funfm(t, lam: int, itype: int) -> np.ndarray:
    """FUNFM branch itype for arguments t: returns F[m, n] for m = 0..lam."""
    t = np.ascontiguousarray(t, dtype=np.float64)
    out = np.zeros((lam + 1, t.size))
    _oracle_lib().orc_funfm(t.ctypes.data, t.size, lam, itype, out.ctypes.data)
    return out


def rmake(qp, aph, lam: int) -> np.ndarray:
    """RMAKE: R[tuv, n] for Q-P vectors qp (n,3) and exponents aph (n)."""
    qp = np.ascontiguousarray(np.asarray(qp, dtype=np.float64).T)
    aph = np.ascontiguousarray(aph, dtype=np.float64)
    n = aph.size
    nt = (lam + 1) * (lam + 2) * (lam + 3) // 6
    out = np.zeros((nt, n))
    _oracle_lib().orc_rmake(qp.ctypes.data, aph.ctypes.data, n, lam, out.ctypes.data)
    return out


class Oracle:
    def __init__(self, system, intsym: bool = True, racah1: bool = False):
        xyz, lmax, nfnc, exps, lstride, mbs = system.arrays()
        self._keep = (xyz, lmax, nfnc, exps)
        lib = _oracle_lib()
        self._h = lib.orc_create(system.ncnt, xyz.ctypes.data, lmax.ctypes.data, nfnc.ctypes.data,
                                 lstride, exps.ctypes.data, mbs, SHAPE[system.shape],
                                 1 if intsym else 0, 1 if racah1 else 0)
        self.ndim = lib.orc_ndim(self._h)
        self.nskp = lib.orc_nskp(self._h)
        self.system = system

    def close(self):
        if self._h:
            _oracle_lib().orc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def labels(self):
        n = self.nskp
        a = [np.zeros(n, dtype=np.int32) for _ in range(3)]
        _oracle_lib().orc_labels(self._h, a[0].ctypes.data, a[1].ctypes.data, a[2].ctypes.data)
        return a

    def etable(self, itt, ica, icb, ka, kb, ma, mb):
        """E0LL (itt 1), E0SS (itt 4): [4][ntuv*maxab]; EILS (itt 2): [12][...]."""
        nc = 12 if itt == 2 else 4
        cap = nc * 400000
        out = np.zeros(cap)
        n = _oracle_lib().orc_etable(self._h, itt, ica, icb, ka, kb, ma, mb, out.ctypes.data, cap)
        if n < 0:
            raise ValueError("etable capacity")
        return out[:nc * n].reshape(nc, n)

    def erifast(self, icnt, kidx, midx, ibas, jbas, itn, ilev=3):
        lc, ld = kidx[2] // 2, kidx[3] // 2
        maxcd = self.system.nfnc_of(icnt[2] - 1, lc) * self.system.nfnc_of(icnt[3] - 1, ld)
        re = np.zeros(16 * maxcd)
        im = np.zeros(16 * maxcd)
        a = [np.asarray(v, dtype=np.int32) for v in (icnt, kidx, midx, itn)]
        _oracle_lib().orc_erifast(self._h, a[0].ctypes.data, a[1].ctypes.data, a[2].ctypes.data,
                                  ibas, jbas, a[3].ctypes.data, ilev, re.ctypes.data,
                                  im.ctypes.data)
        return (re + 1j * im).reshape(16, maxcd)

    def clmdiag(self):
        """CLMDIAG's GDSC (ndim x ndim real)."""
        g = np.zeros((self.ndim, self.ndim), order="F")
        _oracle_lib().orc_clmdiag(self._h, g.ctypes.data)
        return g

    def oneel(self, znuc, xi, cv=137.035999084):
        """OVRLP/ONEEL: (OVLP, HNUC, HKIN); xi[c] <= 0 is a point nucleus."""
        n = self.ndim
        z = np.ascontiguousarray(znuc, dtype=np.float64)
        x = np.ascontiguousarray(xi, dtype=np.float64)
        out = [np.zeros((n, n), dtype=np.complex128, order="F") for _ in range(3)]
        _oracle_lib().orc_oneel(self._h, z.ctypes.data, x.ctypes.data, cv,
                                *[o.ctypes.data for o in out])
        return tuple(out)

    def biifast(self, icnt, kidx, midx, ibas, jbas, gaunt=False):
        """BIIFAST batch RR(16, maxcd) at a fresh class."""
        maxcd = (self.system.nfnc_of(icnt[2] - 1, kidx[2] // 2) *
                 self.system.nfnc_of(icnt[3] - 1, kidx[3] // 2))
        re = np.zeros(16 * maxcd)
        im = np.zeros(16 * maxcd)
        a = [np.asarray(v, dtype=np.int32) for v in (icnt, kidx, midx)]
        _oracle_lib().orc_set_gaunt(self._h, 1 if gaunt else 0)
        _oracle_lib().orc_biifast(self._h, a[0].ctypes.data, a[1].ctypes.data, a[2].ctypes.data,
                                  ibas, jbas, re.ctypes.data, im.ctypes.data)
        return (re + 1j * im).reshape(16, maxcd)

    def brtfast(self, dent, ilev=3, nthreads=1, gaunt=False):
        """BRTFAST: (BDIR, BXCH) for total density dent."""
        n = self.ndim
        d = np.asfortranarray(dent, dtype=np.complex128)
        b = np.zeros((n, n), dtype=np.complex128, order="F")
        x = np.zeros((n, n), dtype=np.complex128, order="F")
        _oracle_lib().orc_set_gaunt(self._h, 1 if gaunt else 0)
        if nthreads > 1:
            _oracle_lib().orc_brtfast_mt(self._h, ilev, d.ctypes.data, b.ctypes.data,
                                         x.ctypes.data, nthreads)
        else:
            _oracle_lib().orc_brtfast(self._h, ilev, d.ctypes.data, b.ctypes.data, x.ctypes.data)
        return b, x

    def brtfast_sample(self, dent, ilev=3, nthreads=1, stride=1, gaunt=False):
        """Timing sample of BRTFAST, as clmfast_sample: (fraction of the
        build's BIIFAST calls made, BDIR, BXCH)."""
        n = self.ndim
        d = np.asfortranarray(dent, dtype=np.complex128)
        b = np.zeros((n, n), dtype=np.complex128, order="F")
        x = np.zeros((n, n), dtype=np.complex128, order="F")
        _oracle_lib().orc_set_gaunt(self._h, 1 if gaunt else 0)
        frac = _oracle_lib().orc_brtfast_sample(self._h, ilev, d.ctypes.data, b.ctypes.data,
                                                x.ctypes.data, nthreads, stride)
        return frac, b, x

    def pairbound(self):
        """sqrt|(ab|ba)| table (ndim x ndim real)."""
        g = np.zeros((self.ndim, self.ndim), order="F")
        _oracle_lib().orc_pairbound(self._h, g.ctypes.data)
        return g

    def schwarz_stats(self):
        """(largest dropped digestion term, number dropped) of the last
        single-threaded screened build."""
        a, b = C.c_double(), C.c_double()
        _oracle_lib().orc_schwarz_stats(self._h, C.byref(a), C.byref(b))
        return a.value, b.value

    def schwarz_probe(self, g, dent, nlab, ibas, jbas, nb3, nb4, sens):
        """SCHWARZ's masks for one ERIFAST call on bound table g: int[nb3*nb4],
        bit b keeps batch b+1, 0 drops the kl (kl = k * nb4 + l)."""
        gf = np.asfortranarray(g, dtype=np.float64)
        d = np.asfortranarray(dent, dtype=np.complex128)
        lab = np.ascontiguousarray(nlab, dtype=np.int32)
        msk = np.zeros(nb3 * nb4, dtype=np.int32)
        _oracle_lib().orc_schwarz_probe(self._h, gf.ctypes.data, d.ctypes.data, lab.ctypes.data,
                                        ibas, jbas, nb3, nb4, sens, msk.ctypes.data)
        return msk

    def clmfast_sample(self, dent, ilev=3, nthreads=1, stride=1):
        """Timing sample of CLMFAST (classes as work items, every stride-th
        primitive pair ij): (fraction of the build's ERIFAST calls made,
        GDIR, GXCH partial sums; the whole build for stride 1)."""
        n = self.ndim
        d = np.asfortranarray(dent, dtype=np.complex128)
        g = np.zeros((n, n), dtype=np.complex128, order="F")
        x = np.zeros((n, n), dtype=np.complex128, order="F")
        _oracle_lib().orc_set_schwarz(self._h, 0, 0.0)
        frac = _oracle_lib().orc_clmfast_sample(self._h, ilev, d.ctypes.data, g.ctypes.data,
                                                x.ctypes.data, nthreads, stride)
        return frac, g, x

    def clmfast(self, dent, ilev=3, nthreads=1, schwarz=False, sens=0.0):
        n = self.ndim
        d = np.asfortranarray(dent, dtype=np.complex128)
        g = np.zeros((n, n), dtype=np.complex128, order="F")
        x = np.zeros((n, n), dtype=np.complex128, order="F")
        mode = 2 if schwarz == "rigorous" else (1 if schwarz else 0)
        _oracle_lib().orc_set_schwarz(self._h, mode, sens)
        if nthreads > 1:
            _oracle_lib().orc_clmfast_mt(self._h, ilev, d.ctypes.data, g.ctypes.data,
                                         x.ctypes.data, nthreads)
        else:
            _oracle_lib().orc_clmfast(self._h, ilev, d.ctypes.data, g.ctypes.data, x.ctypes.data)
        return g, x

    def coulomb1(self, dent, g, x):
        """COULOMB1 for every centre on top of a CLMFAST result g, x (built by
        an Oracle with racah1=True); returns the completed matrices."""
        d = np.asfortranarray(dent, dtype=np.complex128)
        g = np.array(g, dtype=np.complex128, order="F")
        x = np.array(x, dtype=np.complex128, order="F")
        _oracle_lib().orc_coulomb1(self._h, d.ctypes.data, g.ctypes.data, x.ctypes.data)
        return g, x

    def breit1(self, dent, b, x):
        """BREIT1 for every centre on top of a BRTFAST result (b, x) of an
        Oracle with racah1=True; returns the completed matrices."""
        d = np.asfortranarray(dent, dtype=np.complex128)
        b = np.array(b, dtype=np.complex128, order="F")
        x = np.array(x, dtype=np.complex128, order="F")
        _oracle_lib().orc_breit1(self._h, d.ctypes.data, b.ctypes.data, x.ctypes.data)
        return b, x

    def brtfast_racah(self, dent, ilev=3, nthreads=1):
        """HFSCF's Breit build with RACAH1 on: BRTFAST without one-centre
        quartets, then BREIT1 per centre (construct with racah1=True)."""
        b, x = self.brtfast(dent, ilev, nthreads)
        return self.breit1(dent, b, x)

    def clmfast_racah(self, dent, ilev=3, nthreads=1):
        """HFSCF's Coulomb build with RACAH1 on: CLMFAST without one-centre
        quartets, then COULOMB1 per centre (construct with racah1=True)."""
        g, x = self.clmfast(dent, ilev, nthreads)
        return self.coulomb1(dent, g, x)


_rlib = None


def reference_available() -> bool:
    return os.path.exists(REF_LIB)


def _ref_lib():
    global _rlib
    if _rlib is None:
        _rlib = C.CDLL(REF_LIB)
    return _rlib


class Reference:
    """The reference's CLMFAST/ERIFAST/EQSAVE, driven through ref_driver.f.
    One system at a time (the routines keep their state in COMMON)."""
    MBS, MEL = 26, 6

    def __init__(self, system, intsym: bool = True, racah1: bool = False):
        lib = _ref_lib()
        n = system.ncnt
        if n > 2:
            raise ValueError("the reference build holds at most 2 centres (parameters.h MCT)")
        xyz = np.ascontiguousarray(system.xyz, dtype=np.float64)
        lmax = np.array([system.lmax(c) for c in range(n)], dtype=np.int32)
        nf = np.zeros((n, self.MEL + 1), dtype=np.int32)
        ex = np.zeros((n, self.MEL + 1, self.MBS))
        for c in range(n):
            for l, e in enumerate(system.shells[c]):
                if len(e) > self.MBS:
                    raise ValueError("more than 26 exponents in a shell (parameters.h MBS)")
                nf[c, l] = len(e)
                ex[c, l, :len(e)] = e
        nd, ns = C.c_int(), C.c_int()
        lib.ref_setup(C.c_int(n), C.c_void_p(xyz.ctypes.data), C.c_void_p(lmax.ctypes.data),
                      C.c_void_p(nf.ctypes.data), C.c_void_p(ex.ctypes.data),
                      C.c_int(SHAPE[system.shape]), C.c_int(1 if system.hmlt == "DHFR" else 2),
                      C.c_int(1 if racah1 else 0), C.c_int(1 if intsym else 0),
                      C.byref(nd), C.byref(ns))
        self.ndim, self.nskp, self.system = nd.value, ns.value, system

    def labels(self):
        n = self.nskp
        a = [np.zeros(n, dtype=np.int32) for _ in range(3)]
        lrge = np.zeros(2 * 13 * 14, dtype=np.int32)
        _ref_lib().ref_labels(C.c_void_p(a[0].ctypes.data), C.c_void_p(a[1].ctypes.data),
                              C.c_void_p(a[2].ctypes.data), C.c_void_p(lrge.ctypes.data))
        return a

    def etable(self, itt, ica, icb, ka, kb, ma, mb):
        nc = 12 if itt == 2 else 4
        cap = nc * 400000
        out = np.zeros(cap)
        nt, mx = C.c_int(), C.c_int()
        _ref_lib().ref_etable(C.c_int(itt), C.c_int(ica), C.c_int(icb), C.c_int(ka), C.c_int(kb),
                              C.c_int(ma), C.c_int(mb), C.c_int(cap), C.c_void_p(out.ctypes.data),
                              C.byref(nt), C.byref(mx))
        n = nt.value * mx.value
        return out[:nc * n].reshape(nc, n)

    def erifast(self, icnt, kidx, midx, ibas, jbas, itn, ilev=3):
        a = [np.asarray(v, dtype=np.int32) for v in (icnt, kidx, midx, itn)]
        lc, ld = kidx[2] // 2, kidx[3] // 2
        maxcd = self.system.nfnc_of(icnt[2] - 1, lc) * self.system.nfnc_of(icnt[3] - 1, ld)
        re = np.zeros(16 * maxcd)
        im = np.zeros(16 * maxcd)
        mo = C.c_int()
        _ref_lib().ref_erifast(*[C.c_void_p(x.ctypes.data) for x in a[:3]], C.c_int(ibas),
                               C.c_int(jbas), C.c_void_p(a[3].ctypes.data), C.c_int(ilev),
                               C.c_void_p(re.ctypes.data), C.c_void_p(im.ctypes.data),
                               C.byref(mo))
        return (re + 1j * im).reshape(16, maxcd)

    def biifast(self, icnt, kidx, midx, ibas, jbas, gaunt=False):
        a = [np.asarray(v, dtype=np.int32) for v in (icnt, kidx, midx)]
        maxcd = (self.system.nfnc_of(icnt[2] - 1, kidx[2] // 2) *
                 self.system.nfnc_of(icnt[3] - 1, kidx[3] // 2))
        re = np.zeros(16 * maxcd)
        im = np.zeros(16 * maxcd)
        mo = C.c_int()
        _ref_lib().ref_biifast(*[C.c_void_p(x.ctypes.data) for x in a], C.c_int(ibas),
                               C.c_int(jbas), C.c_int(1 if gaunt else 0),
                               C.c_void_p(re.ctypes.data), C.c_void_p(im.ctypes.data),
                               C.byref(mo))
        return (re + 1j * im).reshape(16, maxcd)

    def brtfast(self, dent, ilev=3):
        n = self.ndim
        d = np.asfortranarray(dent, dtype=np.complex128)
        dre = np.asfortranarray(d.real)
        dim = np.asfortranarray(d.imag)
        out = [np.zeros((n, n), order="F") for _ in range(4)]
        _ref_lib().ref_brtfast(C.c_int(ilev), C.c_void_p(dre.ctypes.data),
                               C.c_void_p(dim.ctypes.data),
                               *[C.c_void_p(o.ctypes.data) for o in out])
        return out[0] + 1j * out[1], out[2] + 1j * out[3]

    def oneel(self, znuc, anuc, gauss=False):
        """ONEEL: (OVLP, HNUC, HKIN) for charges znuc, mass numbers anuc."""
        n = self.ndim
        z = np.ascontiguousarray(znuc, dtype=np.float64)
        a = np.ascontiguousarray(anuc, dtype=np.float64)
        out = [np.zeros(n * n) for _ in range(6)]
        _ref_lib().ref_oneel(C.c_void_p(z.ctypes.data), C.c_void_p(a.ctypes.data),
                             C.c_int(1 if gauss else 0), *[C.c_void_p(o.ctypes.data) for o in out])
        m = [o.reshape(n, n, order="F") for o in out]
        return m[0] + 1j * m[1], m[2] + 1j * m[3], m[4] + 1j * m[5]

    def schwarz_masks(self, g, dent, nlab, ibas, jbas, nb3, nb4, sens):
        """The reference's SCHWARZ on bound table g (passed as its GDIR
        argument) and density dent: (IMTX[11][nb3*nb4], ISCR[nb3*nb4], MAXN)."""
        n = self.ndim
        gf = np.asfortranarray(g, dtype=np.float64)
        d = np.asfortranarray(dent, dtype=np.complex128)
        dre = np.asfortranarray(d.real)
        dim = np.asfortranarray(d.imag)
        lab = np.ascontiguousarray(nlab, dtype=np.int32)
        m = nb3 * nb4
        imtx = np.zeros(11 * m, dtype=np.int32)
        iscr = np.zeros(m, dtype=np.int32)
        maxn = C.c_int()
        assert gf.shape == (n, n)
        _ref_lib().ref_schwarz_masks(C.c_void_p(gf.ctypes.data), C.c_void_p(dre.ctypes.data),
                                     C.c_void_p(dim.ctypes.data), C.c_void_p(lab.ctypes.data),
                                     C.c_int(ibas), C.c_int(jbas), C.c_int(nb3), C.c_int(nb4),
                                     C.c_double(sens), C.c_void_p(imtx.ctypes.data),
                                     C.c_void_p(iscr.ctypes.data), C.byref(maxn))
        return imtx.reshape(11, m), iscr, maxn.value

    @staticmethod
    def _planes(m):
        m = np.asfortranarray(m, dtype=np.complex128)
        return np.asfortranarray(m.real), np.asfortranarray(m.imag)

    def densty(self, coef, icls):
        """DENSTY with COEF and closed-shell spinors icls (1-based, counted
        from NSKP): DENT."""
        n = self.ndim
        cre, cim = self._planes(coef)
        ic = np.ascontiguousarray(icls, dtype=np.int32)
        dre, dim = np.zeros((n, n), order="F"), np.zeros((n, n), order="F")
        _ref_lib().ref_densty(C.c_void_p(cre.ctypes.data), C.c_void_p(cim.ctypes.data),
                              C.c_int(ic.size), C.c_void_p(ic.ctypes.data),
                              C.c_void_p(dre.ctypes.data), C.c_void_p(dim.ctypes.data))
        return dre + 1j * dim

    def levshft(self, fock, ovlp, coef, nocc, shlv):
        """LEVSHFT: the shifted FOCK."""
        fre, fim = self._planes(fock)
        ore, oim = self._planes(ovlp)
        cre, cim = self._planes(coef)
        _ref_lib().ref_levshft(*[C.c_void_p(x.ctypes.data) for x in (fre, fim, ore, oim, cre, cim)],
                               C.c_int(nocc), C.c_double(shlv))
        return fre + 1j * fim

    def enrgies(self, dent, hkin, hnuc, fock, gdir, gxch, bdir, bxch, enuc, breit):
        """ENRGIES, closed shell: (ETOT, EONE, EHKN, EHNC, EGDR, EGXC, EBDR,
        EBXC)."""
        planes = []
        for m in (dent, hkin, hnuc, fock, gdir, gxch, bdir, bxch):
            planes += list(self._planes(m))
        a = np.concatenate([p.reshape(-1, order="F") for p in planes])
        out = np.zeros(8)
        _ref_lib().ref_enrgies(C.c_void_p(a.ctypes.data), C.c_double(enuc),
                               C.c_int(2 if breit else 1), C.c_void_p(out.ctypes.data))
        return out

    def clmdiag(self):
        """CLMDIAG's GDSC (ndim x ndim real)."""
        g = np.zeros(self.ndim * self.ndim)
        _ref_lib().ref_clmdiag(C.c_void_p(g.ctypes.data))
        return g.reshape(self.ndim, self.ndim, order="F")

    def clmfast(self, dent, ilev=3, schwarz=False):
        n = self.ndim
        _ref_lib().ref_schwarz(C.c_int(1 if schwarz else 0))
        d = np.asfortranarray(dent, dtype=np.complex128)
        dre = np.asfortranarray(d.real)
        dim = np.asfortranarray(d.imag)
        out = [np.zeros((n, n), order="F") for _ in range(4)]
        _ref_lib().ref_clmfast(C.c_int(ilev), C.c_void_p(dre.ctypes.data),
                               C.c_void_p(dim.ctypes.data),
                               *[C.c_void_p(o.ctypes.data) for o in out])
        return out[0] + 1j * out[1], out[2] + 1j * out[3]

    def brtfast_racah(self, dent, ilev=3):
        """BRTFAST followed by BREIT1 for every centre, as HFSCF runs them
        with RACAH1 on (construct the Reference with racah1=True)."""
        n = self.ndim
        d = np.asfortranarray(dent, dtype=np.complex128)
        dre = np.asfortranarray(d.real)
        dim = np.asfortranarray(d.imag)
        out = [np.zeros((n, n), order="F") for _ in range(4)]
        _ref_lib().ref_brtfast_racah(C.c_int(ilev), C.c_void_p(dre.ctypes.data),
                                     C.c_void_p(dim.ctypes.data),
                                     *[C.c_void_p(o.ctypes.data) for o in out])
        return out[0] + 1j * out[1], out[2] + 1j * out[3]

    def clmfast_racah(self, dent, ilev=3):
        """CLMFAST followed by COULOMB1 for every centre, as HFSCF runs them
        with RACAH1 on (construct the Reference with racah1=True)."""
        n = self.ndim
        d = np.asfortranarray(dent, dtype=np.complex128)
        dre = np.asfortranarray(d.real)
        dim = np.asfortranarray(d.imag)
        out = [np.zeros((n, n), order="F") for _ in range(4)]
        _ref_lib().ref_schwarz(C.c_int(0))
        _ref_lib().ref_clmfast_racah(C.c_int(ilev), C.c_void_p(dre.ctypes.data),
                                     C.c_void_p(dim.ctypes.data),
                                     *[C.c_void_p(o.ctypes.data) for o in out])
        return out[0] + 1j * out[1], out[2] + 1j * out[3]

    def funfm(self, t, lam, itype):
        t = np.ascontiguousarray(t, dtype=np.float64)
        out = np.zeros((lam + 1) * t.size)
        _ref_lib().ref_funfm(C.c_void_p(t.ctypes.data), C.c_int(t.size), C.c_int(lam),
                             C.c_int(itype), C.c_void_p(out.ctypes.data))
        return out.reshape(lam + 1, t.size)

    def rmake(self, qp, aph, lam):
        qp = np.asfortranarray(np.asarray(qp, dtype=np.float64))
        aph = np.ascontiguousarray(aph, dtype=np.float64)
        n = aph.size
        nt = (lam + 1) * (lam + 2) * (lam + 3) // 6
        out = np.zeros(nt * n)
        _ref_lib().ref_rmake(C.c_void_p(qp.ctypes.data), C.c_void_p(aph.ctypes.data),
                             C.c_int(n), C.c_int(lam), C.c_void_p(out.ctypes.data))
        return out.reshape(nt, n)
