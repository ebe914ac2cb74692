"""Oracle-backed matrices for the SCF driver (test infrastructure): the same
HFSCF iteration run on the oracle's CPU restatement, to compare against the
engine-backed run."""
import numpy as np
import torch


class OracleBackend:
    def __init__(self, pkg, orc, system, nthreads: int = 8, racah1: bool = False):
        self.pkg = pkg
        self.system = system
        self.racah1 = racah1
        self.o = orc.Oracle(system, racah1=racah1)
        self.n, self.nskp = self.o.ndim, self.o.nskp
        self.dev = torch.device("cpu")
        self.nthreads = nthreads

    def one_electron(self, cv):
        xi = self.pkg.systems.nuclear_exponents(self.system)
        xi = np.zeros(self.system.ncnt) if xi is None else xi
        mats = self.o.oneel(np.array(self.system.charges, float), xi, cv)
        return [torch.from_numpy(np.ascontiguousarray(x)) for x in mats]

    def coulomb(self, dens, ilev):
        if self.racah1:  # CLMFAST without one-centre quartets, then COULOMB1
            g, x = self.o.clmfast_racah(dens.numpy(), ilev, nthreads=self.nthreads)
        else:
            g, x = self.o.clmfast(dens.numpy(), ilev, nthreads=self.nthreads)
        return torch.from_numpy(np.ascontiguousarray(g)), torch.from_numpy(np.ascontiguousarray(x))

    def breit(self, dens, ilev):
        if self.racah1:  # BRTFAST without one-centre quartets, then BREIT1
            b, x = self.o.brtfast_racah(dens.numpy(), ilev, nthreads=self.nthreads)
        else:
            b, x = self.o.brtfast(dens.numpy(), ilev, nthreads=self.nthreads)
        return torch.from_numpy(np.ascontiguousarray(b)), torch.from_numpy(np.ascontiguousarray(x))
