"""Autotuner of the quartet kernels' launch shapes.

For a system, times the stage-3 Coulomb build under every launch shape of
fock_kernel (waves per SIMD, group capacity, persistent workgroups per CU)
and, for Dirac-Coulomb-Breit systems, the Breit build on the matrix-core
kernel (where the system's path has it) and on the vector-ALU breit_kernel
under every workgroup count and AB-combination group, and records the
fastest in tuned_launch.json beside this file, keyed by the system's
signature: integral path (real for
ATOMIC/DIATOM/LINEAR, complex otherwise), Hamiltonian, largest primitive
count per shell and largest l.  Engine(..., tuned=True) applies the entry
of its system's signature through bdf_set_launch.

    python bertha-program_amd/tune.py WORKLOAD [WORKLOAD ...]

(run on the GPU; WORKLOAD is a name of systems.CONFIGS).  No counterpart in
the reference, whose loop nest has no launch parameters.
"""
from __future__ import annotations

import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "tuned_launch.json")

# (waves, cap, workgroups per CU) of fock_kernel: the instantiated shapes
# (three waves at capacity 1, two at capacity 1 or 2) at the occupancies
# their registers allow and below
COULOMB_SHAPES = [(4, 1, 16), (3, 1, 12), (3, 1, 8), (2, 2, 8), (2, 1, 8), (2, 2, 6),
                  (2, 1, 6), (2, 2, 4)]
# (workgroups per CU, AB group); one wave per SIMD for each four workgroups
BREIT_SHAPES = [(16, 1), (12, 1), (8, 1), (8, 2)]


def signature(system) -> str:
    path = "real" if system.shape.upper() in ("ATOMIC", "DIATOM", "LINEAR") else "complex"
    maxn = max(system.nfnc_of(c, l) for c in range(system.ncnt)
               for l in range(system.lmax(c) + 1))
    lmax = max(system.lmax(c) for c in range(system.ncnt))
    return f"{path}/{system.hmlt.upper()}/maxprim{maxn}/lmax{lmax}"


def load_table() -> dict:
    try:
        with open(TABLE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def lookup(system):
    """The tuned launch settings of the system's signature, or None."""
    return load_table().get(signature(system))


def _time(eng, build, reps: int) -> float:
    import torch
    build()
    torch.cuda.synchronize()
    t = []
    for _ in range(reps):
        build()
        torch.cuda.synchronize()
        t.append(eng.stats()[1])
    return float(np.median(t))


def tune(pkg, system, reps: int = 3, log=print) -> dict:
    """Times every launch shape on the GPU; returns the table entry."""
    import torch
    dev = torch.device("cuda:0")
    eng = pkg.Engine(system, tuned=False)
    n = eng.ndim
    d = torch.from_numpy(np.ascontiguousarray(
        pkg.systems.random_density(n, seed=11, real=system.shape != "NOSYMM").T)).to(dev)
    g, x = (torch.zeros((n, n), dtype=torch.complex128, device=dev) for _ in range(2))
    default = eng.launch_info(3)
    coulomb = {}
    for waves, cap, per_cu in COULOMB_SHAPES:
        eng.set_launch(waves=waves, cap=cap, per_cu=per_cu)
        ms = _time(eng, lambda: eng.clmfast_device(d.data_ptr(), g.data_ptr(), x.data_ptr(), 3),
                   reps)
        coulomb[(waves, cap, per_cu)] = ms
        log(f"tune {signature(system)}: fock_kernel waves {waves} cap {cap} per CU {per_cu}: "
            f"{ms:.2f} ms")
    best = min(coulomb, key=coulomb.get)
    dshape = (default[0], default[1], 4 * default[0])
    entry = {"waves": best[0], "cap": best[1], "per_cu": best[2], "coulomb_ms": coulomb[best],
             "default": {"waves": dshape[0], "cap": dshape[1], "per_cu": dshape[2],
                         "coulomb_ms": coulomb.get(dshape)},
             "workload": system.name, "ndim": n,
             "tuned": time.strftime("%Y-%m-%d")}
    if system.hmlt.upper() == "DHFB":
        # (kernel, workgroups per CU, AB group): the matrix-core kernel has
        # one launch shape; the vector-ALU kernel's shapes follow
        breit = {}
        shapes = [(2, 0, 0)] + [(1, per_cu, group) for per_cu, group in BREIT_SHAPES]
        for kern, per_cu, group in shapes:
            try:
                eng.set_launch(waves=best[0], cap=best[1], per_cu=best[2], breit_per_cu=per_cu,
                               breit_group=group, breit_kernel=kern)
            except pkg.BerthaError:
                continue  # no matrix-core kernel on this path
            ms = _time(eng, lambda: eng.breit_device(d.data_ptr(), g.data_ptr(), x.data_ptr(), 3),
                       max(1, reps - 1))
            breit[(kern, per_cu, group)] = ms
            log(f"tune {signature(system)}: Breit kernel {kern} per CU {per_cu} group {group}: "
                f"{ms:.2f} ms")
        bbest = min(breit, key=breit.get)
        entry.update({"breit_kernel": bbest[0], "breit_per_cu": bbest[1], "breit_group": bbest[2],
                      "breit_ms": breit[bbest]})
        dkey = (2, 0, 0) if (2, 0, 0) in breit else (1, 16, 1)
        entry["default"].update({"breit_kernel": dkey[0], "breit_per_cu": dkey[1],
                                 "breit_group": dkey[2], "breit_ms": breit.get(dkey)})
    eng.close()
    return entry


def main(argv) -> int:
    sys.path.insert(0, os.path.dirname(HERE))
    from __graft_entry__ import load_package
    pkg = load_package()
    table = load_table()
    for name in argv or ["tlf", "water", "ring"]:
        kw = {} if name == "ne" else {"seed": 0}
        system = pkg.systems.config_system(name, **kw)
        entry = tune(pkg, system, log=lambda m: print(m, file=sys.stderr, flush=True))
        table[signature(system)] = entry
        with open(TABLE, "w") as f:
            json.dump(table, f, indent=1, sort_keys=True)
        print(json.dumps({signature(system): entry}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
