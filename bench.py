#!/usr/bin/env python3
"""Fock-build throughput of the MI355X two-electron engine.

One step is one stage-3 Dirac-Coulomb Fock build (CLMFAST semantics: GDIR and
GXCH for a density) of the workload, with the density, E tables and outputs
resident in HBM.  Default workload: the water triatomic (O 14s9p4d, H 8s3p,
NDIM 380, seeded geometry).  Prints one JSON line.

Legs:
  timed     K builds back to back on the engine's stream (HIP events and a
            device synchronisation on both sides), FP64 and then mixed
            precision; the headline value is the mixed build when its
            two-electron energy is within 1 microhartree of the FP64 build's
  kernel    average duration of the quartet kernel (HIP events bracketing it
            on the engine's stream) -> roofline
  cpu       the oracle (C restatement of the reference path) timed on the
            host's cores on the same build, and the GPU result checked
            against it (max |dG|, |dE_2e|)
  serve     Poisson-arriving DHF solves through the continuous-batching
            server: solves/s with arrivals above capacity, then p50/p99
            latency with arrivals below it (serving_sustained)
  breit     (--breit K) K Breit builds of the same density, kernel time

The 12-centre ring (NDIM 1368) runs as
  python bench.py --workload ring --schwarz rigorous --density atoms --no-cpu-baseline --serve 0
(Schwarz screening at SENS 1e-12 with every dropped term below it,
superposed atomic densities; its oracle build takes tens of minutes, hence no
CPU leg).
"""
from __future__ import annotations

import argparse
import itertools
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Fock builds/s, mixed vs FP64 at |ΔE|<1µEh; served SCF p50/p99 latency"
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 (vector and matrix)


def profiled_traffic(workload: str):
    """HBM bytes per fock_kernel launch for this workload, from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary (calibrated against kernels of
    known traffic); None if there is none."""
    path = os.path.join(ROOT, "profiles", f"fock_kernel_{workload}_traffic.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        return None
    return None


def atoms_density(pkg, system):
    """Superposed densities of the system's atoms, each solved alone in its
    basis with the bare-nucleus Hamiltonian: Z electrons in the lowest
    positive-energy Kramers pairs, an odd electron shared by the last pair.
    Hermitian, positive and normalised (elements below 1), unlike a random
    test density; centres with the same charge and basis share one solve."""
    import importlib

    import torch
    scf = importlib.import_module(pkg.__name__ + ".scf")
    cache = {}
    blocks = []
    for c in range(system.ncnt):
        z = float(system.charges[c])
        key = (z, tuple(tuple(float(x) for x in e) for e in system.shells[c]))
        if key not in cache:
            atom = pkg.systems.System("atom", np.zeros((1, 3)), [system.shells[c]],
                                      shape="ATOMIC", charges=[z])
            be = scf.EngineBackend(pkg, atom)
            try:
                ovlp, hnuc, hkin = be.one_electron(137.035999084)
                nskp = be.nskp
            finally:
                be.close()
            _, coef, finite = scf.solve_fock(hnuc + hkin, scf.orthogonaliser(ovlp))
            if not bool(finite):
                raise RuntimeError("atomic guess: eigensolver returned non-finite vectors")
            nel = int(round(z))
            full, odd = nel // 2, nel % 2
            occ = coef[:, nskp:nskp + 2 * (full + odd)]
            w = torch.ones(occ.shape[1], dtype=torch.float64, device=occ.device)
            if odd:
                w[-2:] = 0.5
            cache[key] = (occ.conj() * w) @ occ.T
        blocks.append(cache[key])
    return scf.superposed_density(system, blocks)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="water", choices=["water", "ne", "tlf", "ring"])
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--stage", type=int, default=3, choices=[1, 2, 3],
                    help="integral-inclusion stage of every leg (ILEV; 3 = full Dirac-Coulomb)")
    ap.add_argument("--schwarz", nargs="?", const="reference", default=None,
                    choices=["reference", "rigorous"],
                    help="SCHWARZ masks at SENS 1e-12 on every leg, the oracle's included: the "
                         "reference's (GDSC bounds) or rigorous (Q(a,b) Q(c,d) on every group)")
    ap.add_argument("--density", default="random", choices=["random", "atoms"],
                    help="random Hermitian test density, or superposed bare-nucleus atoms")
    ap.add_argument("--breit", type=int, default=0, metavar="K",
                    help="also time K Breit builds (BRTFAST) of the same density (0: skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: min(16, cpu count)")
    ap.add_argument("--split", type=float, default=0.1,
                    help="Schwarz split threshold of the mixed-precision leg (0: skip the leg)")
    ap.add_argument("--serve", type=int, default=32,
                    help="requests of the served-SCF leg (0: skip the leg)")
    ap.add_argument("--serve-rate", type=float, default=8.0,
                    help="Poisson arrivals per second of the served-SCF leg (above capacity: throughput)")
    ap.add_argument("--serve-sustained-rate", type=float, default=1.2,
                    help="arrival rate of a second served-SCF leg below capacity (latency; 0: skip)")
    ap.add_argument("--serve-slots", type=int, default=8)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device-to-device (a kernel of known traffic that "
                         "calibrates FETCH_SIZE/WRITE_SIZE counters in a profiled run)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this engine runs on one GPU; the multi-GPU design (partial J/K "
              "all-reduce) is not built. Use --gpus 1.", file=sys.stderr)
        return 2

    import torch
    from __graft_entry__ import load_package
    pkg = load_package()
    S = pkg.systems
    kw = {"seed": args.seed} if args.workload != "ne" else {}
    if args.workload == "tlf":
        kw["hmlt"] = "DHFR"
    system = S.config_system(args.workload, **kw)
    schwarz = {None: False, "reference": True, "rigorous": "rigorous"}[args.schwarz]
    eng = pkg.Engine(system, schwarz=schwarz)
    n = eng.ndim
    flops, quartets, items = eng.plan_info(args.stage)
    dev = torch.device("cuda:0")
    if args.density == "atoms":
        dens = np.asfortranarray(atoms_density(pkg, system).cpu().numpy())
    else:
        dens = S.random_density(n, seed=11, real=(system.shape != "NOSYMM"))

    if args.copy_calibration > 0:
        nbytes = args.copy_calibration << 20
        # a fill kernel (writes nbytes) and an elementwise scale (reads and
        # writes nbytes each): their counter readings give the bytes-per-count
        # factors of the profiled run
        src = torch.ones(nbytes // 8, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        torch.mul(src, 2.0, out=dst)
        torch.cuda.synchronize()
        print(f"copy calibration: fill and scale of {nbytes} bytes", file=sys.stderr, flush=True)
        del src, dst
    # column-major matrices: a row-major torch tensor of D^T has D's layout
    d_t = torch.from_numpy(np.ascontiguousarray(dens.T)).to(dev)
    g_t = torch.zeros((n, n), dtype=torch.complex128, device=dev)
    x_t = torch.zeros((n, n), dtype=torch.complex128, device=dev)
    stream = torch.cuda.ExternalStream(eng.stream(), device=dev)
    torch.cuda.synchronize()

    def build():
        eng.clmfast_device(d_t.data_ptr(), g_t.data_ptr(), x_t.data_ptr(), args.stage)

    for _ in range(args.warmup):
        build()
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    ev0.record(stream)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        build()
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = ev0.elapsed_time(ev1) / args.steps
    value = 1000.0 / ms_per_step

    # quartet-kernel duration (events bracket only that kernel)
    kms = []
    for _ in range(min(args.steps, 10)):
        build()
        torch.cuda.synchronize()
        kms.append(eng.stats()[1])
    kernel_ms = float(np.mean(kms))
    achieved = flops / (kernel_ms * 1e-3) / 1e12

    g = g_t.cpu().numpy().T
    x = x_t.cpu().numpy().T

    # mixed-precision leg: same build with FP32 items below the Schwarz split
    mixed = None
    if args.split > 0:
        eng_m = pkg.Engine(system, precision="mixed", split_threshold=args.split,
                           schwarz=schwarz)
        stream_m = torch.cuda.ExternalStream(eng_m.stream(), device=dev)
        gm_t = torch.zeros_like(g_t)
        xm_t = torch.zeros_like(x_t)

        def build_m():
            eng_m.clmfast_device(d_t.data_ptr(), gm_t.data_ptr(), xm_t.data_ptr(), args.stage)

        for _ in range(args.warmup):
            build_m()
        torch.cuda.synchronize()
        ev0.record(stream_m)
        for _ in range(args.steps):
            build_m()
        ev1.record(stream_m)
        torch.cuda.synchronize()
        ms_m = ev0.elapsed_time(ev1) / args.steps
        build_m()
        torch.cuda.synchronize()
        kms_m = eng_m.stats()[1]
        gm = gm_t.cpu().numpy().T
        xm = xm_t.cpu().numpy().T
        de_m = 0.5 * np.real(np.sum(dens * ((gm - xm) - (g - x))))
        mixed = {"value": 1000.0 / ms_m, "ms_per_step": ms_m, "split_threshold": args.split,
                 "abs_dE2e_vs_fp64_Eh": float(abs(de_m)),
                 "max_abs_dG_vs_fp64": float(max(np.abs(gm - g).max(), np.abs(xm - x).max())),
                 "speedup_vs_fp64": ms_per_step / ms_m,
                 "quartets_per_s": quartets / (kms_m * 1e-3)}
        eng_m.close()

    # Breit leg: BRTFAST of the same density on an engine with the LS tables
    breit = None
    if args.breit > 0:
        import dataclasses
        eng_b = pkg.Engine(dataclasses.replace(system, hmlt="DHFB"))
        bflops, bquartets, bitems = eng_b.breit_plan_info(args.stage)
        b_t = torch.zeros_like(g_t)
        xb_t = torch.zeros_like(x_t)
        eng_b.breit_device(d_t.data_ptr(), b_t.data_ptr(), xb_t.data_ptr(), args.stage)
        torch.cuda.synchronize()
        bms = []
        for _ in range(args.breit):
            eng_b.breit_device(d_t.data_ptr(), b_t.data_ptr(), xb_t.data_ptr(), args.stage)
            torch.cuda.synchronize()
            bms.append(eng_b.stats()[1])
            print(f"breit: build {len(bms)} of {args.breit}, {bms[-1]:.1f} ms", file=sys.stderr,
                  flush=True)
        bmean = float(np.mean(bms))
        breit = {"builds": args.breit, "kernel_ms": bmean, "flops_per_build": bflops,
                 "quartets_per_build": bquartets, "work_items": bitems,
                 "tflops": bflops / (bmean * 1e-3) / 1e12}
        eng_b.close()

    cpu = None
    parity = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        orc = pyoracle.Oracle(system)
        t0 = time.perf_counter()
        go, xo = orc.clmfast(dens, args.stage, nthreads=threads, schwarz=schwarz)
        tc = time.perf_counter() - t0
        cpu = {"value": 1.0 / tc, "unit": "Fock builds/s", "cores": threads, "kind": "port",
               "sample": f"one full stage-{args.stage} Dirac-Coulomb build of the {system.name} workload "
                         f"(NDIM {n}), same density, oracle/bdf_oracle.c with OpenMP over centre "
                         f"quadruples: {tc:.2f} s"}
        de = 0.5 * np.real(np.sum(dens * ((g - x) - (go - xo))))
        parity = {"max_abs_dGDIR": float(np.abs(g - go).max()),
                  "max_abs_dGXCH": float(np.abs(x - xo).max()),
                  "abs_dE2e_Eh": float(abs(de))}

    # served-SCF leg: DHF solves of atoms and diatomics arriving over time
    serving = None
    serving_sustained = None
    if args.serve > 0:
        import importlib
        sv = importlib.import_module(pkg.__name__ + ".serve")

        def serve_leg(rate):
            reqs = sv.request_mix(args.serve, seed=args.seed, rate=rate)
            finished = itertools.count(1)

            def progress(r):
                k = next(finished)  # atomic under the GIL: workers call this concurrently
                if k % 32 == 0:
                    print(f"serve: {k} of {len(reqs)} solves done (rate {rate}/s)",
                          file=sys.stderr, flush=True)

            server = sv.Server(pkg, slots=args.serve_slots, on_result=progress)
            t0 = time.perf_counter()
            res = server.run(reqs)
            span = time.perf_counter() - t0
            p50, p99 = sv.percentiles(res)
            return {"requests": len(res), "p50_latency_s": p50, "p99_latency_s": p99,
                    "solves_per_s": len(res) / span, "wall_s": span,
                    "converged": sum(1 for r in res if r.converged),
                    "arrival_rate_per_s": rate, "slots": args.serve_slots,
                    "peak_concurrent": server.peak_active,
                    "ndim_range": [min(r.ndim for r in res), max(r.ndim for r in res)],
                    "mix": "60% atoms Z in {2,4,10,12,18,20,36,54}, 40% diatomics of the "
                           "same elements, R ~ U(2.5, 8.0) bohr, Poisson arrivals"}

        # arrivals faster than the solves (throughput), then below capacity
        # (latency of a stream the server keeps up with)
        serving = serve_leg(args.serve_rate)
        if args.serve_sustained_rate > 0:
            serving_sustained = serve_leg(args.serve_sustained_rate)

    # headline: the mixed build when it meets the metric's |dE| < 1 microhartree
    # against the FP64 build, else the FP64 build
    headline = mixed if (mixed is not None and mixed["abs_dE2e_vs_fp64_Eh"] < 1e-6) else None
    out = {
        "metric": METRIC,
        "value": headline["value"] if headline else value,
        "unit": "Fock builds/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": headline["ms_per_step"] if headline else ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32 (Schwarz split)" if headline else "f64",
        "data": ("superposed bare-nucleus atomic densities" if args.density == "atoms"
                 else "synthetic Hermitian density") +
                "; even-tempered basis and seeded geometry of the workload",
        "config": {"workload": system.name, "ndim": n,
                   "precision": (f"mixed, FP32 below split {args.split}" if headline else "fp64"),
                   "stage": args.stage, "schwarz": args.schwarz, "density": args.density,
                   "quartets_per_build": quartets, "work_items": items,
                   "host_wall_s": wall},
        "fp64": {"value": value, "ms_per_step": ms_per_step,
                 "quartets_per_s": quartets / (kernel_ms * 1e-3)},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                     "traffic": profiled_traffic(args.workload),
                     "kernel": "fock_kernel<double> (FP64 build)", "kernel_ms": kernel_ms,
                     "flops_per_launch": flops},
        "cpu_baseline": cpu,
        "parity": parity,
        "mixed": mixed,
        "breit": breit,
        "serving": serving,
        "serving_sustained": serving_sustained,
    }
    print(json.dumps(out), flush=True)
    eng.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
