#!/usr/bin/env python3
"""Fock-build throughput of the MI355X two-electron engine.

One step is one stage-3 Fock build of the workload with the density, E
tables and outputs resident in HBM: for the heavy diatomic (default: Tl-F,
Tl 28s24p18d12f + F 14s9p4d, NDIM 1340, Dirac-Coulomb-Breit) a Dirac-Coulomb
build (CLMFAST semantics: GDIR, GXCH) followed by a Breit build (BRTFAST:
BXCH) of the same density; for the other workloads the Dirac-Coulomb build.
The density is the superposition of the atoms' bare-nucleus ground states
(--density atoms, default) or a random Hermitian test matrix.  Prints one
JSON line.

Legs:
  fp64      K steps back to back on the engine's stream, HIP events on that
            stream and a device synchronisation on both sides
  mixed     the same with the Coulomb build's Schwarz split (FP32 below the
            split, FP64 above); the headline value is this leg when its
            two-electron energy is within 1 microhartree of the FP64 leg's
            and its step is the faster
  kernels   average duration of each quartet kernel (HIP events bracketing
            it on the engine's stream) -> roofline of the dominant one
  cpu       the oracle (the C restatement of the reference path) on the
            same step, timed on a stated sample (every S-th primitive pair
            of every class, classes as OpenMP work items) on all host cores
            and on one core, scaled by the fraction of the build's ERIFAST/
            BIIFAST calls the sample made
  serve     the serving workload's stream of 512 Poisson-arriving DHF
            solves (atoms and diatomics, NDIM 40-748) through the
            continuous-batching server at 3 arrivals/s, near its capacity:
            p50/p99 latency from arrival to converged energy and solves/s;
            then 32 solves at 1.2/s, well below capacity

Other workloads: --workload water|ne|ring (Dirac-Coulomb only), e.g. the
12-centre ring with Schwarz screening at SENS 1e-12:
  python bench.py --workload ring --schwarz rigorous --no-cpu-baseline --serve 0
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


def profiled_traffic(kernel: str, workload: str):
    """HBM bytes per launch of a kernel on a workload from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary (scaled by the counters' reading
    of kernels of known traffic), with the file it came from; (None, None)
    if there is none."""
    path = os.path.join("profiles", f"{kernel}_{workload}_traffic.md")
    try:
        with open(os.path.join(ROOT, path)) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1]), path
    except OSError:
        pass
    return None, None


def atoms_density(pkg, system):
    """Superposed densities of the system's atoms, each solved alone in its
    basis with its bare-nucleus Hamiltonian (Gaussian nuclei where the
    system has them): Z electrons in the lowest positive-energy Kramers
    pairs, an odd electron shared by the last pair.  Hermitian, positive and
    normalised, unlike a random test density; centres with the same charge
    and basis share one solve."""
    import importlib

    import torch
    scf = importlib.import_module(pkg.__name__ + ".scf")
    xi_all = pkg.systems.nuclear_exponents(system)
    cache = {}
    blocks = []
    for c in range(system.ncnt):
        z = float(system.charges[c])
        xi = None if xi_all is None else [float(xi_all[c])]
        key = (z, xi and xi[0], tuple(tuple(float(x) for x in e) for e in system.shells[c]))
        if key not in cache:
            atom = pkg.systems.System("atom", np.zeros((1, 3)), [system.shells[c]],
                                      shape="ATOMIC", charges=[z])
            eng = pkg.Engine(atom)
            try:
                n = eng.ndim
                mats = [torch.zeros((n, n), dtype=torch.complex128, device="cuda")
                        for _ in range(3)]
                eng.one_electron_device(*[m.data_ptr() for m in mats], xi=xi)
                torch.cuda.synchronize()
                # the device arrays hold the column-major matrices, i.e. M^T
                ovlp, hnuc, hkin = (m.T.cpu().contiguous() for m in mats)
                nskp = eng.nskp
            finally:
                eng.close()
            _, coef, _ = scf.solve_fock(hnuc + hkin, scf.orthogonaliser(ovlp))
            nel = int(round(z))
            full, odd = nel // 2, nel % 2
            occ = coef[:, nskp:nskp + 2 * (full + odd)]
            w = torch.ones(occ.shape[1], dtype=torch.float64)
            if odd:
                w[-2:] = 0.5
            cache[key] = (occ.conj() * w) @ occ.T
        blocks.append(cache[key])
    return scf.superposed_density(system, blocks).numpy()


def host_cores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="tlf", choices=["tlf", "water", "ne", "ring"])
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--stage", type=int, default=3, choices=[1, 2, 3],
                    help="integral-inclusion stage of every leg (ILEV; 3 = full Dirac-Coulomb)")
    ap.add_argument("--no-breit", action="store_true",
                    help="Dirac-Coulomb build only, also for the heavy diatomic")
    ap.add_argument("--schwarz", nargs="?", const="reference", default=None,
                    choices=["reference", "rigorous"],
                    help="SCHWARZ masks at SENS 1e-12 on every leg: the reference's (GDSC "
                         "bounds) or rigorous (Q(a,b) Q(c,d) on every group)")
    ap.add_argument("--density", default="atoms", choices=["random", "atoms"],
                    help="superposed bare-nucleus atoms, or a random Hermitian test density")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch every build from the host instead of replaying captured HIP "
                         "graphs")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-stride", type=int, default=0,
                    help="oracle sample: every S-th primitive pair (0: by workload)")
    ap.add_argument("--split", type=float, default=0.1,
                    help="Schwarz split threshold of the mixed-precision leg (0: skip the leg)")
    ap.add_argument("--serve", type=int, default=512,
                    help="requests of the served-SCF leg (0: skip the leg)")
    ap.add_argument("--serve-rate", type=float, default=3.0,
                    help="Poisson arrivals per second of the served-SCF leg (near capacity)")
    ap.add_argument("--serve-sustained-rate", type=float, default=1.2,
                    help="arrival rate of a second served-SCF leg below capacity (0: skip)")
    ap.add_argument("--serve-sustained", type=int, default=32,
                    help="requests of the sustained leg")
    ap.add_argument("--serve-slots", type=int, default=8)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first fill and scale MIB MiB on the device (kernels of known traffic "
                         "that calibrate FETCH_SIZE/WRITE_SIZE counters in a profiled run)")
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
    system = S.config_system(args.workload, **kw)
    breit = system.hmlt == "DHFB" and not args.no_breit
    if system.hmlt == "DHFB" and not breit:
        import dataclasses
        system = dataclasses.replace(system, hmlt="DHFR")
    schwarz = {None: False, "reference": True, "rigorous": "rigorous"}[args.schwarz]
    dev = torch.device("cuda:0")
    t_setup = time.perf_counter()
    graph = not args.no_graph
    eng = pkg.Engine(system, schwarz=schwarz, graph=graph)
    n = eng.ndim
    flops_c, quartets_c, items_c = eng.plan_info(args.stage)
    flops_b = quartets_b = 0.0
    if breit:
        flops_b, quartets_b, _ = eng.breit_plan_info(args.stage)
    print(f"bench: {system.name} NDIM {n}, plans built in {time.perf_counter() - t_setup:.1f} s",
          file=sys.stderr, flush=True)
    if args.density == "atoms":
        dens = atoms_density(pkg, system)
    else:
        dens = S.random_density(n, seed=11, real=(system.shape != "NOSYMM"))

    if args.copy_calibration > 0:
        nbytes = args.copy_calibration << 20
        src = torch.ones(nbytes // 8, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        torch.mul(src, 2.0, out=dst)
        torch.cuda.synchronize()
        print(f"copy calibration: fill and scale of {nbytes} bytes", file=sys.stderr, flush=True)
        del src, dst

    # column-major matrices: a row-major torch tensor of D^T has D's layout
    d_t = torch.from_numpy(np.ascontiguousarray(dens.T)).to(dev)

    def buffers():
        return [torch.zeros((n, n), dtype=torch.complex128, device=dev) for _ in range(4)]

    def timed_leg(e, label):
        g_t, x_t, b_t, bx_t = buffers()
        stream = torch.cuda.ExternalStream(e.stream(), device=dev)

        def step():
            e.clmfast_device(d_t.data_ptr(), g_t.data_ptr(), x_t.data_ptr(), args.stage)
            if breit:
                e.breit_device(d_t.data_ptr(), b_t.data_ptr(), bx_t.data_ptr(), args.stage)

        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        ev1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        ms = ev0.elapsed_time(ev1) / args.steps
        # each quartet kernel's own duration (events bracket only it in a
        # build launched from the host)
        e.set_launch(graph=False)
        e.clmfast_device(d_t.data_ptr(), g_t.data_ptr(), x_t.data_ptr(), args.stage)
        kc = e.stats()[1]
        kb = 0.0
        if breit:
            e.breit_device(d_t.data_ptr(), b_t.data_ptr(), bx_t.data_ptr(), args.stage)
            kb = e.stats()[1]
        torch.cuda.synchronize()
        e.set_launch(graph=graph)
        print(f"bench: {label} {ms:.1f} ms/step (coulomb kernel {kc:.1f} ms, breit kernel "
              f"{kb:.1f} ms)", file=sys.stderr, flush=True)
        mats = [m.cpu().numpy().T for m in (g_t, x_t, b_t, bx_t)]
        return ms, wall, kc, kb, mats

    ms64, wall64, kc64, kb64, (g, x, b, bx) = timed_leg(eng, "fp64")

    def e2(gm, xm, bm, bxm):
        v = np.sum(dens * (gm - xm))
        if breit:
            v = v + np.sum(dens * (bm - bxm))
        return 0.5 * float(np.real(v))

    e64 = e2(g, x, b, bx)
    launch = {"hip_graph": graph, "settings": eng.launch,
              "tuned": eng.tuning is not None}
    # the Breit build's quartet kernel: 2 matrix-core, 1 vector-ALU
    bkern = eng.breit_launch_info(args.stage)[0] if breit else 0
    if breit:
        launch["breit_kernel"] = "breit_tc_kernel" if bkern == 2 else "breit_kernel"
    eng.close()

    mixed = None
    if args.split > 0:
        eng_m = pkg.Engine(system, precision="mixed", split_threshold=args.split,
                           schwarz=schwarz, graph=graph)
        ms_m, _, kc_m, kb_m, (gm, xm, bm, bxm) = timed_leg(eng_m, "mixed")
        eng_m.close()
        dE = abs(e2(gm, xm, bm, bxm) - e64)
        mixed = {"value": 1000.0 / ms_m, "ms_per_step": ms_m, "split_threshold": args.split,
                 "abs_dE2e_vs_fp64_Eh": dE,
                 "max_abs_dG_vs_fp64": float(max(np.abs(gm - g).max(), np.abs(xm - x).max(),
                                                 np.abs(bxm - bx).max())),
                 "speedup_vs_fp64": ms64 / ms_m,
                 "coulomb_kernel_ms": kc_m, "breit_kernel_ms": kb_m,
                 "precision": "Coulomb quartets in FP32 below the Schwarz split "
                              "(Boys function, accumulation and digestion in FP64); Breit FP64"}

    # roofline of the dominant kernel of the FP64 step
    if breit and kb64 > kc64:
        kname = "breit_tc_kernel" if bkern == 2 else "breit_kernel"
        kms, kflops = kb64, flops_b
    else:
        kname, kms, kflops = "fock_kernel", kc64, flops_c
    bound = "mfma" if kname == "breit_tc_kernel" else "valu"
    achieved = kflops / (kms * 1e-3) / 1e12
    traffic, tsrc = profiled_traffic(kname, args.workload)

    cpu = None
    parity = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        cores = host_cores()
        stride = args.cpu_stride or {"tlf": 64, "water": 1, "ne": 1, "ring": 64}[args.workload]
        orc = pyoracle.Oracle(system)

        def cpu_step(threads, stride_):
            t0 = time.perf_counter()
            fr, go, xo = orc.clmfast_sample(dens, args.stage, nthreads=threads, stride=stride_)
            tc = time.perf_counter() - t0
            tb, frb, bo, bxo = 0.0, 1.0, None, None
            if breit:
                t0 = time.perf_counter()
                frb, bo, bxo = orc.brtfast_sample(dens, args.stage, nthreads=threads,
                                                  stride=stride_)
                tb = time.perf_counter() - t0
            # seconds of a whole step, scaled from the sample's share of calls
            return tc / fr + (tb / frb if breit else 0.0), (tc, fr, tb, frb), (go, xo, bo, bxo)

        s_all, (tc, fr, tb, frb), outs = cpu_step(cores, stride)
        s_one, (tc1, fr1, tb1, frb1), _ = cpu_step(1, stride * 8)
        what = "Dirac-Coulomb + Breit" if breit else "Dirac-Coulomb"
        cpu = {"value": 1.0 / s_all, "unit": "Fock builds/s", "cores": cores, "kind": "port",
               "sample": f"oracle/bdf_oracle.c, stage-{args.stage} {what} build of the "
                         f"{system.name} workload (NDIM {n}), same density; every {stride}th "
                         f"primitive pair ij of every class ({fr:.4f} of the ERIFAST calls"
                         + (f", {frb:.4f} of the BIIFAST calls" if breit else "") +
                         f"), {tc + tb:.1f} s on {cores} threads, scaled to one build",
               "one_thread": {"value": 1.0 / s_one, "cores": 1,
                              "sample": f"every {stride * 8}th primitive pair ({fr1:.4f} of the "
                                        f"calls), {tc1 + tb1:.1f} s"},
               "host": cpu_model(), "nproc": cores}
        if stride == 1:
            go, xo, _, _ = outs
            de = 0.5 * np.real(np.sum(dens * ((g - x) - (go - xo))))
            parity = {"max_abs_dGDIR": float(np.abs(g - go).max()),
                      "max_abs_dGXCH": float(np.abs(x - xo).max()),
                      "abs_dE2e_Eh": float(abs(de))}

    serving = None
    serving_sustained = None
    if args.serve > 0:
        import importlib
        sv = importlib.import_module(pkg.__name__ + ".serve")

        def serve_leg(count, rate):
            reqs = sv.request_mix(count, seed=args.seed, rate=rate)
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

        serving = serve_leg(args.serve, args.serve_rate)
        if args.serve_sustained_rate > 0 and args.serve_sustained > 0:
            serving_sustained = serve_leg(args.serve_sustained, args.serve_sustained_rate)

    # the mixed leg is the headline where it keeps the energy and is faster
    headline = mixed if (mixed is not None and mixed["abs_dE2e_vs_fp64_Eh"] < 1e-6
                         and mixed["ms_per_step"] < ms64) else None
    value = headline["value"] if headline else 1000.0 / ms64
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "Fock builds/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": headline["ms_per_step"] if headline else ms64,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32 (Schwarz split)" if headline else "f64",
        "data": ("superposed bare-nucleus atomic densities" if args.density == "atoms"
                 else "synthetic Hermitian density") +
                "; even-tempered basis and seeded geometry of the workload",
        "config": {"workload": system.name,
                   "build": ("stage-%d Dirac-Coulomb + Breit" if breit else
                             "stage-%d Dirac-Coulomb") % args.stage,
                   "ndim": n, "stage": args.stage, "schwarz": args.schwarz,
                   "density": args.density,
                   "precision": (f"mixed, Coulomb FP32 below split {args.split}" if headline
                                 else "fp64"),
                   "quartets_per_build": quartets_c + quartets_b, "work_items": items_c,
                   "host_wall_s_fp64": wall64, "launch": launch},
        "fp64": {"value": 1000.0 / ms64, "ms_per_step": ms64, "coulomb_kernel_ms": kc64,
                 "breit_kernel_ms": kb64, "two_electron_energy_Eh": e64,
                 "coulomb_tflops": flops_c / (kc64 * 1e-3) / 1e12,
                 "breit_tflops": (flops_b / (kb64 * 1e-3) / 1e12) if breit else None},
        "roofline": {"bound": bound, "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                     "traffic": traffic, "traffic_source": tsrc,
                     "kernel": f"{kname} (FP64 step)", "kernel_ms": kms,
                     "flops_per_launch": kflops,
                     "note": ("FP64 matrix-core kernel (16x16x4 FP64 MFMA); achieved counts "
                              "BIIFAST's statements (the plan's flops), of which the kernel "
                              "executes fewer: its gauge tables are shared by the rows of a "
                              "tile" if bound == "mfma" else
                              "FP64 vector-ALU kernel (no MFMA); the MI355X FP64 vector peak "
                              "equals its FP64 matrix peak")},
        "cpu_baseline": cpu,
        "parity": parity,
        "mixed": mixed,
        "serving": serving,
        "serving_sustained": serving_sustained,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
