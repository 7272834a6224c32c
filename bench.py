#!/usr/bin/env python3
"""SD-step benchmark on one MI355X.

Workload (default "h2o-32"): 32 water molecules (96 atoms) in an 18.6 bohr
simple-cubic cell, 96^3 FFT grid, 128 doubly occupied orbitals, O s+p / H s
Kleinman-Bylander projectors (192), LDA, dt = 5.8, fmass = 400000, energy
evaluated every inner = 10 steps, FP64 throughout.  Positions: seeded 4x4x2
lattice with +-0.3 bohr jitter and random orientations; orbitals: seeded
complex Gaussian coefficients on the masked G vectors, orthonormalised.

A "step" is one electronic SD step (ccpsdv3/cpsd.c:169-491) of all orbitals;
the timed region runs exactly --steps of them, resident in HBM, in calls of
cpsd_dev_run(inner <= 10), each ending with the energy tail, with no
instrumentation.  Per-phase GPU times come from HIP events recorded on the
library's stream around each phase during a second pass of min(--steps, 100)
steps right after it (the events cost ~2% of a step, so they stay out of
`value`).  A run from the same start measures the time until
|E(k) - E(k-1)| < 1e-7 Ha between energy evaluations (10 steps apart).

The "mixed" object repeats the steps/s, phase and time-to-converge legs in
the FP32-transform mode (cpsd_dev_set_precision(h, 1): the orbital FFTs and
KB products of 9 of every 10 steps on FP32 values, everything else and every
energy evaluation FP64) and reports its converged energy's error against the
FP64 run's, per atom; "mixed.bf16_projection" does the same for mode 2 (the
KB projection with BF16 inputs).  `value` stays the FP64 rate.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "simple-cpsd_amd"))

import numpy as np  # noqa: E402

WORKLOADS = {
    "h2o-1": dict(nmol=1, L=20.0, nfft=32, ne=4),
    "h2o-32": dict(nmol=32, L=18.6, nfft=96, ne=128),
    "h2o-64": dict(nmol=64, L=23.5, nfft=128, ne=256),
}
HBM_PEAK_GBS = 8000.0
FP64_MFMA_PEAK_TFLOPS = 78.6


def measured_traffic(workload):
    """HBM bytes per launch of each phase's kernel, from the committed
    rocprofv3 FETCH_SIZE + WRITE_SIZE summary of this workload (lines
    'phase <name> <bytes>'), or {} when there is none."""
    path = os.path.join(ROOT, "profiles", "sd_step_%s_hbm_traffic.md" % workload.replace("-", ""))
    out = {}
    if os.path.exists(path):
        for line in open(path):
            t = line.split()
            if len(t) == 3 and t[0] == "phase":
                out[t[1]] = float(t[2])
    return out


def algorithmic(phase, w, ng, nproj, ndnblk):
    """(kind, amount per SD step) of the algorithmic work of one phase:
    kind 'flop' for MFMA contractions (flops over the masked G vectors,
    complex coefficients as 2 reals), 'byte' for FFT / elementwise passes
    (compulsory traffic of the pass: its inputs read once, outputs written
    once)."""
    n, ne = w["nfft"], w["ne"]
    nh = n // 2 + 1
    plane = n * n * nh * 16            # one orbital's z-transformed array
    packed = ng * 16                   # one orbital's masked coefficients
    n3 = n ** 3 * 8
    if phase == "kb_projection":
        return "flop", 2.0 * nproj * ne * 2 * ng
    if phase == "kb_apply_sd":
        return "flop", 2.0 * nproj * ne * 2 * ng
    if phase == "gram":
        if ne % 128 == 0:
            # from the second step of a run <c1|c1> is carried over and the
            # Gram forms <c2|c1> and the upper triangle of <c2|c2>
            return "flop", 2.0 * (ne * ne + ne * (ne + 1) / 2) * 2 * ng
        return "flop", 2.0 * (ne * ne + ne * (ne + 1)) * 2 * ng
    if phase == "constraint":
        return "flop", 2.0 * ne * ne * 2 * ng
    if phase == "zpass_g2r":
        return "byte", ne * (packed + plane)
    if phase == "plane_density":
        return "byte", ne * plane + ndnblk * n3
    if phase == "plane_vpsi":
        return "byte", ne * 2 * plane + n3
    if phase == "zpass_r2g_hpsi":
        return "byte", ne * (plane + 2 * packed)
    return None, None


def run(args):
    import cpsd_amd
    w = WORKLOADS[args.workload]
    t0 = time.perf_counter()
    cell = cpsd_amd.water_cell(w["nmol"], w["L"], w["nfft"], seed=args.seed)
    psi = cell.random_orbitals(w["ne"], seed=args.seed + 1, decay=0.0)
    t_setup = time.perf_counter() - t0

    dev = cpsd_amd.DeviceSD(cell, w["ne"])
    dev.set_graphs(args.graphs)
    dev.set_orbitals(psi)
    sizes = dev.sizes()
    inner = 10

    def steps(k):
        E = de = dc = None
        while k > 0:
            m = min(inner, k)
            E, de, dc = dev.run(m)
            k -= m
        return E, de, dc

    steps(args.warmup)
    # the timed steps run without instrumentation; the per-phase times come
    # from a second, profiled pass of the same steps (HIP events on the
    # library's stream between the phases)
    t0 = time.perf_counter()
    E, de, dc = steps(args.steps)
    elapsed = time.perf_counter() - t0
    lag_it, lag_ad = dev.lagrange_info()
    prof_steps = min(args.steps, 100)
    dev.profile(True)
    t0 = time.perf_counter()
    steps(prof_steps)
    elapsed_prof = time.perf_counter() - t0
    prof = dev.profile_read()
    dev.profile(False)

    # the run epilogue once (orthonormality check, eigenpairs of hml, c1 = c2 V;
    # outside the timed steps)
    t0 = time.perf_counter()
    ep = dev.epilogue()
    epilogue = dict(seconds=round(time.perf_counter() - t0, 4),
                    ortho_max_dev=float(np.max(np.abs(ep["ortho"] - np.eye(w["ne"])))),
                    eig_max=float(ep["eig"][0]), eig_min=float(ep["eig"][-1]))

    # time to |dE| < 1e-7 Ha between energy evaluations, from the same start
    def converge(mode):
        dev.set_precision(mode)
        dev.set_orbitals(psi)
        dev.E[:] = 0.0
        t0 = time.perf_counter()
        done, nsteps, last = False, 0, None
        Eprev = None
        while nsteps < args.converge_steps:
            Ec, dec, dcc = dev.run(inner)
            nsteps += inner
            if Eprev is not None and abs(Ec[0] - Eprev) < 1e-7:
                done = True
                break
            if nsteps % 1000 == 0 and Eprev is not None:
                print("converge(%s): %d steps, E %.10f, dE %.3e" %
                      (("fp64", "mixed", "mixed-bf16")[mode], nsteps, Ec[0], Ec[0] - Eprev),
                      file=sys.stderr, flush=True)
            last = None if Eprev is None else Ec[0] - Eprev
            Eprev = Ec[0]
        tconv = time.perf_counter() - t0
        dev.set_precision(0)
        return dict(converged=done, steps=nsteps, seconds=round(tconv, 4),
                    E=float(Ec[0]), last_dE=None if last is None else float(last))

    conv = converge(0) if args.converge_steps > 0 else None

    # the FP32-transform mode: steps/s, phases, time to converge and the
    # converged energy's error against FP64
    mixed = None
    if args.mixed:
        dev.set_orbitals(psi)
        dev.set_precision(1)
        steps(args.warmup)
        t0 = time.perf_counter()
        Em, _, _ = steps(args.steps)
        el_m = time.perf_counter() - t0
        dev.profile(True)
        steps(prof_steps)
        prof_m = dev.profile_read()
        dev.profile(False)
        dev.set_precision(0)
        mixed = dict(value=round(args.steps / el_m, 3), unit="SD steps/s",
                     ms_per_step=round(el_m * 1000.0 / args.steps, 4),
                     speedup_vs_fp64=round(elapsed / el_m, 4),
                     fp32_steps_per_call=inner - 1, E_last=float(Em[0]),
                     # both legs ran warmup + steps steps from the same start
                     abs_E_diff_vs_fp64_per_atom_at_E_last=abs(float(Em[0]) - float(E[0])) / (3 * w["nmol"]),
                     phases={k: round(t / prof_steps, 4) for k, (t, c) in prof_m.items() if c > 0})
        if args.converge_steps > 0:
            cm = converge(1)
            if conv is not None and conv["converged"] and cm["converged"]:
                cm["abs_E_error_vs_fp64_per_atom"] = abs(cm["E"] - conv["E"]) / (3 * w["nmol"])
                cm["speedup_vs_fp64"] = round(conv["seconds"] / cm["seconds"], 4)
            mixed["time_to_converge"] = cm
        # mode 2: the same with a BF16-input KB projection
        dev.set_orbitals(psi)
        dev.set_precision(2)
        steps(args.warmup)
        t0 = time.perf_counter()
        Eb, _, _ = steps(args.steps)
        el_b = time.perf_counter() - t0
        dev.set_precision(0)
        mixed["bf16_projection"] = dict(
            value=round(args.steps / el_b, 3), unit="SD steps/s", ms_per_step=round(el_b * 1000.0 / args.steps, 4),
            speedup_vs_fp64=round(elapsed / el_b, 4),
            abs_E_diff_vs_fp64_per_atom_at_E_last=abs(float(Eb[0]) - float(E[0])) / (3 * w["nmol"]))
        if args.converge_steps > 0:
            cb = converge(2)
            if conv is not None and conv["converged"] and cb["converged"]:
                cb["abs_E_error_vs_fp64_per_atom"] = abs(cb["E"] - conv["E"]) / (3 * w["nmol"])
                cb["speedup_vs_fp64"] = round(conv["seconds"] / cb["seconds"], 4)
            mixed["bf16_projection"]["time_to_converge"] = cb
    dev.close()

    ms = elapsed * 1000.0 / args.steps
    value = args.steps / elapsed
    # per-phase GPU time per SD step and the roofline of the dominant phase
    ndnblk = min(8, w["ne"])
    phases = {}
    for name, (tms, calls) in prof.items():
        if calls == 0:
            continue
        per_step = tms / prof_steps if name != "energy_tail" else tms / max(1, calls) / inner
        phases[name] = dict(ms_total=round(tms, 3), launches=calls,
                            ms_per_step=round(per_step, 4),
                            launches_per_step=round(calls / prof_steps, 3))
    roofline = None
    traffic = measured_traffic(args.workload)
    for name in sorted(phases, key=lambda n: -phases[n]["ms_per_step"]):
        kind, amount = algorithmic(name, w, sizes["ng"], sizes["nproj"], ndnblk)
        if kind is None:
            continue
        sec = phases[name]["ms_per_step"] / 1000.0
        per_launch = amount / max(1e-9, phases[name]["launches_per_step"])
        if kind == "flop":
            ach = amount / sec / 1e12
            roofline = dict(bound="mfma", kernel=name, achieved=round(ach, 3),
                            peak=FP64_MFMA_PEAK_TFLOPS, unit="TFLOP/s",
                            frac=round(ach / FP64_MFMA_PEAK_TFLOPS, 4), traffic=traffic.get(name),
                            algorithmic_per_launch=per_launch)
        else:
            ach = amount / sec / 1e9
            roofline = dict(bound="hbm", kernel=name, achieved=round(ach, 1), peak=HBM_PEAK_GBS,
                            unit="GB/s", frac=round(ach / HBM_PEAK_GBS, 4), traffic=traffic.get(name),
                            algorithmic_per_launch=per_launch)
        break
    for name in phases:
        kind, amount = algorithmic(name, w, sizes["ng"], sizes["nproj"], ndnblk)
        if kind is not None:
            sec = phases[name]["ms_per_step"] / 1000.0
            phases[name]["achieved"] = round(amount / sec / (1e12 if kind == "flop" else 1e9), 3)
            phases[name]["unit"] = "TFLOP/s" if kind == "flop" else "GB/s"

    cpu = None
    if args.cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib
        t0 = time.perf_counter()
        est = oracle_lib.step_time_sampled(w["nfft"], w["ne"], cell.nion, cell.nproj,
                                           lag_iters=lag_it, inner=inner, sample=args.cpu_sample)
        tcpu = time.perf_counter() - t0
        cpu = dict(value=round(1.0 / est["total"], 6), unit="SD steps/s", cores=1, kind="port",
                   sample=("serial FP64 restatement (oracle), one SD step of this workload "
                           "estimated from %d timed work items per component at full array "
                           "size (3D FFT pairs, orbital passes, 8x KB pairs, overlaps, "
                           "constraint columns), one full Lagrange iteration x %d, LDA, "
                           "energy tail / %d; %.1f s of CPU work" %
                           (args.cpu_sample, lag_it, inner, tcpu)),
                   seconds_per_step=round(est["total"], 4),
                   parts={k: round(v, 5) for k, v in est.items()},
                   nproc=os.cpu_count())
        if conv is not None and conv["converged"]:
            # the CPU would need the same number of steps
            cpu["time_to_converge_s_extrapolated"] = round(conv["steps"] * est["total"], 1)

    out = {
        "metric": "SD steps/s and time-to-|dE|<1e-7 Ha: mixed+learned vs FP64 GPU, 1 MI355X",
        "value": round(value, 3),
        "unit": "SD steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded water positions and Gaussian orbitals)",
        "config": {"workload": args.workload, "waters": w["nmol"], "atoms": 3 * w["nmol"],
                   "cell_bohr": w["L"], "fft_grid": w["nfft"], "orbitals": w["ne"],
                   "projectors": sizes["nproj"], "masked_G": sizes["ng"], "inner": inner,
                   "precision": "fp64", "start": "gaussian", "graphs": bool(args.graphs),
                   "setup_s": round(t_setup, 2)},
        "time_to_converge": conv,
        "mixed": mixed,
        "profiled_pass": {"steps": prof_steps, "ms_per_step": round(elapsed_prof * 1000.0 / prof_steps, 4)},
        "lagrange_iterations_last_step": lag_it,
        "epilogue": epilogue,
        "E_last": float(E[0]),
        "phases": phases,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    return out


def batch_cells(seed):
    """64 seeded water cells of 1-8 molecules: cell edge 12-22 bohr growing with
    the molecule count, grid 24^3-64^3 by edge, 4 orbitals per molecule"""
    rng = np.random.default_rng(seed)
    cells = []
    for _ in range(64):
        nmol = int(rng.integers(1, 9))
        L = 12.0 + 10.0 * (nmol - 1) / 7.0 + float(rng.uniform(-0.4, 0.4))
        n = 24 if L < 14 else 32 if L < 17 else 48 if L < 20 else 64
        cells.append((nmol, L, n))
    return cells


def run_batch(args):
    """--workload batch-64: the 64 cells advance 10-step calls one cell after
    the other (each call waited for), then with every cell's call started
    before any is finished (one stream per cell); cell-steps per second"""
    import cpsd_amd
    t0 = time.perf_counter()
    specs = batch_cells(args.seed)
    devs = []
    for i, (nmol, L, n) in enumerate(specs):
        cell = cpsd_amd.water_cell(nmol, L, n, seed=args.seed + i)
        d = cpsd_amd.DeviceSD(cell, 4 * nmol)
        d.set_graphs(args.graphs)
        d.set_orbitals(cell.random_orbitals(4 * nmol, seed=args.seed + 100 + i, decay=0.05))
        devs.append(d)
    t_setup = time.perf_counter() - t0
    inner = 10
    calls = max(1, args.steps // inner)
    for _ in range(max(1, args.warmup // inner)):
        for d in devs:
            d.run(inner)
        cpsd_amd.run_batch(devs, inner)
    t0 = time.perf_counter()
    for _ in range(calls):
        for d in devs:
            d.run(inner)
    seq = time.perf_counter() - t0
    t0 = time.perf_counter()
    for _ in range(calls):
        cpsd_amd.run_batch(devs, inner)
    bat = time.perf_counter() - t0
    for d in devs:
        d.close()
    cell_steps = len(devs) * calls * inner
    return {
        "metric": "SD cell-steps/s, 64 independent water cells, 1 MI355X",
        "value": round(cell_steps / bat, 1),
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": calls * inner,
        "warmup": args.warmup,
        "ms_per_step": round(bat * 1000.0 / (calls * inner), 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded water cells and Gaussian orbitals)",
        "config": {"workload": "batch-64", "cells": len(devs),
                   "waters": [c[0] for c in specs], "fft_grids": sorted(set(c[2] for c in specs)),
                   "orbitals": "4 per water", "inner": inner, "graphs": bool(args.graphs),
                   "setup_s": round(t_setup, 2)},
        "sequential": {"value": round(cell_steps / seq, 1), "unit": "cell-steps/s",
                       "ms_per_step_all_cells": round(seq * 1000.0 / (calls * inner), 4)},
        "speedup_vs_sequential": round(seq / bat, 4),
    }


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=500)
    p.add_argument("--warmup", type=int, default=20)
    p.add_argument("--workload", default="h2o-32", choices=sorted(WORKLOADS) + ["batch-64"])
    p.add_argument("--seed", type=int, default=0)
    p.add_argument("--converge-steps", type=int, default=10000)
    p.add_argument("--no-cpu-baseline", dest="cpu_baseline", action="store_false")
    p.add_argument("--no-mixed", dest="mixed", action="store_false")
    p.add_argument("--graphs", action="store_true",
                   help="replay each 10-step call from a captured HIP graph")
    p.add_argument("--cpu-sample", type=int, default=16)
    args = p.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the SD step of one cell runs on one MI355X and there "
              "is no multi-GPU path; run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    # the library prints the reference's own notices (e.g. "ierr=10 ...", as
    # ccpsdv3/cpsd.c:477 does) on C stdout; keep stdout for the one JSON line
    sys.stdout.flush()
    saved = os.dup(1)
    os.dup2(2, 1)
    try:
        out = run_batch(args) if args.workload == "batch-64" else run(args)
    finally:
        sys.stdout.flush()
        import ctypes
        ctypes.CDLL(None).fflush(None)   # C stdio buffers go to the redirected fd
        os.dup2(saved, 1)
        os.close(saved)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
