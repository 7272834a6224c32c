#!/usr/bin/env python3
"""SD-step benchmark on one MI355X.

Workload (default "h2o-64", the cell the metric's target is quoted on): 64
water molecules (192 atoms) in a 23.5 bohr simple-cubic cell, 128^3 FFT
grid, 256 doubly occupied orbitals, O s+p / H s Kleinman-Bylander
projectors (384), LDA, dt = 5.8, fmass = 400000, energy evaluated every
inner = 10 steps.  Positions: seeded 4x4x4 lattice with +-0.3 bohr jitter
and random orientations; orbitals: seeded complex Gaussian coefficients on
the masked G vectors, orthonormalised.  Synthetic inputs of the configured
shape (no downloaded data).

A "step" is one electronic SD step (ccpsdv3/cpsd.c:169-491) of all orbitals.
`value` is the FP64 rate: exactly --steps steps in calls of cpsd_dev_run
(inner <= 10, each ending with the energy tail), orbitals resident in HBM, no
instrumentation.  Per-phase GPU times come from HIP events on the library's
stream during a second pass of min(--steps, 100) steps.

The metric's comparison, time to |E(k) - E(k-1)| < 1e-7 Ha between energy
evaluations from the same start:
  time_to_converge.fp64          the reference's FP64 step (plain SD);
  time_to_converge.mixed_learned precision mode 3 (in 9 of 10 steps FP32
                                 orbital FFTs and every product in split
                                 BF16 MFMA: KB projection and apply, the
                                 preconditioned update, the overlaps of the
                                 step c2 - c1 with the carried FP64 <c1|c1>,
                                 the constraint; the Lagrange iteration,
                                 potentials and every energy FP64) with the
                                 learned per-G preconditioner
                                 (data/precond_mlp.json), its converged
                                 energy's error per atom against the FP64
                                 run's, and the speedup (mode 1, FP32
                                 products, where ne is not 128 or 256).
--all-legs adds mode 1 + learned (FP32 KB products and update, FP64 overlaps
and constraint), FP64 + learned, mode 1 alone and mode 2 (BF16 KB
projection).
The drop-in leg times two cpsd() calls on host arrays (the reference's
host-pointer contract, PCIe copies included); the second reuses the device
context.
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
# the reference's serial FP64 build timed on one core of the same host type
# at 96^3 x 128 orbitals; the 128^3 x 256-orbital figure scales it by the
# ratio of dense flops of the two cells
REFERENCE_CPU_S_PER_STEP = {"h2o-32": (113.0, "measured, 1 core Xeon 2.7 GHz"),
                            "h2o-64": (1080.0, "extrapolated from h2o-32 by the dense-flop ratio 9.5")}
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


def step_floor_ms(w, ng, nproj):
    """the per-step roofline floor of the plain FP64 step: FFT traffic
    (2ne+2) 2 n2 8 B and elementwise
    traffic 6 ne n2 8 B at 8 TB/s, contractions 8 nproj ne NG + 8 ne^2 NG +
    4 ne^2 NG flop at 78.6 TFLOP/s"""
    n, ne = w["nfft"], w["ne"]
    n2 = (n + 2) * n * n
    fft_b = (2 * ne + 2) * 2 * n2 * 8
    elem_b = 6 * ne * n2 * 8
    flop = 8.0 * nproj * ne * ng + 8.0 * ne * ne * ng + 4.0 * ne * ne * ng
    return dict(fft_ms=fft_b / HBM_PEAK_GBS / 1e6, elementwise_ms=elem_b / HBM_PEAK_GBS / 1e6,
                contractions_ms=flop / FP64_MFMA_PEAK_TFLOPS / 1e9,
                total_ms=fft_b / HBM_PEAK_GBS / 1e6 + elem_b / HBM_PEAK_GBS / 1e6 + flop / FP64_MFMA_PEAK_TFLOPS / 1e9)


def run(args):
    import cpsd_amd
    import learned
    w = WORKLOADS[args.workload]
    natoms = 3 * w["nmol"]
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

    # the run epilogue once (orthonormality check, eigenpairs of hml, c1 = c2 V)
    t0 = time.perf_counter()
    ep = dev.epilogue()
    epilogue = dict(seconds=round(time.perf_counter() - t0, 4),
                    ortho_max_dev=float(np.max(np.abs(ep["ortho"] - np.eye(w["ne"])))),
                    eig_max=float(ep["eig"][0]), eig_min=float(ep["eig"][-1]))

    model = learned.load_model() if args.converge_steps > 0 else None

    def configure(mode, learned_on):
        dev.set_precision(mode)
        dev.set_preconditioner(None)
        if learned_on:
            dev.set_learned_preconditioner(model["params"], model["alpha"], model["kmin"], model["kmax"])

    def converge(mode, learned_on, tag):
        configure(mode, learned_on)
        dev.set_orbitals(psi)
        dev.E[:] = 0.0
        t0 = time.perf_counter()
        done, nsteps, last, Eprev = False, 0, None, None
        while nsteps < args.converge_steps:
            Ec, dec, dcc = dev.run(inner)
            nsteps += inner
            if Eprev is not None and abs(Ec[0] - Eprev) < 1e-7:
                done = True
                break
            if nsteps % 1000 == 0 and Eprev is not None:
                print("converge(%s): %d steps, E %.10f, dE %.3e" % (tag, nsteps, Ec[0], Ec[0] - Eprev),
                      file=sys.stderr, flush=True)
            last = None if Eprev is None else Ec[0] - Eprev
            Eprev = Ec[0]
        tconv = time.perf_counter() - t0
        fb = dev.fallback_info()
        ortho = float(np.max(np.abs(dev.epilogue()["ortho"] - np.eye(w["ne"]))))
        configure(0, False)
        return dict(converged=done, steps=nsteps, seconds=round(tconv, 4), E=float(Ec[0]),
                    last_dE=None if last is None else float(last), fallbacks=fb["fallbacks"],
                    ortho_max_dev_at_end=ortho)

    def rate(mode, learned_on):
        """steps/s of the configured step over --steps steps (after warmup)"""
        configure(mode, learned_on)
        dev.set_orbitals(psi)
        steps(max(args.warmup, 2 * inner))   # the preconditioner starts with the second call
        t0 = time.perf_counter()
        steps(args.steps)
        el = time.perf_counter() - t0
        dev.profile(True)
        steps(prof_steps)
        ph = dev.profile_read()
        dev.profile(False)
        configure(0, False)
        return dict(value=round(args.steps / el, 3), unit="SD steps/s", ms_per_step=round(el * 1000.0 / args.steps, 4),
                    phases={k: round(t / prof_steps, 4) for k, (t, c) in ph.items() if c > 0})

    conv = {}
    if args.converge_steps > 0:
        conv["fp64"] = converge(0, False, "fp64")
        hmode = 3 if w["ne"] in (128, 256) else 1
        ml = converge(hmode, True, "mixed+learned")
        ml.update(rate(hmode, True))
        ml["precision_mode"] = hmode
        if conv["fp64"]["converged"] and ml["converged"]:
            ml["abs_E_error_vs_fp64_per_atom"] = abs(ml["E"] - conv["fp64"]["E"]) / natoms
            ml["speedup_vs_fp64"] = round(conv["fp64"]["seconds"] / ml["seconds"], 4)
        conv["mixed_learned"] = ml
        if args.all_legs:
            for key, mode, lon in (("mixed1_learned", 1, True), ("fp64_learned", 0, True), ("mixed", 1, False),
                                   ("mixed_bf16", 2, False)):
                r = converge(mode, lon, key)
                if conv["fp64"]["converged"] and r["converged"]:
                    r["abs_E_error_vs_fp64_per_atom"] = abs(r["E"] - conv["fp64"]["E"]) / natoms
                    r["speedup_vs_fp64"] = round(conv["fp64"]["seconds"] / r["seconds"], 4)
                conv[key] = r
    dev.close()

    # the drop-in cpsd() on host arrays (reference layouts, PCIe both ways)
    dropin = None
    if args.dropin:
        E0 = np.zeros(10)
        c2 = psi
        times = []
        for call in range(2):
            t0 = time.perf_counter()
            out = cpsd_amd.cpsd_dropin(cell, w["ne"], inner, c2, E0)
            times.append(time.perf_counter() - t0)
            c2, E0 = out["c2"], out["E"]
            del out
        dropin = dict(first_call_s=round(times[0], 3), second_call_s=round(times[1], 3),
                      ms_per_step_incl_pcie=round(times[1] * 1000.0 / inner, 3),
                      note="second call reuses the device context; each call uploads c2 and copies "
                           "c1, c2, cpsi, dn, hml, lmd back (%.2f GB)" % (
                               (4 * w["ne"] * cell.n2ft3d + cell.n2ft3d) * 8 / 1e9))

    ms = elapsed * 1000.0 / args.steps
    value = args.steps / elapsed
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
    floor = step_floor_ms(w, sizes["ng"], sizes["nproj"])
    step_roofline = dict(floor_ms=round(floor["total_ms"], 3), measured_ms=round(ms, 3),
                         frac=round(floor["total_ms"] / ms, 4),
                         parts_ms={k: round(v, 3) for k, v in floor.items() if k != "total_ms"})

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
        ref = REFERENCE_CPU_S_PER_STEP.get(args.workload)
        if ref is not None:
            cpu["reference_binary_s_per_step"] = ref[0]
            cpu["reference_binary_note"] = ref[1]
        fp = conv.get("fp64")
        if fp is not None and fp["converged"]:
            cpu["time_to_converge_s_extrapolated"] = round(fp["steps"] * est["total"], 1)

    headline = None
    if "mixed_learned" in conv:
        ml, fp = conv["mixed_learned"], conv["fp64"]
        headline = dict(
            fp64_time_to_converge_s=fp["seconds"], fp64_steps=fp["steps"],
            mixed_learned_time_to_converge_s=ml["seconds"], mixed_learned_steps=ml["steps"],
            speedup=ml.get("speedup_vs_fp64"),
            abs_E_error_per_atom=ml.get("abs_E_error_vs_fp64_per_atom"),
            target="speedup >= 3 with |E - E_FP64| <= 1e-6 Ha/atom",
            met=bool(ml.get("speedup_vs_fp64") is not None and ml["speedup_vs_fp64"] >= 3.0 and
                     ml.get("abs_E_error_vs_fp64_per_atom", 1.0) <= 1e-6))

    out = {
        "metric": "SD steps/s and time-to-|ΔE|<1e-7 Ha: mixed+learned vs FP64 GPU, 1 MI355X",
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
        "config": {"workload": args.workload, "waters": w["nmol"], "atoms": natoms,
                   "cell_bohr": w["L"], "fft_grid": w["nfft"], "orbitals": w["ne"],
                   "projectors": sizes["nproj"], "masked_G": sizes["ng"], "inner": inner,
                   "precision": "fp64", "start": "gaussian", "graphs": bool(args.graphs),
                   "setup_s": round(t_setup, 2)},
        "headline": headline,
        "time_to_converge": conv or None,
        "dropin": dropin,
        "profiled_pass": {"steps": prof_steps, "ms_per_step": round(elapsed_prof * 1000.0 / prof_steps, 4)},
        "lagrange_iterations_last_step": lag_it,
        "epilogue": epilogue,
        "E_last": float(E[0]),
        "phases": phases,
        "roofline": roofline,
        "step_roofline": step_roofline,
        "cpu_baseline": cpu,
    }
    return out


def run_batch(args):
    """--workload batch-64: 64 independent water cells of uneven size (grids
    24^3-64^3, 4-32 orbitals) advance 10-step calls in precision mode 1.
    Three ways are timed over the same calls: one cell after the other (each
    call waited for), every cell's call started before any is finished (one
    stream per cell), and the one-graph batch (cpsd_batch_run: every cell's
    call in one HIP graph, one launch per call), after each size class's
    launch plan was chosen by the autotuner.  value: cell-steps/s of the
    one-graph batch."""
    import cpsd_amd
    t0 = time.perf_counter()
    specs = cpsd_amd.water_batch_specs(args.seed)
    cells = []
    for i, (nmol, L, n) in enumerate(specs):
        cell = cpsd_amd.water_cell(nmol, L, n, seed=args.seed + i)
        cells.append((cell, 4 * nmol, cell.random_orbitals(4 * nmol, seed=args.seed + 100 + i, decay=0.05)))
    t_setup = time.perf_counter() - t0

    def make():
        devs = []
        for cell, ne, psi in cells:
            d = cpsd_amd.DeviceSD(cell, ne)
            d.set_precision(1)
            d.set_orbitals(psi)
            devs.append(d)
        return devs

    devs = make()
    t0 = time.perf_counter()
    plans = {}
    for d in devs:
        p = d.autotune(2)
        key = "%d^3/%d" % (d.cell.nfft, d.ne)
        if not p["from_table"]:
            plans[key] = p
    t_tune = time.perf_counter() - t0
    inner = 10
    calls = max(1, args.steps // inner)
    warm = max(1, args.warmup // inner)
    for _ in range(warm):
        for d in devs:
            d.run(inner)
    t0 = time.perf_counter()
    for _ in range(calls):
        for d in devs:
            d.run(inner)
    seq = time.perf_counter() - t0
    for _ in range(warm):
        cpsd_amd.run_batch(devs, inner)
    t0 = time.perf_counter()
    for _ in range(calls):
        cpsd_amd.run_batch(devs, inner)
    con = time.perf_counter() - t0
    b = cpsd_amd.Batch(devs)
    for _ in range(warm + 1):
        b.run(inner)
    t0 = time.perf_counter()
    for _ in range(calls):
        b.run(inner)
    one = time.perf_counter() - t0
    info = b.info()
    b.close()
    for d in devs:
        d.close()
    cell_steps = len(devs) * calls * inner
    return {
        "metric": "SD cell-steps/s, 64 independent water cells, 1 MI355X",
        "value": round(cell_steps / one, 1),
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": calls * inner,
        "warmup": args.warmup,
        "ms_per_step": round(one * 1000.0 / (calls * inner), 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32/f64",
        "data": "synthetic (seeded water cells and Gaussian orbitals)",
        "config": {"workload": "batch-64", "cells": len(devs),
                   "waters": [c[0] for c in specs], "fft_grids": sorted(set(c[2] for c in specs)),
                   "orbitals": "4 per water", "inner": inner, "precision": "mode 1 (FP32 FFT/KB, FP64 rest)",
                   "setup_s": round(t_setup, 2), "autotune_s": round(t_tune, 2)},
        "one_graph": {"value": round(cell_steps / one, 1), "graph_info": info},
        "concurrent_streams": {"value": round(cell_steps / con, 1)},
        "sequential": {"value": round(cell_steps / seq, 1)},
        "speedup_vs_sequential": round(seq / one, 4),
        "autotuned_plans": plans,
    }


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=50)
    p.add_argument("--warmup", type=int, default=20)
    p.add_argument("--workload", default="h2o-64", choices=sorted(WORKLOADS) + ["batch-64"])
    p.add_argument("--seed", type=int, default=0)
    p.add_argument("--converge-steps", type=int, default=10000)
    p.add_argument("--no-cpu-baseline", dest="cpu_baseline", action="store_false")
    p.add_argument("--all-legs", action="store_true",
                   help="also time FP64 + learned, mode 1 and mode 2 to convergence")
    p.add_argument("--no-dropin", dest="dropin", action="store_false")
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
