#!/usr/bin/env python3
"""Benchmark: bond-updates/s of the RK4 dynamic solve on one MI355X, FP32-compensated
against FP64, plus the static-solve wall time.

Default workload ("slit"): the 200 x 40 x 40 bar (320,000 nodes, dx = 1 mm,
delta = 3.015 dx, 73,728,544 family entries), node jitter U(-0.01, 0.01) dx from seed
0, E = 72 GPa, rho = 2440, x = 0 layer fixed, 1 MPa +x traction on the x = L layer,
RK4 at 0.8 x the stable dt, with one planar slit through the whole section at the bar
centre, normal at 30 degrees to x in the x-y plane (it cuts 0.86 % of the bonds).  One
step = 4 stages, each a full bond-force gather.  "bar" is the same bar without the slit;
the default run also measures it (plain_bar) in FP64 and FP32C.

The paths run the same K steps after W warm-up steps: FP64 (reference arithmetic),
FP32C (FP32 bond force, FP32 family sums) and FP32C2 (two-float family sums), all with
FP64 RK4 state.  `value` is the FP32C rate; the line carries the FP64 and FP32C2 rates
and each FP32 path's max relative displacement error against FP64 after the W + K steps.
Inputs are resident in HBM before the timed region (pdm_init builds them on the GPU);
the timed region is pdm_step_rk4: hipGraph replays of the four fused stage kernels per
step, in chunks of gcd(W, K) steps (at most 50), so the warm-up has captured every
graph the timed steps launch.

Static solve: the same bar with one planar edge crack from the y = min face to the
mid-plane, normal at 30 degrees to x (the through-section slit leaves the far half
unsupported, and the static problem singular); matrix-free PCG to 1e-10 relative
residual in FP64, FP32 matvecs with FP64 residual replacement, and the FP64 solve for
the error figure.  Wall seconds of pdm_solve_static.

value counts bond-updates as the reference performs them: 4 per unbroken family entry
per RK4 step (pdm_num_entries less the cut ones).  The gathers evaluate only the
evaluated_entries (pdm_num_live_entries): an entry whose CCI beta is 0 contributes
exactly 0 to every force and energy, and the body force still counts it.
evaluated_bond_updates_per_s gives the rate of the evaluated ones.

roofline: dominant kernel = the fused bond-force / RK4-stage kernel.  Its average
launch time comes from HIP events recorded on the library's stream around the timed
region (pdm_last_kernel_timing).  Algorithmic bytes per launch use the per-unit figures
of the bond-update unit over the entries the kernel evaluates: 8 B per evaluated entry
for FP32 coefficients (int32 neighbour + FP32 coefficient; 12 B with FP64
coefficients) plus 264 B per node (trial u, X, Vbuf: 72 B, and the fused stage update,
8 vectors of 24 B).  The library's own format is smaller (uint16 LDS slot + one
correction byte from which the FP32 kernels form the coefficient: 3 B per entry; uint16
slot + FP64 coefficient: 10 B); `format_bytes` gives that figure.  `traffic` is
FETCH_SIZE + WRITE_SIZE per launch from rocprofv3 counter runs on this workload, scaled
by a plain copy kernel's readings, as recorded under profiles/ for this workload (the
200 x 40 x 40 bar with and without the slit, the 400 x 400 x 50 plate; null for any
other), and `measured_gbs` is that traffic over the launch time.

cpu_baseline: the oracle (a serial C restatement of the reference's RK4 loop; the
reference itself does not build here, so the figure is not calibrated against it) on the
same bar, a few steps on 1 thread of this host (value, cores 1), then the next few steps
with its node loop on all the cores given to this process (all_cores; OMP_NUM_THREADS,
else the CPU count).
"""

import argparse
import json
import math
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
ENTRY_BYTES = {"fp32c": 8, "fp64": 12}        # int32 neighbour + coefficient
FORMAT_ENTRY_BYTES = {"fp32c": 3, "fp64": 10}  # uint16 slot + correction byte / FP64 coefficient (this library's format)
NODE_BYTES = 72 + 8 * 24                       # trial u, X, Vbuf + fused stage update (8 x 24 B)
DEFAULT_DIMS = (200, 40, 40)
TRAFFIC_FILES = {("slit", DEFAULT_DIMS): os.path.join(ROOT, "profiles", "k_stage_slitbar200x40x40_traffic.txt"),
                 ("bar", DEFAULT_DIMS): os.path.join(ROOT, "profiles", "k_stage_bar200x40x40_traffic.txt"),
                 ("plate", (400, 400, 50)): os.path.join(ROOT, "profiles", "k_stage_plate400x400x50_traffic.txt"),
                 ("batch", 512): os.path.join(ROOT, "profiles", "k_stage_batch512_traffic.txt")}
PRECISIONS = ("fp64", "fp32c", "fp32c2")


def workload_spec(args, steps, workload=None):
    import pdmtel_amd as pm
    from pdmtel_amd.deck import plate_spec

    workload = workload or args.workload
    if workload == "plate":
        return plate_spec(args.nx, args.ny, args.nz, jitter=0.01, seed=0, steps=steps)
    spec = pm.bar_spec(args.nx, args.ny, args.nz, jitter=0.01, seed=0, steps=steps)
    if workload == "slit":
        spec["slits"] = [pm.mid_slit(spec, 30.0)]
    return spec


def graph_chunk(args):
    """Steps per captured graph: all K timed steps in one graph up to 200 steps, else the largest
    divisor of K not above 200 (or --chunk when given and dividing K).  The graph of that length
    is captured before the warm-up (prepare), so the timed steps replay it with no capture or
    host sync between chunks."""
    k = args.steps
    if args.chunk > 0 and k % args.chunk == 0:
        return args.chunk
    if k <= 200:
        return max(1, k)
    return next(c for c in range(200, 0, -1) if k % c == 0)


def precision_code(name):
    import pdmtel_amd as pm

    return {"fp64": pm.PDM_FP64, "fp32c": pm.PDM_FP32C, "fp32c2": pm.PDM_FP32C2}[name]


def run_precision(args, spec, precision):
    import pdmtel_amd as pm

    text = pm.deck_text(spec)
    t0 = time.perf_counter()
    layout = {}
    if args.layout:  # "edge,waves,families": fixed tile layout instead of the autotuner's
        e, w, f = (int(v) for v in args.layout.split(","))
        layout = dict(tile_edge=e, waves_per_tile=w, families_per_wave=f)
    prob = pm.Problem(text, precision=precision, graph_chunk=graph_chunk(args), energies=1, **layout)
    prob.prepare(graph_chunk(args))
    t_init = time.perf_counter() - t0
    if args.warmup:
        prob.step(args.warmup)
    t0 = time.perf_counter()
    done = prob.step(args.steps)  # returns after the stream has drained
    t1 = time.perf_counter()
    assert done == args.steps, f"ran {done} of {args.steps} steps"
    ms_kernel, launches = prob.kernel_timing()
    U, _ = prob.state()
    ncut = int(prob.csr()[2].sum())
    res = dict(n=prob.n, E=prob.E, E_live=prob.E_live, E_unbroken=prob.E - ncut, seconds=t1 - t0,
               init_seconds=t_init, kernel_ms=ms_kernel, launches=launches, U=U, energy=prob.energy(),
               tuning=prob.tuning(), launch=prob.stage_launch())
    prob.close()
    return res


def run_static(args, names=("fp32c", "fp32c2", "fp64")):
    import numpy as np

    import pdmtel_amd as pm

    spec = workload_spec(args, 1, "bar" if args.workload in ("bar", "slit") else None)
    if args.workload in ("bar", "slit"):
        spec["slits"] = [pm.mid_slit(spec, 30.0, edge=True)]
    text = pm.deck_text(spec)
    out = {}
    for name, prec in (("fp32c", pm.PDM_FP32C), ("fp32c2", pm.PDM_FP32C2), ("fp64", pm.PDM_FP64)):
        if name not in names:
            continue
        prob = pm.Problem(text, precision=prec, verbose=1)  # progress lines on stderr
        t0 = time.perf_counter()
        st = prob.solve_static(rel_tol=1e-10, max_iterations=200000)
        wall = time.perf_counter() - t0
        U, _ = prob.state()
        out[name] = dict(wall=wall, iterations=st.iterations, refinements=st.refinements,
                         rel_residual=st.rel_residual, U=U)
        print(f"# static {name}: {wall:.3f} s, {st.iterations} iterations, {st.refinements} refinements, "
              f"residual {st.rel_residual:.2e}", file=sys.stderr, flush=True)
        prob.close()
    err = {k: float(np.abs(out[k]["U"] - out["fp64"]["U"]).max() / np.abs(out["fp64"]["U"]).max())
           for k in ("fp32c", "fp32c2") if k in out and "fp64" in out}
    return out, err


FGI_FLOPS_PER_POINT = 55  # FP64 operations of one Gauss-point bond force (sqrt and division as one)
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3


def run_fgi(args):
    """Fixed Gaussian integration (fgi_on 1, intg_point n): RK4 on the bar in FP64 and with FP32
    bond arithmetic (FP32C, FP32C2; FP64 Gauss-point displacements and state), one JSON line whose
    value is the FP32C rate (FP64 alone with --precision fp64)."""
    import numpy as np
    import pdmtel_amd as pm

    n1 = args.intg_point
    spec = pm.bar_spec(args.nx, args.ny, args.nz, jitter=0.01, seed=0, steps=args.warmup + args.steps + 10)
    spec["cci"] = None
    spec["fgi"] = (1, n1)
    names = ["fp64", "fp32c", "fp32c2"] if args.precision == "all" else [args.precision]
    res = {}
    for name in names:
        t0 = time.perf_counter()
        prob = pm.Problem(pm.deck_text(spec), precision=precision_code(name), graph_chunk=graph_chunk(args),
                          energies=1)
        prob.prepare(graph_chunk(args))
        t_init = time.perf_counter() - t0
        print(f"# fgi n={n1} {name}: init {t_init:.2f} s, {prob.n} nodes, {prob.E} entries", file=sys.stderr,
              flush=True)
        if args.warmup:
            prob.step(args.warmup)
        t0 = time.perf_counter()
        done = prob.step(args.steps)
        sec = time.perf_counter() - t0
        assert done == args.steps, f"ran {done} of {args.steps} steps"
        ms, launches = prob.kernel_timing()
        res[name] = dict(sec=sec, ms=ms, launches=launches, n=prob.n, E=prob.E, P=prob.live_gauss_pairs,
                         U=prob.state()[0], init=t_init)
        print(f"# fgi n={n1} {name}: {sec:.3f} s for {args.steps} steps, stage {ms:.4f} ms", file=sys.stderr,
              flush=True)
        prob.close()
    head = "fp32c" if "fp32c" in res else names[0]
    r = res[head]
    n, E, ng, P = r["n"], r["E"], n1 ** 3, r["P"]
    flops = P * FGI_FLOPS_PER_POINT  # the pairs the kernel evaluates (beta > 0)
    achieved = flops / (r["ms"] * 1e-3) / 1e12
    line = {
        "metric": f"bond-updates/s (RK4, fixed Gaussian integration {n1}x{n1}x{n1}, {head})",
        "value": 4.0 * E * args.steps / r["sec"],
        "unit": "bond-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["sec"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if head != "fp64" else "f64",
        "data": "synthetic",
        "config": {"workload": f"bar {args.nx}x{args.ny}x{args.nz}, {n} nodes, {E} family entries, delta=3.015dx, "
                               f"jitter 0.01dx seed 0, fgi_on 1 intg_point {n1}, RK4 {args.steps} steps",
                   "precision": head, "gauss_points_per_node": ng, "evaluated_pairs": P, "pairs": E * ng,
                   "init_seconds": r["init"]},
        "gauss_point_evaluations_per_s": 4.0 * E * ng * args.steps / r["sec"],
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS if head == "fp64" else
                     FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / (FP64_PEAK_TFLOPS if head == "fp64" else FP32_PEAK_TFLOPS), "traffic": None,
                     "kernel": "k_fgi_stage + k_fgi_gauss (one RK4 stage)", "kernel_avg_ms": r["ms"],
                     "launches": r["launches"], "flops_per_launch": flops},
        "cpu_baseline": None,
    }
    for name, q in res.items():
        line[f"{name}_value"] = 4.0 * E * args.steps / q["sec"]
        line[f"{name}_stage_ms"] = q["ms"]
        if name != "fp64" and "fp64" in res:
            U64 = res["fp64"]["U"]
            line[f"{name}_over_fp64"] = res["fp64"]["sec"] / q["sec"]
            line[f"{name}_max_rel_disp_err_vs_fp64"] = float(np.abs(q["U"] - U64).max() / np.abs(U64).max())
    print(json.dumps(line))


AI_FLOPS_PER_POINT = 55  # FP64 operations of one point's bond force (sqrt and division as one)


def run_ai(args):
    """Adaptive integration (ai_on 1 1, intg_point n): FP64 RK4 on the bar (20 x 10 x 10 unless
    dimensions are given), one JSON line.  The stage kernel interpolates every point's
    displacement from its MLS coefficients, so its bytes are the coefficient stream."""
    import pdmtel_amd as pm

    n1 = args.intg_point
    dims = (20, 10, 10) if (args.nx, args.ny, args.nz) == DEFAULT_DIMS else (args.nx, args.ny, args.nz)
    spec = pm.bar_spec(*dims, jitter=0.01, seed=0, steps=args.warmup + args.steps + 10)
    spec["cci"] = None
    spec["ai"] = (1, 1, n1)
    t0 = time.perf_counter()
    prob = pm.Problem(pm.deck_text(spec), precision=pm.PDM_FP64, graph_chunk=graph_chunk(args), energies=1)
    prob.prepare(graph_chunk(args))
    t_init = time.perf_counter() - t0
    P, M = prob.ai_counts
    print(f"# ai n={n1}: init {t_init:.2f} s, {prob.n} nodes, {prob.E} entries, {P} points, {M} coefficients",
          file=sys.stderr, flush=True)
    if args.warmup:
        prob.step(args.warmup)
    t0 = time.perf_counter()
    done = prob.step(args.steps)
    sec = time.perf_counter() - t0
    assert done == args.steps, f"ran {done} of {args.steps} steps"
    ms, launches = prob.kernel_timing()
    n, E = prob.n, prob.E
    # per launch: the coefficients (8 B) and their contributing-node ids (4 B), each point's
    # position / weight (32 B), entry offset (4 B) and coefficient offset (8 B), the node bytes
    nbytes = 12 * M + 44 * P + NODE_BYTES * n
    achieved = nbytes / (ms * 1e-3) / 1e9
    line = {
        "metric": f"bond-updates/s (RK4, adaptive integration, trapezoidal index {n1}, FP64)",
        "value": 4.0 * E * args.steps / sec,
        "unit": "bond-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"bar {dims[0]}x{dims[1]}x{dims[2]}, {n} nodes, {E} family entries, delta=3.015dx, "
                               f"jitter 0.01dx seed 0, ai_on 1 1 intg_point {n1}, RK4 {args.steps} steps",
                   "points": P, "mls_coefficients": M, "init_seconds": t_init},
        "point_evaluations_per_s": 4.0 * P * args.steps / sec,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_fgi_stage<AI> (one RK4 stage)", "kernel_avg_ms": ms, "launches": launches,
                     "bytes_per_launch": nbytes,
                     "fp64_tflops": P * (AI_FLOPS_PER_POINT + 6 * M / max(P, 1)) / (ms * 1e-3) / 1e12},
        "cpu_baseline": None,
    }
    prob.close()
    print(json.dumps(line))


def run_batch(args):
    """Batched specimens: FP32C, FP32C2 and FP64 runs of the same random specimens."""
    import numpy as np

    import pdmtel_amd as pm
    from pdmtel_amd.deck import specimen_specs

    specs = specimen_specs(args.specimens, seed=0, steps=args.steps + args.warmup + 10)
    decks = [pm.deck_text(s) for s in specs]
    layout = {}
    if args.layout:  # "edge,waves,families": every tile class takes this layout
        e, w, f = (int(v) for v in args.layout.split(","))
        layout = dict(tile_edge=e, waves_per_tile=w, families_per_wave=f)
    res = {}
    for name, prec in (("fp32c", pm.PDM_FP32C), ("fp32c2", pm.PDM_FP32C2), ("fp64", pm.PDM_FP64)):
        if args.precision not in ("all", name):
            continue
        print(f"# batch {name}: building {len(decks)} specimens", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        b = pm.Batch(decks, precision=prec, graph_chunk=graph_chunk(args), energies=1, verbose=1, **layout)
        b.prepare(graph_chunk(args))
        t_init = time.perf_counter() - t0
        if args.warmup:
            b.step(args.warmup)
        t0 = time.perf_counter()
        done = b.step(args.steps)
        t1 = time.perf_counter()
        assert done == args.steps
        ms, launches = b.kernel_timing()
        U = [b.state(k)[0] for k in range(b.nspec)]
        res[name] = dict(seconds=t1 - t0, init=t_init, kernel_ms=ms, launches=launches, U=U, E=sum(b.E),
                         E_live=sum(b.E_live), n=sum(b.n), tuning=b.tuning())
        print(f"# batch {name}: {args.specimens} specimens, {sum(b.n)} nodes, {sum(b.E)} entries, init {t_init:.1f} s, "
              f"{t1 - t0:.3f} s for {args.steps} steps, stage {ms:.4f} ms", file=sys.stderr, flush=True)
        b.close()
    if len(res) < 3:  # one precision (profiling runs): its rate and stage time only
        name, r = next(iter(res.items()))
        ks = r["kernel_ms"] * 1e-3
        bytes_launch = r["E_live"] * ENTRY_BYTES["fp64" if name == "fp64" else "fp32c"] + r["n"] * NODE_BYTES
        traffic = recorded_traffic(args, name)
        print(json.dumps({
            "metric": "bond-updates/s (RK4)", "value": 4.0 * r["E"] * args.steps / r["seconds"],
            "unit": "bond-updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": r["seconds"] / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64" if name == "fp64" else "f32+f64", "data": "synthetic",
            "config": {"workload": f"batch of {args.specimens} random specimens (seed 0)", "precision": name,
                       "layout": r["tuning"]},
            "roofline": {"bound": "hbm", "achieved": bytes_launch / ks / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": bytes_launch / ks / 1e9 / HBM_PEAK_GBS, "traffic": traffic,
                         "measured_gbs": traffic / ks / 1e9 if traffic else None,
                         "kernel": "k_stage, the tile classes' launches of one stage",
                         "kernel_avg_ms": r["kernel_ms"], "bytes_per_launch": bytes_launch},
            "cpu_baseline": None}))
        return
    E, n, E_live = res["fp32c"]["E"], res["fp32c"]["n"], res["fp32c"]["E_live"]
    errs = np.array([float(np.abs(a - c).max() / max(np.abs(c).max(), 1e-300))
                     for a, c in zip(res["fp32c"]["U"], res["fp64"]["U"])])
    errs2 = np.array([float(np.abs(a - c).max() / max(np.abs(c).max(), 1e-300))
                      for a, c in zip(res["fp32c2"]["U"], res["fp64"]["U"])])
    err = float(errs.max())
    rate = 4.0 * E * args.steps / res["fp32c"]["seconds"]
    bytes_launch = E_live * ENTRY_BYTES["fp32c"] + n * NODE_BYTES
    ks = res["fp32c"]["kernel_ms"] * 1e-3
    line = {
        "metric": "bond-updates/s (RK4), mixed vs FP64 at <=1e-6 rel err; static-solve wall s",
        "value": rate, "unit": "bond-updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": res["fp32c"]["seconds"] / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": f"batch of {args.specimens} random specimens (seed 0): nodes log-uniform in "
                               f"[1e3, 1e5], aspect ratios U[1,5], delta/dx in {{3.015, 4.015}}, edge crack at "
                               f"U[0,90] deg, E U[50,200] GPa; {n} nodes, {E} family entries; RK4 {args.steps} steps",
                   "precision": "fp32c", "nodes": n, "family_entries": E, "evaluated_entries": E_live,
                   "specimens": args.specimens,
                   "layout": res["fp32c"]["tuning"], "layout_fp32c2": res["fp32c2"]["tuning"],
                   "layout_fp64": res["fp64"]["tuning"]},
        "evaluated_bond_updates_per_s": 4.0 * E_live * args.steps / res["fp32c"]["seconds"],
        "fp64_value": 4.0 * E * args.steps / res["fp64"]["seconds"],
        "fp32c_over_fp64": res["fp64"]["seconds"] / res["fp32c"]["seconds"],
        "fp32c_max_rel_disp_err_vs_fp64": err,
        "fp32c_rel_err_per_specimen": {"median": float(np.median(errs)), "p99": float(np.quantile(errs, 0.99)),
                                       "over_1e-6": int((errs > 1e-6).sum())},
        "fp32c2_value": 4.0 * E * args.steps / res["fp32c2"]["seconds"],
        "fp32c2_over_fp64": res["fp64"]["seconds"] / res["fp32c2"]["seconds"],
        "fp32c2_max_rel_disp_err_vs_fp64": float(errs2.max()),
        "fp32c2_rel_err_per_specimen": {"median": float(np.median(errs2)), "p99": float(np.quantile(errs2, 0.99)),
                                        "over_1e-6": int((errs2 > 1e-6).sum())},
        "roofline": {"bound": "hbm", "achieved": bytes_launch / ks / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": bytes_launch / ks / 1e9 / HBM_PEAK_GBS, "traffic": recorded_traffic(args, "fp32c"),
                     "measured_gbs": (recorded_traffic(args, "fp32c") / ks / 1e9
                                      if recorded_traffic(args, "fp32c") else None),
                     "kernel": "k_stage (fused bond-force gather + RK4 stage), the tile classes' launches of "
                               "one stage over all specimens",
                     "kernel_avg_ms": res["fp32c"]["kernel_ms"], "bytes_per_launch": bytes_launch},
        "init_seconds": {"fp32c": res["fp32c"]["init"], "fp32c2": res["fp32c2"]["init"], "fp64": res["fp64"]["init"]},
        "cpu_baseline": None,
    }
    print(json.dumps(line))


def cpu_baseline(spec, steps):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    cores = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    o = pyoracle.Oracle(spec, bond_method=1, nthreads=cores)  # set-up is not timed
    o.nthreads = 1
    t0 = time.perf_counter()
    o.run(steps)
    dt = time.perf_counter() - t0
    # the same restatement with its node loop on all host cores given to this process (OpenMP)
    o.nthreads = cores
    t0 = time.perf_counter()
    o.run(steps)
    dt_all = time.perf_counter() - t0
    return dict(value=4.0 * o.E * steps / dt, seconds=dt, E=o.E, n=o.n, steps=steps,
                value_all=4.0 * o.E * steps / dt_all, seconds_all=dt_all, cores_all=cores)


def recorded_traffic(args, prec, workload=None):
    """FETCH_SIZE + WRITE_SIZE bytes per stage launch (a batch: per stage, its tile classes'
    launches together) recorded for this workload, else None."""
    wl = workload or args.workload
    path = TRAFFIC_FILES.get((wl, args.specimens if wl == "batch" else (args.nx, args.ny, args.nz)))
    if path is None or not os.path.exists(path):
        return None
    for line in open(path):
        m = re.match(r"^traffic_bytes_per_launch\s+(\w+)\s+([0-9.eE+]+)", line)
        if m and m.group(1) == prec:
            return float(m.group(2))
    return None


def calib_copy(nbytes):
    """A plain copy kernel of known size, for scaling FETCH_SIZE / WRITE_SIZE."""
    import torch

    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1.0)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"calib_copy_bytes": nbytes}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--nx", type=int, default=DEFAULT_DIMS[0])
    ap.add_argument("--ny", type=int, default=DEFAULT_DIMS[1])
    ap.add_argument("--nz", type=int, default=DEFAULT_DIMS[2])
    ap.add_argument("--chunk", type=int, default=0,
                    help="steps per graph (a divisor of --steps; default: automatic, up to 200)")
    ap.add_argument("--precision", choices=["all", "fp64", "fp32c", "fp32c2"], default="all")
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-static", action="store_true")
    ap.add_argument("--skip-plain", action="store_true", help="slit workload: leave out the plain-bar figures")
    ap.add_argument("--calib-copy", type=int, default=0, help="run only a copy kernel of this many bytes")
    ap.add_argument("--workload", choices=["slit", "bar", "batch", "plate", "fgi", "ai"], default="slit",
                    help="slit: 200x40x40 bar with a 30 deg through-section slit (default); bar: the same bar "
                         "without it; batch: random specimens; plate: --nx --ny --nz plate "
                         "with two crossing slits (400 400 50 for the 8M-node plate); fgi: the bar with fixed "
                         "Gaussian integration (--intg-point), FP64; ai: adaptive integration with MLS "
                         "(--intg-point), FP64, on the 20x10x10 bar unless --nx --ny --nz are given")
    ap.add_argument("--intg-point", type=int, default=2)
    ap.add_argument("--specimens", type=int, default=512)
    ap.add_argument("--layout", default="", help="edge,waves,families: fixed tile layout (default: autotuned)")
    ap.add_argument("--static-only", action="store_true",
                    help="only the static solve of the edge-cracked bar in --precision (one handle)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this library runs on one GPU; --gpus {args.gpus} is not supported "
              "(the path is not sharded across GPUs)", file=sys.stderr)
        sys.exit(2)
    if args.calib_copy:
        calib_copy(args.calib_copy)
        return
    if args.workload == "batch":
        run_batch(args)
        return
    if args.workload == "fgi":
        run_fgi(args)
        return
    if args.workload == "ai":
        run_ai(args)
        return
    if args.static_only:  # one precision's static solve, one handle in the process (profiling)
        name = "fp32c" if args.precision == "all" else args.precision
        st, _ = run_static(args, (name,))
        r = st[name]
        print(json.dumps({"metric": "static-solve wall s", "value": r["wall"], "unit": "s", "n_gpus": 1,
                          "higher_is_better": False, "dtype": "f64" if name == "fp64" else "f32+f64",
                          "data": "synthetic",
                          "config": {"workload": ("static solve of the edge-cracked bar" if args.workload in
                                                  ("bar", "slit") else f"static solve of the {args.workload}"),
                                     "precision": name},
                          "iterations": r["iterations"], "rel_residual": r["rel_residual"]}))
        return

    import numpy as np

    spec = workload_spec(args, args.warmup + args.steps + 10)
    precs = PRECISIONS if args.precision == "all" else (args.precision,)
    res = {}
    for name in precs:
        res[name] = run_precision(args, spec, precision_code(name))
        r = res[name]
        print(f"# {args.workload} {name}: {r['seconds']:.3f} s for {args.steps} steps, kernel {r['kernel_ms']:.4f} "
              f"ms/launch, init {r['init_seconds']:.2f} s, layout {r['tuning']}, launch {r['launch']}",
              file=sys.stderr, flush=True)
    head_prec = "fp32c" if "fp32c" in res else precs[0]
    head = res[head_prec]
    E, n, E_live, E_ub = head["E"], head["n"], head["E_live"], head["E_unbroken"]

    def rate(r):
        return 4.0 * r["E_unbroken"] * args.steps / r["seconds"]

    def roof(name, r):
        eb = ENTRY_BYTES["fp64" if name == "fp64" else "fp32c"]
        fb = FORMAT_ENTRY_BYTES["fp64" if name == "fp64" else "fp32c"]
        b = r["E_live"] * eb + r["n"] * NODE_BYTES
        ks = r["kernel_ms"] * 1e-3
        traffic = recorded_traffic(args, name)
        ach = b / ks / 1e9 if ks > 0 else 0.0
        return {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                "traffic": traffic, "measured_gbs": (traffic / ks / 1e9) if (traffic and ks > 0) else None,
                "kernel": "k_stage (fused bond-force gather + RK4 stage)", "kernel_avg_ms": r["kernel_ms"],
                "launches": r["launches"], "bytes_per_launch": b, "format_bytes": r["E_live"] * fb + r["n"] * NODE_BYTES}

    def err_vs(name, ref="fp64"):
        if name not in res or ref not in res:
            return None
        U64, U = res[ref]["U"], res[name]["U"]
        return float(np.abs(U - U64).max() / np.abs(U64).max())

    desc = {"slit": ", one through-section slit at the centre, normal 30 deg to x",
            "bar": "", "plate": ", two crossing slits"}[args.workload]
    line = {
        "metric": "bond-updates/s (RK4), mixed vs FP64 at <=1e-6 rel err; static-solve wall s",
        "value": rate(head),
        "unit": "bond-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["seconds"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if head_prec != "fp64" else "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{'bar' if args.workload == 'slit' else args.workload} {args.nx}x{args.ny}x{args.nz}, "
                        f"{n} nodes, {E} family entries ({E - E_ub} cut), delta=3.015dx, jitter 0.01dx seed 0"
                        + desc + f", RK4 {args.steps} steps",
            "precision": head_prec,
            "nodes": n,
            "family_entries": E,
            "unbroken_entries": E_ub,
            "evaluated_entries": E_live,
            "graph_chunk": graph_chunk(args),
            "layout": {k: r["tuning"] for k, r in res.items()},
            "stage_launch": {k: r["launch"] for k, r in res.items()},
        },
        "evaluated_bond_updates_per_s": 4.0 * E_live * args.steps / head["seconds"],
        "fp64_value": rate(res["fp64"]) if "fp64" in res else None,
        "fp32c_over_fp64": (res["fp64"]["seconds"] / res["fp32c"]["seconds"]) if {"fp32c", "fp64"} <= set(res) else None,
        "fp32c_max_rel_disp_err_vs_fp64": err_vs("fp32c"),
        "fp32c2_value": rate(res["fp32c2"]) if "fp32c2" in res else None,
        "fp32c2_over_fp64": (res["fp64"]["seconds"] / res["fp32c2"]["seconds"]) if {"fp32c2", "fp64"} <= set(res) else None,
        "fp32c2_max_rel_disp_err_vs_fp64": err_vs("fp32c2"),
        "roofline": roof(head_prec, head),
        "cpu_baseline": None,
    }
    for name in res:
        if name != head_prec:
            line[f"roofline_{name}"] = roof(name, res[name])
    if args.workload == "slit" and not args.skip_plain:
        # the bar without the slit (the FP64 workload), same steps, FP64 and FP32C
        pspec = workload_spec(args, args.warmup + args.steps + 10, "bar")
        pr = {name: run_precision(args, pspec, precision_code(name)) for name in ("fp64", "fp32c")}
        for name, r in pr.items():
            print(f"# bar {name}: {r['seconds']:.3f} s for {args.steps} steps, kernel {r['kernel_ms']:.4f} ms/launch",
                  file=sys.stderr, flush=True)
        line["plain_bar"] = {
            "workload": f"bar {args.nx}x{args.ny}x{args.nz} without the slit, {pr['fp64']['E']} family entries",
            "fp64_value": rate(pr["fp64"]), "fp32c_value": rate(pr["fp32c"]),
            "fp32c_over_fp64": pr["fp64"]["seconds"] / pr["fp32c"]["seconds"],
            "fp32c_max_rel_disp_err_vs_fp64": float(np.abs(pr["fp32c"]["U"] - pr["fp64"]["U"]).max() /
                                                    np.abs(pr["fp64"]["U"]).max()),
            "roofline_fp64": roof("fp64", pr["fp64"]), "roofline_fp32c": roof("fp32c", pr["fp32c"]),
        }
        line["plain_bar"]["roofline_fp64"]["traffic"] = recorded_traffic(args, "fp64", "bar")
        line["plain_bar"]["roofline_fp32c"]["traffic"] = recorded_traffic(args, "fp32c", "bar")
    if not args.skip_static:
        st, serr = run_static(args)
        line["static_solve"] = {
            "workload": (f"bar {args.nx}x{args.ny}x{args.nz} jitter 0.01dx seed 0, edge crack to the mid-plane, "
                         "normal 30 deg to x, PCG to 1e-10 relative residual") if args.workload in ("bar", "slit") else
                        (f"plate {args.nx}x{args.ny}x{args.nz} jitter 0.01dx seed 0, two crossing slits, "
                         "PCG to 1e-10 relative residual"),
            "wall_s": st["fp32c"]["wall"], "iterations": st["fp32c"]["iterations"],
            "refinements": st["fp32c"]["refinements"], "rel_residual": st["fp32c"]["rel_residual"],
            "fp64_wall_s": st["fp64"]["wall"], "fp64_iterations": st["fp64"]["iterations"],
            "max_rel_disp_err_vs_fp64": serr["fp32c"],
            "fp32c2_wall_s": st["fp32c2"]["wall"], "fp32c2_iterations": st["fp32c2"]["iterations"],
            "fp32c2_max_rel_disp_err_vs_fp64": serr["fp32c2"],
        }
    if not args.skip_cpu:
        cb = cpu_baseline(spec, steps=args.cpu_steps)
        line["cpu_baseline"] = {
            "value": cb["value"],
            "unit": "bond-updates/s",
            "cores": 1,
            "kind": "port",
            "sample": f"the oracle (serial C restatement of the reference RK4 loop; not calibrated against the "
                      f"reference binary, which does not build here), same {cb['n']}-node workload, "
                      f"{cb['steps']} RK4 steps ({4 * cb['E'] * cb['steps']} bond-updates) in {cb['seconds']:.1f} s",
            "all_cores": {"value": cb["value_all"], "cores": cb["cores_all"],
                          "sample": f"the next {cb['steps']} steps with the node loop on {cb['cores_all']} OpenMP "
                                    f"threads, {cb['seconds_all']:.1f} s"},
        }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
