#!/usr/bin/env python3
"""Benchmark: bond-updates/s of the RK4 dynamic solve on one MI355X.

Workload (default): the 200 x 40 x 40 bar (320,000 nodes, dx = 1 mm,
delta = 3.015 dx, 73,728,544 family entries), node jitter U(-0.01, 0.01) dx
from seed 0, E = 72 GPa, rho = 2440, x = 0 face fixed, 1 MPa +x traction on
the x = L layer, RK4 at 0.8 x the stable dt.  One step = 4 stages, each a
full bond-force gather: 4 E bond-updates.

Both paths run the same K steps after W warm-up steps: the FP64 path
(reference arithmetic) and the FP32-compensated path.  `value` is the
FP32-compensated rate; the line also carries the FP64 rate and the max
relative displacement error of FP32C vs FP64 after the K + W steps.
Inputs are resident in HBM before the timed region (pdm_init builds them
on the GPU); the timed region is pdm_step_rk4 (hipGraph replays).

roofline: dominant kernel = the fused bond-force / RK4-stage kernel; its
average launch time comes from HIP events on the library's stream around
the timed region (pdm_last_kernel_timing); algorithmic bytes per launch =
E x (4 B col + coefficient bytes) + N x 288 B per-node state traffic.

cpu_baseline: the oracle (faithful serial restatement of the reference's
RK4 loop) on the same bar for one step, 1 thread, on this host.
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
NODE_BYTES = 288  # per node per stage: NodeRec read 64 + row_ptr 16 + mat/bc/nb 12 + Vb/V r+w 64
#                  + U r/w 64 + k r/w 64 + next-stage displacement write 32, rounded to 288


def workload_spec(args):
    import pdmtel_amd as pm

    total = args.warmup + args.steps + 10
    spec = pm.bar_spec(args.nx, args.ny, args.nz, jitter=0.01, seed=0, steps=total)
    return spec


def run_precision(args, spec, precision):
    import numpy as np

    import pdmtel_amd as pm

    text = pm.deck_text(spec)
    t0 = time.perf_counter()
    prob = pm.Problem(text, precision=precision, graph_chunk=args.chunk, energies=1)
    t_init = time.perf_counter() - t0
    if args.warmup:
        prob.step(args.warmup)
    t0 = time.perf_counter()
    done = prob.step(args.steps)
    t1 = time.perf_counter()
    assert done == args.steps, f"ran {done} of {args.steps} steps"
    ms_kernel, launches = prob.kernel_timing()
    U, _ = prob.state()
    res = dict(n=prob.n, E=prob.E, seconds=t1 - t0, init_seconds=t_init, kernel_ms=ms_kernel, launches=launches,
               U=U, energy=prob.energy())
    prob.close()
    return res


def cpu_baseline(spec, steps=1):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    o = pyoracle.Oracle(spec, bond_method=1, nthreads=1)
    t0 = time.perf_counter()
    o.run(steps)
    dt = time.perf_counter() - t0
    return dict(value=4.0 * o.E * steps / dt, seconds=dt, E=o.E, n=o.n)


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
    ap.add_argument("--nx", type=int, default=200)
    ap.add_argument("--ny", type=int, default=40)
    ap.add_argument("--nz", type=int, default=40)
    ap.add_argument("--chunk", type=int, default=50)
    ap.add_argument("--precision", choices=["both", "fp64", "fp32c"], default="both")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--calib-copy", type=int, default=0, help="run only a copy kernel of this many bytes")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding is described in DESIGN.md, not launched)", file=sys.stderr)
        sys.exit(2)
    if args.calib_copy:
        calib_copy(args.calib_copy)
        return

    import numpy as np

    import pdmtel_amd as pm

    spec = workload_spec(args)
    res = {}
    if args.precision in ("both", "fp64"):
        res["fp64"] = run_precision(args, spec, pm.PDM_FP64)
        print(f"# fp64: {res['fp64']['seconds']:.3f} s for {args.steps} steps, "
              f"kernel {res['fp64']['kernel_ms']:.4f} ms/launch, init {res['fp64']['init_seconds']:.2f} s",
              file=sys.stderr, flush=True)
    if args.precision in ("both", "fp32c"):
        res["fp32c"] = run_precision(args, spec, pm.PDM_FP32C)
        print(f"# fp32c: {res['fp32c']['seconds']:.3f} s for {args.steps} steps, "
              f"kernel {res['fp32c']['kernel_ms']:.4f} ms/launch", file=sys.stderr, flush=True)
    head = res.get("fp32c") or res["fp64"]
    head_prec = "fp32c" if "fp32c" in res else "fp64"
    E, n = head["E"], head["n"]
    rate = 4.0 * E * args.steps / head["seconds"]
    coef_bytes = 4 if head_prec == "fp32c" else 8
    bytes_launch = E * (4 + coef_bytes) + n * NODE_BYTES
    kernel_s = head["kernel_ms"] * 1e-3
    achieved = bytes_launch / kernel_s / 1e9 if kernel_s > 0 else 0.0
    err = None
    if "fp32c" in res and "fp64" in res:
        U64, U32 = res["fp64"]["U"], res["fp32c"]["U"]
        err = float(np.abs(U32 - U64).max() / np.abs(U64).max())
    line = {
        "metric": "bond-updates/s (RK4), mixed vs FP64 at <=1e-6 rel err; static-solve wall s",
        "value": rate,
        "unit": "bond-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["seconds"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if head_prec == "fp32c" else "f64",
        "data": "synthetic",
        "config": {
            "workload": f"bar {args.nx}x{args.ny}x{args.nz}, {n} nodes, {E} family entries, "
                        f"delta=3.015dx, jitter 0.01dx seed 0, RK4 {args.steps} steps",
            "precision": head_prec,
            "nodes": n,
            "family_entries": E,
        },
        "fp64_value": (4.0 * E * args.steps / res["fp64"]["seconds"]) if "fp64" in res else None,
        "fp32c_max_rel_disp_err_vs_fp64": err,
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "kernel": "k_stage (fused bond-force gather + RK4 stage)",
            "kernel_avg_ms": head["kernel_ms"],
            "bytes_per_launch": bytes_launch,
        },
        "cpu_baseline": None,
    }
    if not args.skip_cpu:
        cb = cpu_baseline(spec, steps=1)
        line["cpu_baseline"] = {
            "value": cb["value"],
            "unit": "bond-updates/s",
            "cores": 1,
            "kind": "port",
            "sample": f"oracle serial restatement of the reference RK4 loop, same {cb['n']}-node bar, "
                      f"1 RK4 step ({4 * cb['E']} bond-updates) in {cb['seconds']:.1f} s",
        }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
