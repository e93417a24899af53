#!/usr/bin/env python3
"""Benchmark of the lattice-evolution hot path on one MI355X.

Workload (the metric's configuration): Klein-Gordon in a radiation-dominated FRW
background, N=512^3 lattice, L=512, m=1, a(t) = (t/t_i)^(1/2) with t_i = 1
(a1=1, H1=0.5, t1=1), initial phi a Gaussian random field (seed 2, P(k) ~
exp(-k^2/(2*0.25^2)), rms 1), dphi/dt = 0, classic RK4 with dt = 0.05.

A "step" is one RK4 step over the whole lattice. Two runs are timed on the same
initial state: run A stores the fields in FP32 with compensated hi+lo state and
double arithmetic (the headline `value`), run B is full FP64. Inputs are
resident in HBM before the timed region. The line also carries the roofline of
the step's kernels (HIP events on the launch stream) and the oracle (CPU
restatement of the reference path) timed on this host on a bounded sample.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cosmic-fields-lite_amd"))

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per site-step of one RK4 step (per site: phi and pi), the
# canonical schedule of one fused kernel per stage over state x, stage input y
# and increment D:  s1 read x, write y, D (3 passes); s2, s3 read y, x, D, write
# y, D (5 each); s4 read y, x, D, write x (4).
#   FP64: 17 passes x 2 x 8 B = 272 B.
#   FP32-compensated: s4 also reads and writes the float compensation term:
#   19 passes x 2 x 4 B = 152 B.
# The FP64 kernels here keep the partial RK4 sum instead of D, so s4 does not
# read x and they move 16 passes (256 B); the canonical 272 B is still the
# figure used for `achieved`.
BYTES_PER_SITE_STEP = {"f64": 272, "f32c": 152}
DESIGN_BYTES_PER_SITE_STEP = {"f64": 256, "f32c": 152}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-steps", type=int, default=8)
    ap.add_argument("--calib-copy", action="store_true",
                    help="first run 3 device-to-device copies of 1 GiB (cf_d2d), the known-size "
                         "kernel that calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE")
    return ap.parse_args()


def cpu_baseline(N, steps):
    """Oracle (CPU restatement of the reference path, 1 thread) on a bounded
    sample of the same workload: `steps` FP64 RK4 steps at N^3."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import cf_oracle as O
    O.set_threads(1)
    eq = O.Equation(N, float(N), 1.0, 0.0, O.Cosmo.radiation(1.0, 0.5, 1.0))
    rng = np.random.default_rng(2)
    x = np.concatenate([rng.standard_normal(N ** 3), np.zeros(N ** 3)])
    eq._w(5).fill(0.0)  # touch the work arrays before timing
    t = time.perf_counter()
    eq.rk4_run(x, 1.0, 0.05, steps)
    el = time.perf_counter() - t
    return {"value": N ** 3 * steps / el / 1e9, "unit": "GSite*RK4-step/s", "cores": 1, "kind": "port",
            "sample": f"oracle FP64 RK4, KG-FRW N={N}^3, {steps} step(s), 1 thread, {el:.1f} s"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "see DESIGN.md for how the lattice would shard", file=sys.stderr)
        sys.exit(2)
    import torch
    import cfl_amd as cfl

    if args.calib_copy:
        a = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            cfl.check(cfl.lib.cf_d2d(cfl._ptr(b), cfl._ptr(a), 1 << 30, cfl._stream()))
        torch.cuda.synchronize()
        del a, b

    N, L, m, dt, t0 = args.n, float(args.n), 1.0, 0.05, 1.0
    cosmo = cfl.radiation(1.0, 0.5, 1.0)
    p = cfl.kg_params(N, L, m, 0.0, cosmo)
    stream = torch.cuda.current_stream()
    plan = cfl.SpectrumPlan(N)
    x0 = plan.init_grf_state(2, cfl.CF_SPEC_GAUSS, (0.25,), L, 1.0)
    torch.cuda.synchronize()
    del plan

    results = {}
    for mode in ("f32c", "f64"):
        if mode == "f64":
            x = x0.clone()
            work = cfl.alloc_work(N, cfl.CF_PREC_F64)
            run = lambda k0, n: cfl.rk4_run_f64(p, x, t0, k0, n, dt, work)  # noqa: E731
        else:
            hi, lo = cfl.split_f64(x0)
            work = cfl.alloc_work(N, cfl.CF_PREC_F32C)
            run = lambda k0, n: cfl.rk4_run_f32c(p, hi, lo, t0, k0, n, dt, work)  # noqa: E731
        run(0, args.warmup)
        torch.cuda.synchronize()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        w0 = time.perf_counter()
        ev0.record(stream)
        run(args.warmup, args.steps)
        ev1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - w0
        ms = ev0.elapsed_time(ev1) / args.steps
        rate = N ** 3 / (ms * 1e-3) / 1e9
        results[mode] = {"ms_per_step": ms, "gsite_steps": rate, "wall_s": wall,
                         "achieved_GBs": BYTES_PER_SITE_STEP[mode] * N ** 3 / (ms * 1e-3) / 1e9}
        del work
        torch.cuda.synchronize()

    A, B = results["f32c"], results["f64"]
    line = {
        "metric": "GSite*RK4-step/s + % HBM roofline, FP32-comp vs FP64 at N=512^3",
        "value": A["gsite_steps"],
        "unit": "GSite*RK4-step/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": A["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "storage": "f32 hi + f32 compensation (run A)",
        "data": "synthetic (Gaussian random field, seed 2)",
        "config": {"workload": f"KG-FRW radiation era N={N}^3 L={L:g} m=1 dt=0.05 RK4, FP32-comp (A) vs FP64 (B)",
                   "N": N, "L": L, "dt": dt, "parallelism": "single GPU"},
        "fp64": {"value": B["gsite_steps"], "ms_per_step": B["ms_per_step"],
                 "roofline_frac": B["achieved_GBs"] / HBM_PEAK_GBS,
                 "bytes_per_site_step": BYTES_PER_SITE_STEP["f64"],
                 "design_bytes_per_site_step": DESIGN_BYTES_PER_SITE_STEP["f64"]},
        "speedup_f32c_vs_f64": A["gsite_steps"] / B["gsite_steps"],
        "roofline": {"bound": "hbm", "kernel": "rk4 step, FP32-comp (4 fused stage kernels)",
                     "achieved": A["achieved_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": A["achieved_GBs"] / HBM_PEAK_GBS,
                     "bytes_per_site_step": BYTES_PER_SITE_STEP["f32c"], "traffic": None},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(N, args.cpu_steps)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
