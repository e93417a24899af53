#!/usr/bin/env python3
"""Benchmark of the lattice-evolution hot path on one MI355X.

Workload (the metric's configuration): Klein-Gordon in a radiation-dominated FRW
background, N=512^3 lattice, L=512, m=1, a(t) = (t/t_i)^(1/2) with t_i = 1
(a1=1, H1=0.5, t1=1), initial phi a Gaussian random field (seed 2, P(k) ~
exp(-k^2/(2*0.25^2)), rms 1), dphi/dt = 0, classic RK4 with dt = 0.05.

A "step" is one RK4 step over the whole lattice. Run A stores the fields in
FP32 with a compensated hi+lo state, in two arithmetic forms: FP32-comp (f32c, all
arithmetic in double) and the single-precision stencil (f32s, float forces on the
float stage inputs, the increment and the state update in double); run B is full
FP64. They are timed with one temporally blocked kernel per step (tb_*) and, for f32c
and FP64, with four fused stage kernels per step; every path produces the bits of
its scheme (FP64: the reference's arithmetic order). Before the f32s path is timed,
the library's tuner picks the chunk length and the pace of its lock-step for this
lattice (untimed set-up, about 2.7 s at N=512). The headline `value` is run A's fastest path. Inputs are resident in HBM before the timed region, and the GPU
has run 0.3 s of untimed steps before the first path, so that every path is timed at
settled clocks (each then does its own --warmup steps). The line
also carries the roofline of the headline kernel (HIP events on the launch
stream) and the oracle (CPU restatement of the reference path) timed on this
host on a bounded sample.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N] [--no-cpu] [--paths ...]
    python bench.py --workload precision --steps 1000     # run A vs run B errors
    python bench.py --workload ensemble --steps 500       # 96 uneven-N realisations
    python bench.py --workload adjoint                    # loss + gradient, N=128, 1024 steps
    python bench.py --workload variants --steps 10        # fixed-curvature / sqrt-potential steps
    python bench.py --workload slab                       # one GPU's slab of N=512 (64 planes + ghosts)
    python bench.py --workload pacing --steps 20          # the f32s step at a list of paces
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
BYTES_PER_SITE_STEP = {"f64": 272, "f32c": 152, "f32s": 152}
DESIGN_BYTES_PER_SITE_STEP = {"f64": 256, "f32c": 152, "tb_f64": 32, "tb_f32c": 32, "tb_f32s": 32}
KERNEL_NAMES = {"f64": "cf::k_march<double, 1, S, false> (4 stage kernels)",
                "f32c": "cf::k_march<float, 2, S, false> (4 stage kernels)",
                "tb_f64": "cf::k_rk4_tb_ring<false>",
                "tb_f32c": "cf::k_rk4_tb<float, true, false>",
                "tb_f32s": "cf::k_rk4_tb_sa<false, {paced}, {rows}, {cols}>"}  # region shape per N
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "rk4_tb_N512_pmc_traffic.md")
# What bounds each path's kernel at N=512, from the committed PMC summaries: the FP32-comp
# blocked kernel issues VALU 76% of its cycles (float <-> double conversions on top of the
# FP64 arithmetic); the FP64 blocked kernel runs as fast with its arithmetic cut to a
# skeleton, so its time is the memory traffic including the halo re-reads.
LIMITER = {"tb_f32c": "vector-ALU issue", "tb_f64": "memory: HBM + L2 misses on halo re-reads",
           "tb_f32s": "memory: halo-inclusive plane loads at the paced lock-step (L2 + HBM)",
           "f32c": "hbm", "f64": "hbm"}


def traffic_per_launch(kernel, N):
    """Calibrated HBM bytes per launch (FETCH_SIZE + WRITE_SIZE) of `kernel` from the
    committed rocprofv3 PMC summary (lines `traffic <kernel template> <N> <bytes>`), or None."""
    if not os.path.exists(TRAFFIC_PROFILE):
        return None
    key = kernel.split(" (")[0]
    for line in open(TRAFFIC_PROFILE):
        if line.startswith("traffic ") and key in line:
            parts = line.split()
            if int(parts[-2]) == N:
                return float(parts[-1])
    return None


VALU_PROFILE = os.path.join(ROOT, "profiles", "rk4_tb_N512_pmc_valu.md")


def valu_busy(kernel, N):
    """VALU-busy fraction of `kernel` at size N from the committed rocprofv3 PMC
    summary (lines `valu_busy <kernel template> <N> <fraction>`), or None."""
    if not os.path.exists(VALU_PROFILE):
        return None
    key = kernel.split(" (")[0]
    for line in open(VALU_PROFILE):
        if line.startswith("valu_busy ") and key in line:
            parts = line.split()
            if int(parts[-2]) == N:
                return float(parts[-1])
    return None


SETTLE_S = 0.3  # untimed sustained load before the first timed path


def settle(cfl, p, x0, seconds=SETTLE_S):
    """Untimed blocked FP64 steps for about `seconds` of wall time. Under sustained load
    the GPU's clocks settle over the first tens of milliseconds (the first ~20 launches of
    whichever path is timed first ran up to 25% slower), so every path is then timed at
    settled clocks instead of the first one paying for the ramp."""
    import torch
    a, b = x0.clone(), torch.empty_like(x0)
    t_end = time.perf_counter() + seconds
    k = 0
    while time.perf_counter() < t_end:
        for _ in range(10):
            cfl.rk4_step_tb_f64(p, a, b, 1.0 + k * 0.05, 0.05)
            a, b = b, a
            k += 1
        torch.cuda.synchronize()
    del a, b


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-steps", type=int, default=8)
    ap.add_argument("--chunk", type=int, default=0, help="a-planes per block of the temporally blocked kernel")
    ap.add_argument("--paths", default="tb_f32s,tb_f32c,tb_f64,f32c,f64",
                    help="headline: kernels to time (tb_f32s, tb_f32c, tb_f64: one temporally blocked kernel "
                         "per step; f32c, f64: four fused stage kernels per step)")
    ap.add_argument("--run-a", choices=("f32s", "f32c"), default="f32s",
                    help="precision, ensemble: the arithmetic of the compensated storage (f32s: float "
                         "stencil, double increment; f32c: FP64)")
    ap.add_argument("--save-dir", default="", help="ensemble: where to save the averaged density spectra (.npy)")
    ap.add_argument("--workload", choices=("headline", "ensemble", "precision", "adjoint", "variants", "slab",
                                           "pacing"),
                    default="headline",
                    help="headline: N=512 FRW, FP32-comp vs FP64; ensemble: 96 realisations of N=64/96/128; "
                         "precision: run A (FP32-comp) vs run B (FP64) errors every 100 steps; "
                         "adjoint: loss + gradient through 1024 steps at N=128; pacing: the f32s step at "
                         "a list of paces")
    ap.add_argument("--paces", default="0,100,110,115,120,125,130,135,140,150,160",
                    help="pacing: ticks of the 100 MHz real-time counter per plane iteration (0: unpaced); "
                         "the tuner's pick is timed as well")
    ap.add_argument("--report-every", type=int, default=100, help="precision: steps between error reports")
    ap.add_argument("--calib-copy", action="store_true",
                    help="first run 3 device-to-device copies of 1 GiB (cf_d2d), the known-size "
                         "kernel that calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE")
    return ap.parse_args()


def cpu_baseline(N, steps):
    """Oracle (CPU restatement of the reference path) on a bounded sample of the
    same workload: `steps` FP64 RK4 steps at N^3, on one thread (the reference CPU
    path is single-threaded) and, beside it, with the a-loop split over the host's
    cores (same per-site arithmetic, so the same bits)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import cf_oracle as O
    eq = O.Equation(N, float(N), 1.0, 0.0, O.Cosmo.radiation(1.0, 0.5, 1.0))
    rng = np.random.default_rng(2)
    x = np.concatenate([rng.standard_normal(N ** 3), np.zeros(N ** 3)])
    eq._w(5).fill(0.0)  # touch the work arrays before timing

    def timed(threads):
        O.set_threads(threads)
        t = time.perf_counter()
        eq.rk4_run(x, 1.0, 0.05, steps)
        return time.perf_counter() - t

    el = timed(1)
    cores = max(1, min(16, len(os.sched_getaffinity(0))))
    el_mt = timed(cores)
    out = {"value": N ** 3 * steps / el / 1e9, "unit": "GSite*RK4-step/s", "cores": 1, "kind": "port",
           "sample": f"oracle FP64 RK4, KG-FRW N={N}^3, {steps} step(s), 1 thread, {el:.1f} s",
           "all_cores": {"value": N ** 3 * steps / el_mt / 1e9, "cores": cores,
                         "sample": f"same, a-loop over {cores} threads, {el_mt:.1f} s"}}
    ref = reference_integrator_timing(N, x, cores)
    if ref:
        out["reference_integrator"] = ref
    return out


def reference_integrator_timing(N, x, cores, steps=2):
    """The reference's own integrator (vendored Boost.odeint integrate_const +
    runge_kutta4_classic + src/odeint_eigen/eigen_operations.hpp, built from the
    reference tree into oracle/_ref) around the oracle's right-hand side, timed inside
    the binary (std::chrono::steady_clock around integrate_const), once under the
    oracle's flags and once under the reference's (-O3 -ffast-math; x86-64-v3 for
    -march=native). None when the binaries were not built."""
    import subprocess
    import tempfile
    bins = {"reference_flags": os.path.join(ROOT, "oracle", "_ref", "odeint_ref_fastmath"),
            "oracle_flags": os.path.join(ROOT, "oracle", "_ref", "odeint_ref")}
    if not all(os.path.exists(b) for b in bins.values()):
        return None
    res = {"what": f"reference integrator around the oracle RHS, KG-FRW N={N}^3, {steps} RK4 steps per run "
                   "(equations.cpp itself needs fftw3.h, absent here)", "unit": "GSite*RK4-step/s"}
    with tempfile.TemporaryDirectory() as d:
        x.tofile(os.path.join(d, "in.bin"))
        for name, exe in bins.items():
            res[name] = {}
            for th in (1, cores):
                r = subprocess.run([exe, os.path.join(d, "in.bin"), d, str(N), repr(float(N)), "1.0", "0.0", "1.0",
                                    "0.5", "1.0", "1.0", repr(1.0 + steps * 0.05), "0.05", str(th)],
                                   capture_output=True, text=True, check=True)
                el = float(r.stderr.split("elapsed_s")[1].split()[0])
                res[name][f"{th}_threads"] = N ** 3 * int(r.stdout) / el / 1e9
    return res


def ensemble_workload(args):
    """96 seeded realisations (seeds 100-131: N=64, 132-163: N=96, 164-195: N=128,
    L=N), FRW Klein-Gordon with m=1 plus lambda phi^4/4 (lambda=0.1), compensated FP32
    storage (--run-a: f32s, float forces with a double increment and paced blocks, or
    f32c, FP64 arithmetic), P(k) ~ exp(-k^2/(2*0.5^2)), rms 0.5, RK4 dt=0.05, advanced
    by one launch per step captured in a hipGraph, with the plan (grid, and per N the
    block region shape and a-chunk; for f32s the pace) picked by the autotuner. Reports
    aggregate GSite*step/s and the ensemble-averaged density spectrum of each N group."""
    import numpy as np
    import torch
    import cfl_amd as cfl
    Ns = [64] * 32 + [96] * 32 + [128] * 32
    seeds = list(range(100, 196))
    dt, t0 = 0.05, 1.0
    steps = args.steps
    prec = cfl.CF_PREC_F32S if args.run_a == "f32s" else cfl.CF_PREC_F32C
    ens = cfl.Ensemble(Ns, [float(n) for n in Ns], 1.0, 0.1, cfl.radiation(1.0, 0.5, 1.0), prec)
    ens.init_grf(seeds, cfl.CF_SPEC_GAUSS, (0.5,), 0.5)
    tuned_grid = ens.tune(dt, reps=3)
    plan = {n: ens.get_plan(n)[0] for n in (64, 96, 128)}
    regions = {n: "24x40" if ens.get_region(n) == cfl.CF_ENS_REGION_24X40 else "32x32" for n in (64, 96, 128)}
    # capture + first launch outside the timed region (the graph is reused)
    ens.run(t0, 0, steps, dt, use_graph=True)
    torch.cuda.synchronize()
    ens.init_grf(seeds, cfl.CF_SPEC_GAUSS, (0.5,), 0.5)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    w0 = time.perf_counter()
    ev0.record(stream)
    ens.run(t0, 0, steps, dt, use_graph=True)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    ms = ev0.elapsed_time(ev1)
    rate = ens.sites * steps / (ms * 1e-3) / 1e9
    t_end = t0 + steps * dt
    groups = {}
    for r, n in enumerate(Ns):
        sp = ens.density_spectrum(r, t_end).cpu().numpy()
        groups.setdefault(n, []).append(sp)
    summary = {}
    for n, lst in groups.items():
        avg = np.zeros_like(lst[0])
        for sp in lst:  # fixed order
            avg += sp
        avg /= len(lst)
        if args.save_dir:
            np.save(os.path.join(args.save_dir, f"ensemble_rho_spectrum_N{n}.npy"), avg)
        summary[str(n)] = {"bins": int(avg.size), "sum": float(avg.sum()), "first_bins": [float(v) for v in avg[1:5]]}
    line = {
        "metric": f"aggregate GSite*RK4-step/s, ensemble of 96 uneven-N realisations (FP32-comp {args.run_a}, "
                  "hipGraph)",
        "value": rate, "unit": "GSite*RK4-step/s", "n_gpus": 1, "steps": steps, "warmup": 0,
        "ms_per_step": ms / steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 stencil, f64 increment" if args.run_a == "f32s" else "f64",
        "storage": "f32 hi + f32 compensation", "data": "synthetic (Gaussian random fields, seeds 100-195)",
        "config": {"workload": "ensemble 32xN=64 + 32xN=96 + 32xN=128, L=N, FRW KG m=1 lambda=0.1, dt=0.05",
                   "sites": ens.sites, "tuned_grid": "flat (one block per unit)" if tuned_grid == cfl.CF_ENS_GRID_FLAT else tuned_grid, "tuned_chunks": plan, "tuned_regions": regions, "tuned_pace": ens.get_pace(), "graph_wall_s": wall},
        "ensemble_avg_density_spectra": summary,
    }
    print(json.dumps(line), flush=True)


def observer_timing(cfl, plan, x, N, L, m, cosmo, t, stream, reps=5):
    """One observation of ConstIntervalObserver (observer.hpp:87-124) on the
    device: the field (mode) spectrum, the energy density and its spectrum, the
    density slice and axis average. Milliseconds per observation (HIP events)."""
    import torch
    a_t = cfl.cosmo_a(cosmo, t)
    rho = torch.empty(N ** 3, dtype=torch.float64, device="cuda")
    avg = torch.empty(N * N, dtype=torch.float64, device="cuda")

    def observe():
        plan.mode_power_spectrum(x, L, m, a_t)
        cfl.energy_density_f64(N, x, rho, m, 0.0, cfl.inv_ah_sqr(N, L, a_t))
        cfl.axis_average(N, rho, avg)
        plan.power_spectrum(rho, consume=True)  # the observer is done with rho

    observe()
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(reps):
        observe()
    ev1.record(stream)
    torch.cuda.synchronize()
    return {"ms_per_observation": ev0.elapsed_time(ev1) / reps,
            "what": "mode spectrum (2 D2Z + binning) + energy density + density spectrum (D2Z + binning) "
                    "+ density axis average, FP64"}


def precision_workload(args):
    """Run A (FP32 hi + FP32 compensation; --run-a f32s: float stencil arithmetic with a
    double increment, f32c: FP64 arithmetic) and run B (FP64) of the N=512 FRW workload from the same initial state (seed 2, P(k) ~ exp(-k^2/(2*0.25^2)),
    rms 1, dt 0.05) for --steps steps (1000 for the full study). Every
    --report-every steps: relative L2 error of phi and dphi/dt of A against B, and
    the largest relative error over non-empty bins of the field (mode) spectrum
    and of the energy-density spectrum. Both runs use the temporally blocked step
    kernels (bit-identical to the per-stage kernels)."""
    import torch
    import cfl_amd as cfl
    N, L, m, dt, t0 = args.n, float(args.n), 1.0, 0.05, 1.0
    cosmo = cfl.radiation(1.0, 0.5, 1.0)
    p = cfl.kg_params(N, L, m, 0.0, cosmo)
    plan = cfl.SpectrumPlan(N)
    xB = plan.init_grf_state(2, cfl.CF_SPEC_GAUSS, (0.25,), L, 1.0)
    hi, lo = cfl.split_f64(xB)
    wA = cfl.alloc_tb_work(N, cfl.CF_PREC_F32C)
    wB = cfl.alloc_tb_work(N, cfl.CF_PREC_F64)
    run_a = cfl.rk4_run_tb_f32s if args.run_a == "f32s" else cfl.rk4_run_tb_f32c
    rho = torch.empty(N ** 3, dtype=torch.float64, device="cuda")
    n3 = N ** 3

    def spectra(x, t):
        a_t = cfl.cosmo_a(cosmo, t)
        field = plan.mode_power_spectrum(x, L, m, a_t).clone()
        cfl.energy_density_f64(N, x, rho, m, 0.0, cfl.inv_ah_sqr(N, L, a_t))
        return field, plan.power_spectrum(rho).clone()

    def rel_bins(a, b):
        """(sup-norm relative error, largest per-bin relative error over the bins
        within 6 decades of the peak). Bins far below the peak (the DC bin of a
        field with P(0) = 0, the exp(-k^2/2sigma^2) tail near Nyquist at ~1e-35 of
        the peak) hold only round-off of the FP32 stage values and are not
        resolved by either run."""
        d = (a - b).abs()
        peak = b.abs().max()
        sig = b.abs() >= 1e-6 * peak
        return float(d.max() / peak), float((d[sig] / b[sig].abs()).max())

    rows = []
    done = 0
    w0 = time.perf_counter()
    while done < args.steps:
        k = min(args.report_every, args.steps - done)
        run_a(p, hi, lo, t0, done, k, dt, wA)
        cfl.rk4_run_tb_f64(p, xB, t0, done, k, dt, wB)
        done += k
        t = t0 + done * dt
        xA = cfl.join_f64(hi, lo)
        row = {"step": done, "t": t}
        for name, sl in (("phi", slice(0, n3)), ("dphi", slice(n3, 2 * n3))):
            row[f"state_rel_l2_{name}"] = float(torch.linalg.norm(xA[sl] - xB[sl]) / torch.linalg.norm(xB[sl]))
        fA, rA = spectra(xA, t)
        fB, rB = spectra(xB, t)
        row["field_spectrum_sup_rel"], row["field_spectrum_max_rel_6dec"] = rel_bins(fA, fB)
        row["rho_spectrum_sup_rel"], row["rho_spectrum_max_rel_6dec"] = rel_bins(rA, rB)
        del xA
        rows.append(row)
        print(json.dumps({"progress": row}), file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    last = rows[-1]
    worst = max(last["field_spectrum_max_rel_6dec"], last["rho_spectrum_max_rel_6dec"])
    line = {
        "metric": f"relative error of FP32-comp run A ({args.run_a}) vs FP64 (run B), N=512^3 FRW",
        "value": worst,
        "unit": "largest per-bin relative spectrum error (bins within 6 decades of the peak), last step",
        "higher_is_better": False,
        "n_gpus": 1, "steps": args.steps,
        "dtype": "A: f32 stencil, f64 increment; B: f64" if args.run_a == "f32s" else "f64",
        "storage": "A: f32 hi + f32 compensation; B: f64",
        "data": "synthetic (Gaussian random field, seed 2)",
        "config": {"workload": f"KG-FRW radiation era N={N}^3 L={L:g} m=1 dt=0.05 RK4, A vs B every "
                               f"{args.report_every} steps", "wall_s": time.perf_counter() - w0},
        "gate_final_spectrum_1e-6": bool(worst <= 1e-6),
        "per_report": rows,
    }
    print(json.dumps(line), flush=True)


def adjoint_workload(args):
    """Loss L = sum_s w_s P_rho(s, t_end), w = 1/(1+k^2), and its gradient with
    respect to phi(t0) and m through 1024 FP64 RK4 steps of KG-FRW at N=128, L=128,
    m=1, dt=0.05 (seed 3, rms 0.3), checkpoints every 32 steps: the forward pass,
    the loss adjoint and the checkpointed backward pass (every segment but the last
    recomputed from its checkpoint).
    Reported as wall time per gradient and as forward-equivalent site-steps/s."""
    import torch
    import cfl_amd as cfl
    from cfl_amd.adjoint import AdjointProblem
    N, L, m, t0, dt, K, ck = 128, 128.0, 1.0, 1.0, 0.05, 1024, 32
    plan = cfl.SpectrumPlan(N)
    phi0 = plan.init_grf_state(3, cfl.CF_SPEC_GAUSS, (0.5,), L, 0.3)[: N ** 3].clone()
    del plan
    prob = AdjointProblem(N, L, 0.0, cfl.radiation(1.0, 0.5, 1.0), t0, dt, K, ck)
    prob.loss_and_grads(phi0, m)  # warm-up (plans, first launches)
    torch.cuda.synchronize()
    reps = max(1, args.steps // 100)
    w0 = time.perf_counter()
    for _ in range(reps):
        Lv, g, dm = prob.loss_and_grads(phi0, m)
    torch.cuda.synchronize()
    sec = (time.perf_counter() - w0) / reps
    w1 = time.perf_counter()
    prob.loss(phi0, m)
    torch.cuda.synchronize()
    fwd = time.perf_counter() - w1
    line = {
        "metric": "loss + gradient wall time, spectral loss through 1024 RK4 steps, N=128^3 FP64",
        "value": sec, "unit": "s per gradient", "higher_is_better": False, "n_gpus": 1, "dtype": "f64",
        "data": "synthetic (Gaussian random field, seed 3)",
        "config": {"workload": f"KG-FRW N={N}^3 L={L:g} m=1 dt=0.05, {K} steps, checkpoints every {ck}",
                   "repetitions": reps},
        "forward_only_s": fwd, "gradient_over_forward": sec / fwd,
        "forward_equivalent_gsite_steps": N ** 3 * K / sec / 1e9,
        "loss": Lv, "dL_dm": dm,
    }
    print(json.dumps(line), flush=True)


def pacing_workload(args):
    """The single-precision-stencil step (f32s) of the N=512 FRW workload at each pace of
    --paces and at the tuner's pick, --steps timed steps per pace after 4 untimed ones,
    two passes over the list. Paced and unpaced launches are different kernels
    (k_rk4_tb_sa<LAM, true, ...> / <LAM, false, ...>), so a profiler separates them."""
    import torch
    import cfl_amd as cfl
    N, L, dt, t0 = args.n, float(args.n), 0.05, 1.0
    p = cfl.kg_params(N, L, 1.0, 0.0, cfl.radiation(1.0, 0.5, 1.0))
    plan = cfl.SpectrumPlan(N)
    x0 = plan.init_grf_state(2, cfl.CF_SPEC_GAUSS, (0.25,), L, 1.0)
    del plan
    settle(cfl, p, x0)
    hi, lo = cfl.split_f64(x0)
    del x0
    bufs = [(hi, lo), tuple(torch.empty(2 * N ** 3, dtype=torch.float32, device="cuda") for _ in range(2))]
    work = cfl.alloc_tb_work(N, cfl.CF_PREC_F32S)
    t_tune = time.perf_counter()
    tuned = cfl.rk4_tb_tune_pace(p, hi, lo, t0, dt, work, args.chunk)
    tune_s = time.perf_counter() - t_tune
    chunk = args.chunk if args.chunk > 0 else cfl.rk4_tb_get_chunk(N)  # the tuner's pick for 0
    del work
    paces = [int(q) for q in args.paces.split(",") if q] + [tuned]
    stream = torch.cuda.current_stream()
    rates = {}
    for _ in range(2):
        for pace in paces:
            cfl.rk4_tb_set_pace(N, pace, args.chunk)
            for k in range(4):
                i, o = bufs[k & 1], bufs[(k + 1) & 1]
                cfl.rk4_step_tb_f32s(p, i[0], i[1], o[0], o[1], t0 + k * dt, dt, args.chunk)
            torch.cuda.synchronize()
            ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            ev0.record(stream)
            for k in range(args.steps):
                i, o = bufs[k & 1], bufs[(k + 1) & 1]
                cfl.rk4_step_tb_f32s(p, i[0], i[1], o[0], o[1], t0 + k * dt, dt, args.chunk)
            ev1.record(stream)
            torch.cuda.synchronize()
            ms = ev0.elapsed_time(ev1) / args.steps
            rates.setdefault(str(pace), []).append(N ** 3 / (ms * 1e-3) / 1e9)
    cfl.rk4_tb_set_pace(N, tuned, args.chunk)
    best = max(rates, key=lambda q: max(rates[q]))
    line = {
        "metric": "lattice site-steps per second of the f32s step by pace",
        "value": max(rates[str(tuned)]), "unit": "GSite*RK4-step/s at the tuned pace", "higher_is_better": True,
        "chunk": chunk,
        "n_gpus": 1, "steps": args.steps, "dtype": "f32 stencil, f64 increment",
        "data": "synthetic (Gaussian random field, seed 2)",
        "config": {"workload": f"KG-FRW radiation era N={N}^3 L={L:g} m=1 dt=0.05 RK4, f32s step",
                   "pace_unit": "ticks of the 100 MHz real-time counter per plane iteration"},
        "tuned_pace": tuned, "tune_s": tune_s, "best_pace": int(best),
        "rates": {q: [round(r, 2) for r in v] for q, v in rates.items()},
    }
    print(json.dumps(line), flush=True)


def slab_workload(args):
    """One GPU's share of a slab decomposition along a: the blocked step (compensated
    storage in both forms, and FP64 beside it) on 64 of the 512 planes of the headline
    lattice (KG-FRW N=512, eight GPUs' worth of slabs), with four ghost planes on each
    side taken from the whole lattice once (their refresh between steps is the exchange,
    not timed here). One launch per step (one chunk of 64 planes per tile), ping-ponging
    two slab buffers; HIP events on the launch stream. The float-force form is paced at
    the pace the tuner picks for the whole lattice with that chunk."""
    import torch
    import cfl_amd as cfl
    N, G = args.n, 8
    n = N // G
    L, dt, t0 = float(N), 0.05, 1.0
    p = cfl.kg_params(N, L, 1.0, 0.0, cfl.radiation(1.0, 0.5, 1.0))
    plan = cfl.SpectrumPlan(N)
    x0 = plan.init_grf_state(2, cfl.CF_SPEC_GAUSS, (0.25,), L, 1.0)
    del plan
    idx = torch.tensor([(i - 4) % N for i in range(n + 8)], device="cuda")
    slab = torch.cat([x0[: N ** 3].view(N, N * N)[idx].reshape(-1), x0[N ** 3:].view(N, N * N)[idx].reshape(-1)])
    settle(cfl, p, x0)
    hi_w, lo_w = cfl.split_f64(x0)
    del x0
    tw = cfl.alloc_tb_work(N, cfl.CF_PREC_F32S)
    pace = cfl.rk4_tb_tune_pace(p, hi_w, lo_w, t0, dt, tw, n)  # kept for (N, chunk n)
    del hi_w, lo_w, tw
    stream = torch.cuda.current_stream()
    sites = N * N * n
    paths = {}
    for mode in ("tb_f32s", "tb_f32c", "tb_f64"):
        if mode == "tb_f64":
            bufs = [(slab,), (slab.clone(),)]
            step = lambda k, i, o: cfl.rk4_step_tb_f64_slab(p, n, i[0], o[0], t0 + k * dt, dt, n)  # noqa: E731
        else:
            bufs = [cfl.split_f64(slab), tuple(torch.empty(slab.numel(), dtype=torch.float32, device="cuda")
                                               for _ in range(2))]
            f = cfl.rk4_step_tb_f32c_slab if mode == "tb_f32c" else cfl.rk4_step_tb_f32s_slab
            step = (lambda f: lambda k, i, o: f(p, n, i[0], i[1], o[0], o[1], t0 + k * dt, dt, n))(f)  # noqa: E731
        if mode != "tb_f64":  # the ghost planes of the second buffer (never written by a step)
            for dst, src in zip(bufs[1], bufs[0]):
                dst.copy_(src)
        for k in range(args.warmup):
            step(k, bufs[k & 1], bufs[(k + 1) & 1])
        torch.cuda.synchronize()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        for k in range(args.warmup, args.warmup + args.steps):
            step(k, bufs[k & 1], bufs[(k + 1) & 1])
        ev1.record(stream)
        torch.cuda.synchronize()
        ms = ev0.elapsed_time(ev1) / args.steps
        paths[mode] = {"value": sites / (ms * 1e-3) / 1e9, "ms_per_step": ms}
        bufs = step = None
    head = paths["tb_f32s"]
    line = {
        "metric": "GSite*RK4-step/s of one GPU's slab (temporally blocked; FP32-comp, FP64 beside it)",
        "value": head["value"], "unit": "GSite*RK4-step/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": head["ms_per_step"], "higher_is_better": True,
        "dtype": "f32 stencil, f64 increment", "pace_ticks_per_plane": pace,
        "storage": "f32 hi + f32 compensation", "data": "synthetic (Gaussian random field, seed 2)",
        "config": {"workload": f"KG-FRW N={N}^3 L={L:g} m=1 dt=0.05, slab of {n} planes + 2 x 4 ghost planes "
                               f"(one of {G} slabs along a)", "slab_planes": n},
        "paths": paths,
        "ghost_bytes_per_neighbour_per_step": {"tb_f32s": 4 * N * N * 16, "tb_f32c": 4 * N * N * 16,
                                               "tb_f64": 4 * N * N * 16},
    }
    print(json.dumps(line), flush=True)


def variants_workload(args):
    """FP64 RK4 steps of the fixed-curvature (Psi: a Gaussian random field, rms 0.1)
    and sqrt-potential (f_a = 0.7) equations at N=512, L=512, m=1.1, radiation FRW,
    dt=0.01, from the headline's initial state, and of the comoving-curvature
    equations at the reference main run's size: the four fused stage kernels per step
    against the temporally blocked kernel (same bits), HIP events on the launch stream."""
    import torch
    import cfl_amd as cfl
    N = args.n
    p = cfl.kg_params(N, float(N), 1.1, 0.0, cfl.radiation(a1=1.0, H1=0.5, t1=1.0))
    plan = cfl.SpectrumPlan(N)
    x0 = plan.init_grf_state(2, cfl.CF_SPEC_GAUSS, (0.25,), float(N), 1.0)
    psi = plan.init_grf_state(3, cfl.CF_SPEC_GAUSS, (0.25,), float(N), 0.1)[: N ** 3].clone()
    del plan
    settle(cfl, p, x0)
    K, W = args.steps, max(1, args.warmup)
    out = {}
    for name, kind, field in (("fixed_curvature", cfl.CF_VAR_FIXED_CURVATURE, psi),
                              ("sqrt_potential", cfl.CF_VAR_SQRT_POTENTIAL, None)):
        v = cfl.Variant(p, kind, field, f_a=0.7)
        for sched in ("stage_kernels", "tb"):
            x = x0.clone()
            run = v.rk4_run if sched == "stage_kernels" else v.rk4_run_tb
            run(x, 1.0, 0, W, 0.01)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            run(x, 1.0, W, K, 0.01)
            e1.record()
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / K
            out[f"{name}_{sched}"] = {"ms_per_step": ms, "gsite_steps_per_s": N ** 3 / ms / 1e6}
            del x
        v.close()
    # the reference main run's comoving-curvature equations (N=384, L=307.2, m=1,
    # radiation era a1=1, H1=0.05, t1=10, dt=0.5; coarse grid M=128), R a Gaussian
    # random field of rms 0.02
    Nc, Lc = 384, 307.2
    pc = cfl.kg_params(Nc, Lc, 1.0, 0.0, cfl.radiation(a1=1.0, H1=0.05, t1=10.0))
    plan = cfl.SpectrumPlan(Nc)
    Rk = plan.d2z(plan.init_grf_state(5, cfl.CF_SPEC_GAUSS, (0.5,), Lc, 0.02)[: Nc ** 3].clone())
    xc0 = plan.init_grf_state(6, cfl.CF_SPEC_GAUSS, (0.5,), Lc, 1.0)
    del plan
    for name, kind, M in (("comoving_curvature_N384", cfl.CF_VAR_COMOVING_CURVATURE, 0),
                          ("approx_comoving_N384_M128", cfl.CF_VAR_APPROX_COMOVING_CURVATURE, 128)):
        for sched in ("stage_kernels", "tb"):
            v = cfl.Variant(pc, kind, Rk, M=M)
            x = xc0.clone()
            run = v.rk4_run if sched == "stage_kernels" else v.rk4_run_tb
            run(x, 10.0, 0, W, 0.5)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            run(x, 10.0, W, K, 0.5)
            e1.record()
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / K
            out[f"{name}_{sched}"] = {"ms_per_step": ms, "gsite_steps_per_s": Nc ** 3 / ms / 1e6}
            v.close()
            del x
    line = {
        "metric": "lattice site-steps per second, equation variants", "value": out["fixed_curvature_tb"][
            "gsite_steps_per_s"], "unit": "GSite*step/s (fixed curvature, temporally blocked)",
        "higher_is_better": True, "n_gpus": 1, "steps": K, "warmup": W, "dtype": "f64",
        "data": "synthetic (Gaussian random fields, seeds 2 and 3)",
        "config": {"workload": f"fixed-curvature and sqrt-potential FRW equations N={N}^3 L={N} m=1.1 f_a=0.7 "
                               "dt=0.01 RK4; comoving-curvature equations of the reference main run N=384^3 "
                               "L=307.2 dt=0.5 (coarse grid M=128)"},
        "paths": out,
    }
    print(json.dumps(line), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this library runs on one GPU only (asked for --gpus {args.gpus}); the "
              "decomposition over several GPUs (slabs along a, one ghost phi-plane per stage) "
              "is designed, not built", file=sys.stderr)
        sys.exit(2)
    if args.workload == "variants":
        variants_workload(args)
        return
    if args.workload == "adjoint":
        adjoint_workload(args)
        return
    if args.workload == "slab":
        slab_workload(args)
        return
    if args.workload == "pacing":
        pacing_workload(args)
        return
    if args.workload == "precision":
        precision_workload(args)
        return
    if args.workload == "ensemble":
        ensemble_workload(args)
        return
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
    observer = observer_timing(cfl, plan, x0, N, L, m, cosmo, t0, stream)
    del plan
    settle(cfl, p, x0)

    paths = [q for q in args.paths.split(",") if q]
    results = {}
    for mode in paths:
        if mode.startswith("tb_"):
            # one out-of-place launch per step, ping-ponging between two state
            # buffers: the timed region holds exactly `steps` kernel launches
            if mode == "tb_f64":
                bufs = [(x0.clone(),), (torch.empty_like(x0),)]
                step = lambda k, i, o: cfl.rk4_step_tb_f64(p, i[0], o[0], t0 + k * dt, dt, args.chunk)  # noqa: E731
            else:
                bufs = [cfl.split_f64(x0), tuple(torch.empty(2 * N ** 3, dtype=torch.float32, device="cuda")
                                                 for _ in range(2))]
                f = cfl.rk4_step_tb_f32c if mode == "tb_f32c" else cfl.rk4_step_tb_f32s
                step = (lambda f: lambda k, i, o: f(p, i[0], i[1], o[0], o[1], t0 + k * dt, dt,  # noqa: E731
                                                    args.chunk))(f)
                if mode == "tb_f32s":
                    # untimed set-up: the pace of the lock-step for this lattice (the state is
                    # left as it was; the tuner uses its own output buffers)
                    tw = cfl.alloc_tb_work(N, cfl.CF_PREC_F32S)
                    t_tune = time.perf_counter()
                    pace = cfl.rk4_tb_tune_pace(p, bufs[0][0], bufs[0][1], t0, dt, tw, args.chunk)
                    tune_s = time.perf_counter() - t_tune
                    del tw

            def run(k0, n, bufs=bufs, step=step):
                for k in range(k0, k0 + n):
                    step(k, bufs[k & 1], bufs[(k + 1) & 1])
        elif mode == "f64":
            x = x0.clone()
            work = cfl.alloc_work(N, cfl.CF_PREC_F64)
            run = lambda k0, n: cfl.rk4_run_f64(p, x, t0, k0, n, dt, work)  # noqa: E731
        else:
            hi, lo = cfl.split_f64(x0)
            work = cfl.alloc_work(N, cfl.CF_PREC_F32C)
            run = lambda k0, n: cfl.rk4_run_f32c(p, hi, lo, t0, k0, n, dt, work)  # noqa: E731
        run(0, args.warmup)
        torch.cuda.synchronize()
        # HIP events on the stream the kernels are launched on (torch's current stream)
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        w0 = time.perf_counter()
        ev0.record(stream)
        run(args.warmup, args.steps)
        ev1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - w0
        ms = ev0.elapsed_time(ev1) / args.steps
        rate = N ** 3 / (ms * 1e-3) / 1e9
        canon = BYTES_PER_SITE_STEP[mode.replace("tb_", "")]
        kb = DESIGN_BYTES_PER_SITE_STEP[mode]
        kname = KERNEL_NAMES[mode]
        if mode == "tb_f32s":
            rows, cols = cfl.rk4_tb_f32s_region(N)
            kname = kname.format(paced="true" if pace > 0 else "false", rows=rows, cols=cols)
        results[mode] = {"value": rate, "ms_per_step": ms, "wall_s": wall,
                         "kernel": kname,
                         "kernel_bytes_per_site_step": kb,
                         "kernel_frac_of_8TBs": kb * rate / HBM_PEAK_GBS,
                         "canonical_bytes_per_site_step": canon,
                         "canonical_frac_of_8TBs": canon * rate / HBM_PEAK_GBS}
        if mode == "tb_f32s":
            results[mode]["pace_ticks_per_plane"] = pace
            results[mode]["pace_tune_s"] = tune_s
            results[mode]["chunk"] = args.chunk if args.chunk > 0 else cfl.rk4_tb_get_chunk(N)
        print(json.dumps({"path": mode, **results[mode]}), file=sys.stderr, flush=True)
        run = step = bufs = work = None
        torch.cuda.synchronize()

    # Headline: run A (FP32-compensated storage), the fastest of its kernels; the
    # FP64 paths (run B) are reported beside it.
    a_modes = [q for q in ("tb_f32s", "tb_f32c", "f32c") if q in results]
    b_modes = [q for q in ("tb_f64", "f64") if q in results]
    best_a = max(a_modes, key=lambda q: results[q]["value"]) if a_modes else None
    best_b = max(b_modes, key=lambda q: results[q]["value"]) if b_modes else None
    head_mode = best_a if best_a is not None else max(results, key=lambda q: results[q]["value"])
    head = results[head_mode]
    # roofline of the headline kernel: the bytes it must move per site-step (state
    # read once and written once for the temporally blocked step) over its average
    # launch time (HIP events around `steps` launches on the launch stream)
    kb = head["kernel_bytes_per_site_step"]
    achieved = kb * N ** 3 / (head["ms_per_step"] * 1e-3) / 1e9
    canon = head["canonical_bytes_per_site_step"]
    traffic = traffic_per_launch(head["kernel"], N)
    valu = valu_busy(head["kernel"], N)
    line = {
        "metric": "GSite*RK4-step/s + % HBM roofline, FP32-comp vs FP64 at N=512^3",
        "value": head["value"],
        "unit": "GSite*RK4-step/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 stencil, f64 increment" if head_mode == "tb_f32s" else "f64",
        "storage": "f32 hi + f32 compensation" if ("f32c" in head_mode or "f32s" in head_mode) else "f64",
        "data": "synthetic (Gaussian random field, seed 2)",
        "config": {"workload": f"KG-FRW radiation era N={N}^3 L={L:g} m=1 dt=0.05 RK4 "
                               f"(run A FP32-comp, run B FP64)", "N": N, "L": L, "dt": dt,
                   "headline_path": head_mode, "parallelism": "single GPU"},
        "paths": results,
        "observer": observer,
        "roofline": {"bound": "hbm", "kernel": head["kernel"],
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic,
                     "algorithmic_bytes_per_site_step": kb,
                     "traffic_bytes_per_site_step": traffic / N ** 3 if traffic else None,
                     "traffic_frac": traffic / (head["ms_per_step"] * 1e-3) / 1e9 / HBM_PEAK_GBS if traffic
                     else None,
                     "canonical_bytes_per_site_step": canon,
                     "canonical_frac": canon * head["value"] / HBM_PEAK_GBS,
                     "limiter": LIMITER.get(head_mode, "hbm"),
                     "valu_busy": valu,
                     "note": "achieved = the bytes the kernel must move per site-step (read and write the "
                             "state once) x sites / average launch time; traffic = rocprofv3 FETCH_SIZE + "
                             "WRITE_SIZE per launch (calibrated, profiles/); canonical_frac uses the "
                             "one-kernel-per-stage schedule's bytes (152 B FP32-comp, 272 B FP64) and can "
                             "exceed 1 for a temporally blocked step; valu_busy = SQ_ACTIVE_INST_VALU / "
                             "(CUs x kernel cycles) from profiles/"},
    }
    if best_a and best_b:
        line["speedup_run_a_vs_fastest_f64"] = results[best_a]["value"] / results[best_b]["value"]
    for a, b in (("tb_f32s", "tb_f64"), ("tb_f32c", "tb_f64"), ("f32c", "f64")):
        if a in results and b in results:
            line[f"speedup_{a}_vs_{b}"] = results[a]["value"] / results[b]["value"]
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(N, args.cpu_steps)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
