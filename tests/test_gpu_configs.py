"""Two of the workload configurations at their own sizes on the GPU.

* The precision study: KG-FRW (radiation era) at N=512^3, L=512, m=1, GRF initial
  field (seed 2, P(k) ~ exp(-k^2/(2*0.25^2)), rms 1), dt=0.05, 1000 RK4 steps. Run A
  stores phi and dphi/dt as FP32 hi + FP32 compensation term (FP64 arithmetic), run B
  is FP64; both use the temporally blocked step. Every 100 steps the field (mode)
  spectrum and the energy-density spectrum of A must lie within 1e-6 relative of B's,
  and the states within 1e-6 (relative L2). Run A in both compensated forms: FP64
  arithmetic (f32c) and float stencil arithmetic with a double increment (f32s).
* The ensemble: 96 realisations, 32 each at N=64, 96, 128 (L=N, seeds 100-195 in
  that order), FRW Klein-Gordon with m=1 and lambda=0.1, FP32-comp (both arithmetic
  forms), P(k) ~
  exp(-k^2/(2*0.5^2)), rms 0.5, dt=0.05, the autotuned plan and graph replay, for 5
  steps. Every realisation must equal its single-lattice run bit for bit, one per N
  must equal the oracle's compensated step bit for bit, and the per-N average of
  the density spectra must equal the oracle's average to 1e-10 per bin.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from test_gpu_parity import assert_bitwise, host

pytestmark = pytest.mark.gpu

FRW = dict(a1=1.0, H1=0.5, t1=1.0)


def rel_bins(a, b):
    """(sup-norm relative error max|a-b| / max|b|, largest per-bin relative error
    over the bins within 6 decades of the peak). Bins further below the peak (the
    DC bin of a field with P(0) = 0, the exp(-k^2/2 sigma^2) tail near Nyquist at
    ~1e-35 of the peak) hold round-off of FP32 stage values only, which neither run
    resolves."""
    d = (a - b).abs()
    peak = b.abs().max()
    sig = b.abs() >= 1e-6 * peak
    return float(d.max() / peak), float((d[sig] / b[sig].abs()).max())


@pytest.mark.parametrize("mode", ["f32c", "f32s"])
def test_precision_run_a_vs_run_b_n512(cfl, mode):
    N, L, m, dt, t0 = 512, 512.0, 1.0, 0.05, 1.0
    print(f"case precision {mode} N={N} 1000 steps", flush=True)
    run_a = cfl.rk4_run_tb_f32c if mode == "f32c" else cfl.rk4_run_tb_f32s
    cosmo = cfl.radiation(**FRW)
    p = cfl.kg_params(N, L, m, 0.0, cosmo)
    plan = cfl.SpectrumPlan(N)
    xB = plan.init_grf_state(2, cfl.CF_SPEC_GAUSS, (0.25,), L, 1.0)
    hi, lo = cfl.split_f64(xB)
    wA = cfl.alloc_tb_work(N, cfl.CF_PREC_F32C)
    wB = cfl.alloc_tb_work(N, cfl.CF_PREC_F64)
    rho = torch.empty(N ** 3, dtype=torch.float64, device="cuda")
    n3 = N ** 3

    def spectra(x, t):
        a_t = cfl.cosmo_a(cosmo, t)
        field = plan.mode_power_spectrum(x, L, m, a_t).clone()
        cfl.energy_density_f64(N, x, rho, m, 0.0, cfl.inv_ah_sqr(N, L, a_t))
        return field, plan.power_spectrum(rho).clone()

    rows = []
    for done in range(0, 1000, 100):
        run_a(p, hi, lo, t0, done, 100, dt, wA)
        cfl.rk4_run_tb_f64(p, xB, t0, done, 100, dt, wB)
        t = t0 + (done + 100) * dt
        xA = cfl.join_f64(hi, lo)
        row = [done + 100]
        for sl in (slice(0, n3), slice(n3, 2 * n3)):
            row.append(float(torch.linalg.norm(xA[sl] - xB[sl]) / torch.linalg.norm(xB[sl])))
        fA, rA = spectra(xA, t)
        fB, rB = spectra(xB, t)
        row += [*rel_bins(fA, fB), *rel_bins(rA, rB)]
        del xA
        print("step %d: state phi %.2e dphi %.2e; field spectrum sup %.2e bins %.2e; "
              "rho spectrum sup %.2e bins %.2e" % tuple(row), flush=True)
        rows.append(row)
    torch.cuda.synchronize()
    for row in rows:
        assert max(row[1:]) <= 1e-6, f"step {row[0]}: {row[1:]}"
    plan.close()


@pytest.mark.parametrize("mode", ["f32c", "f32s"])
def test_ensemble_96_uneven_n(cfl, O, mode):
    Ns = [64] * 32 + [96] * 32 + [128] * 32
    seeds = list(range(100, 196))
    dt, t0, steps = 0.05, 1.0, 5
    print(f"case ensemble 96 realisations {mode}, 5 steps", flush=True)
    prec = cfl.CF_PREC_F32C if mode == "f32c" else cfl.CF_PREC_F32S
    run_single = cfl.rk4_run_tb_f32c if mode == "f32c" else cfl.rk4_run_tb_f32s
    ens = cfl.Ensemble(Ns, [float(n) for n in Ns], 1.0, 0.1, cfl.radiation(**FRW), prec)
    ens.init_grf(seeds, cfl.CF_SPEC_GAUSS, (0.5,), 0.5)

    def raw(r):
        n = 2 * Ns[r] ** 3
        x, lo = ens.state_ptrs(r)
        hi_t = torch.empty(n, dtype=torch.float32, device="cuda")
        lo_t = torch.empty_like(hi_t)
        cfl.check(cfl.lib.cf_d2d(cfl._ptr(hi_t), C.c_void_p(x), n * 4, cfl._stream()))
        cfl.check(cfl.lib.cf_d2d(cfl._ptr(lo_t), C.c_void_p(lo), n * 4, cfl._stream()))
        return hi_t, lo_t

    init = [raw(r) for r in range(96)]
    ens.tune(dt, reps=2)
    ens.run(t0, 0, steps, dt, use_graph=True)
    torch.cuda.synchronize()

    # every realisation against its single-lattice run (the temporally blocked step)
    works = {n: cfl.alloc_tb_work(n, prec) for n in (64, 96, 128)}
    for r, n in enumerate(Ns):
        hi, lo = init[r]
        if r % 32 == 0:
            hi0, lo0 = host(hi).copy(), host(lo).copy()
        p = cfl.kg_params(n, float(n), 1.0, 0.1, cfl.radiation(**FRW))
        run_single(p, hi, lo, t0, 0, steps, dt, works[n])
        ghi, glo = raw(r)
        assert_bitwise(host(ghi), host(hi), f"hi r={r} N={n}")
        assert_bitwise(host(glo), host(lo), f"lo r={r} N={n}")
        if r % 32 == 0:  # the first realisation of each N against the oracle
            eq = O.Equation(n, float(n), 1.0, 0.1, O.Cosmo.radiation(**FRW))
            oracle_step = eq.rk4_step_f32c if mode == "f32c" else eq.rk4_step_f32s
            for k in range(steps):
                oracle_step(hi0, lo0, t0 + k * dt, dt)
            assert_bitwise(host(ghi), hi0, f"hi r={r} N={n} vs oracle")
            assert_bitwise(host(glo), lo0, f"lo r={r} N={n} vs oracle")
    del works, init

    # ensemble-averaged density spectra per N, fixed order, against the oracle's
    t_end = t0 + steps * dt
    for n in (64, 96, 128):
        g_avg = np.zeros(O.nbins(n))
        o_avg = np.zeros(O.nbins(n))
        eq = O.Equation(n, float(n), 1.0, 0.1, O.Cosmo.radiation(**FRW))
        for r in (i for i, v in enumerate(Ns) if v == n):
            g_avg += host(ens.density_spectrum(r, t_end))
            o_avg += O.power_spectrum(n, eq.energy_density(host(ens.state(r)), t_end))
        g_avg /= 32
        o_avg /= 32
        assert np.all(g_avg[o_avg == 0] == 0), f"N={n}: empty bins"
        ok = np.abs(g_avg - o_avg) <= 1e-10 * o_avg + 1e-18 * o_avg.max()
        assert ok.all(), f"N={n}: {np.count_nonzero(~ok)} bins beyond 1e-10"
    ens.close()
