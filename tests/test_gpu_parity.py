"""GPU parity of the HIP kernels against the oracle (CPU restatement).

FP64 state evolution, the right-hand side and the energy density are compared
bit for bit: the kernels evaluate the reference expressions in the same order
with no FMA contraction. Spectra go through hipFFT, whose rounding differs from
the oracle's FFT, so they are compared to 1e-10 relative over bins with nonzero
multiplicity, with zero-multiplicity bins exactly 0 on both sides.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

SPEC_TOL = 1e-10


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def rand_state(N, seed, amp=1.0):
    rng = np.random.default_rng(seed)
    phi = amp * rng.standard_normal(N ** 3)
    pi = 0.3 * amp * rng.standard_normal(N ** 3)
    return np.concatenate([phi, pi])


def assert_bitwise(a, b, what):
    if not np.array_equal(a, b):
        d = np.abs(a - b)
        i = int(np.argmax(d))
        raise AssertionError(f"{what}: {np.count_nonzero(a != b)} of {a.size} differ; worst at {i}: "
                             f"{a[i]!r} vs {b[i]!r}")


# Absolute floor of the spectrum comparison, relative to the largest bin: bins
# whose true power is below ~1e-18 of the peak (e.g. the DC bin of a Gaussian
# random field, zero by construction) hold only FFT roundoff in FP64.
SPEC_FLOOR = 1e-18


def spec_close(gpu, ref, what, tol=SPEC_TOL):
    """|gpu - ref| <= tol*|ref| + SPEC_FLOOR*max|ref| per bin; bins of zero
    multiplicity (ref exactly 0) must be exactly 0."""
    assert gpu.shape == ref.shape, what
    zero = ref == 0
    assert np.all(gpu[zero] == 0), f"{what}: zero-multiplicity bins not zero"
    err = np.abs(gpu - ref)
    bound = tol * np.abs(ref) + SPEC_FLOOR * np.abs(ref).max()
    bad = err > bound
    assert not bad.any(), (f"{what}: {bad.sum()} bins out of tolerance; worst rel "
                           f"{(err / np.where(zero, 1, np.abs(ref))).max():.3e}")


FRW = dict(a1=1.0, H1=0.5, t1=1.0)


@pytest.mark.parametrize("N,lam", [(4, 0.0), (8, 0.0), (16, 0.1), (64, 0.0), (96, 0.1)])
def test_rhs_bitwise(cfl, O, N, lam):
    print(f"case rhs N={N} lam={lam}", flush=True)
    x = rand_state(N, 1)
    m, H, inv = 1.3, 0.37, 0.81
    ref = O.rhs(N, x, m, lam, H, inv)
    xd = dev(x)
    out = torch.empty_like(xd)
    cfl.rhs_f64(N, xd, out, m, lam, H, inv)
    assert_bitwise(host(out), ref, "rhs")


@pytest.mark.parametrize("N,lam", [(8, 0.0), (64, 0.0), (96, 0.1)])
def test_energy_density_bitwise(cfl, O, N, lam):
    print(f"case rho N={N} lam={lam}", flush=True)
    x = rand_state(N, 2)
    ref = O.energy_density(N, x, 0.9, lam, 0.44)
    rho = torch.empty(N ** 3, dtype=torch.float64, device="cuda")
    cfl.energy_density_f64(N, dev(x), rho, 0.9, lam, 0.44)
    assert_bitwise(host(rho), ref, "rho")


@pytest.mark.parametrize("N,L,lam,frw", [(8, 8.0, 0.0, False), (16, 20.0, 0.1, True),
                                         (64, 64.0, 0.0, True), (96, 96.0, 0.1, True)])
def test_rk4_f64_bitwise(cfl, O, N, L, lam, frw):
    print(f"case rk4 f64 N={N} lam={lam} frw={frw}", flush=True)
    cosmo_o = O.Cosmo.radiation(**FRW) if frw else O.Cosmo.minkowski()
    cosmo_g = cfl.radiation(**FRW) if frw else cfl.minkowski()
    eq = O.Equation(N, L, 1.0, lam, cosmo_o)
    x = rand_state(N, 3, amp=0.5)
    xd = dev(x)
    p = cfl.kg_params(N, L, 1.0, lam, cosmo_g)
    work = cfl.alloc_work(N)
    t0, dt, steps = 1.0, 0.05, 5
    eq.rk4_run(x, t0, dt, steps)
    cfl.rk4_run_f64(p, xd, t0, 0, steps, dt, work)
    assert_bitwise(host(xd), x, "rk4 state")


def test_rk4_f64_matches_unfused_odeint_flow(cfl, O):
    """The fused stage kernels equal sys() + scale_sum2/scale_sum5 on the GPU."""
    N, L, dt, t = 32, 32.0, 0.1, 1.0
    cosmo = cfl.radiation(**FRW)
    x = dev(rand_state(N, 4))
    x_fused = x.clone()
    p = cfl.kg_params(N, L, 1.0, 0.0, cosmo)
    cfl.rk4_step_f64(p, x_fused, t, dt, cfl.alloc_work(N))
    k = [torch.empty_like(x) for _ in range(4)]
    y = torch.empty_like(x)
    dh = dt / 2

    def f(xx, tt, out):
        a = cfl.cosmo_a(cosmo, tt)
        cfl.rhs_f64(N, xx, out, 1.0, 0.0, cfl.cosmo_H(cosmo, tt), cfl.inv_ah_sqr(N, L, a))

    f(x, t, k[0])
    cfl.scale_sum(y, [x, k[0]], [1.0, dh])
    f(y, t + dh, k[1])
    cfl.scale_sum(y, [x, k[1]], [1.0, dh])
    f(y, t + dh, k[2])
    cfl.scale_sum(y, [x, k[2]], [1.0, dt])
    f(y, t + dt, k[3])
    out = torch.empty_like(x)
    cfl.scale_sum(out, [x] + k, [1.0, dt / 6, dt / 3, dt / 3, dt / 6])
    assert_bitwise(host(x_fused), host(out), "fused vs unfused")


@pytest.mark.parametrize("N,lam", [(16, 0.0), (64, 0.1)])
def test_rk4_f32c_bitwise(cfl, O, N, lam):
    print(f"case rk4 f32c N={N} lam={lam}", flush=True)
    eq = O.Equation(N, float(N), 1.0, lam, O.Cosmo.radiation(**FRW))
    x = rand_state(N, 5, amp=0.5)
    hi = x.astype(np.float32)
    lo = (x - hi.astype(np.float64)).astype(np.float32)
    p = cfl.kg_params(N, float(N), 1.0, lam, cfl.radiation(**FRW))
    hd, ld = dev(hi.copy()), dev(lo.copy())
    work = cfl.alloc_work(N, cfl.CF_PREC_F32C)
    t0, dt = 1.0, 0.05
    for k in range(4):
        t = t0 + k * dt
        eq.rk4_step_f32c(hi, lo, t, dt)
        cfl.rk4_step_f32c(p, hd, ld, t, dt, work)
    assert_bitwise(host(hd), hi, "f32c hi")
    assert_bitwise(host(ld), lo, "f32c lo")


def test_split_join(cfl):
    x = rand_state(16, 6)
    hi, lo = cfl.split_f64(dev(x))
    back = host(cfl.join_f64(hi, lo))
    rel = np.max(np.abs(back - x) / np.abs(x))
    assert rel < 2.0 ** -46


@pytest.mark.parametrize("N", [4, 6, 8, 16, 64, 96])
def test_power_spectrum(cfl, O, N):
    print(f"case power spectrum N={N}", flush=True)
    f = np.random.default_rng(7).standard_normal(N ** 3)
    plan = cfl.SpectrumPlan(N)
    fd = dev(f)
    g = host(plan.power_spectrum(fd))
    spec_close(g, O.power_spectrum(N, f), "power spectrum")
    assert_bitwise(host(fd), f, "input preserved")
    # deterministic binning: a second call gives identical bits
    assert_bitwise(host(plan.power_spectrum(fd)), g, "repeat")
    # the consuming form (no input copy) gives the same bins
    assert_bitwise(host(plan.power_spectrum(fd, consume=True)), g, "consume")


@pytest.mark.parametrize("N", [6, 8, 64, 96])
def test_mode_power_spectrum(cfl, O, N):
    print(f"case mode spectrum N={N}", flush=True)
    x = rand_state(N, 8)
    L, m, a_t = 1.7 * N, 0.8, 1.9
    plan = cfl.SpectrumPlan(N)
    g = host(plan.mode_power_spectrum(dev(x), L, m, a_t))
    spec_close(g, O.mode_power_spectrum(N, L, m, a_t, x), "mode spectrum")


@pytest.mark.parametrize("N", [16, 64])
def test_binning_on_identical_fft(cfl, O, N):
    """Given the same half spectrum, the binning kernels match the reference loop
    (fdm3d.cpp:11-28) to rounding of the summation order."""
    rng = np.random.default_rng(9)
    fk = rng.standard_normal(2 * N * N * (N // 2 + 1))
    dk = rng.standard_normal(fk.size)
    plan = cfl.SpectrumPlan(N)
    spec_close(host(plan.bin_power(dev(fk))), O.bin_power(N, fk), "bin_power", 1e-13)
    L, m, a_t = 33.0, 1.1, 2.0
    ref = np.empty(O.nbins(N))
    import ctypes as C
    O.lib().orc_bin_mode_power(N, L, m, a_t, O._p(fk), O._p(dk), O._p(ref))
    spec_close(host(plan.bin_mode_power(dev(fk), dev(dk), L, m, a_t)), ref, "bin_mode_power", 1e-13)


def test_radial_bins(cfl, O):
    N = 32
    plan = cfl.SpectrumPlan(N)
    sb = np.random.default_rng(10).random(O.nbins(N))
    g = host(plan.radial_bins(dev(sb)))
    np.testing.assert_allclose(g, O.radial_bins(N, sb), rtol=1e-14)


@pytest.mark.parametrize("N,kind,params", [(16, 0, (0.5,)), (64, 1, ()), (96, 0, (0.5,))])
def test_grf_matches_oracle(cfl, O, N, kind, params):
    print(f"case grf N={N} kind={kind}", flush=True)
    L, seed, rms = float(N), 11, 0.3
    ref = O.initial_state(N, L, seed, kind, params, rms)
    plan = cfl.SpectrumPlan(N)
    g = host(plan.init_grf_state(seed, kind, params, L, rms))
    assert np.all(g[N ** 3:] == 0)
    err = np.max(np.abs(g[: N ** 3] - ref[: N ** 3])) / np.max(np.abs(ref[: N ** 3]))
    assert err < 1e-13, err


def test_scale_sum_ops(cfl):
    rng = np.random.default_rng(12)
    n = 1000
    vs = [rng.standard_normal(n) for _ in range(14)]
    al = rng.standard_normal(14)
    for K in (1, 2, 5, 14):
        out = torch.empty(n, dtype=torch.float64, device="cuda")
        cfl.scale_sum(out, [dev(v) for v in vs[:K]], list(al[:K]))
        ref = al[0] * vs[0]
        for k in range(1, K):
            ref = ref + al[k] * vs[k]
        assert_bitwise(host(out), ref, f"scale_sum{K}")
    assert cfl.norm_inf(dev(vs[3])) == np.max(np.abs(vs[3]))
