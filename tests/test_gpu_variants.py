"""Curvature / potential variants (cf_variant_*) against the oracle.

Bars: the sqrt-potential right-hand side and its RK4 steps use only IEEE
operations (+, -, *, /, sqrt) and must match the oracle bit for bit; the cut-off
spectrum is an exact scaling and must match bit for bit. The other variants call
exp (and the comoving ones sin/cos and an FFT), whose last bit differs between
the device math library and glibc, so they are held to |gpu - oracle| <=
1e-12 * max|oracle| per element (state, right-hand side, Psi) -- a dropped term,
a wrong sign or a swapped neighbour moves them by O(1e-2) or more. The fused RK4
step must equal the same library's unfused right-hand side + scale_sum ops bit
for bit (same device exp on both sides).
"""
import numpy as np
import pytest
import torch

from test_gpu_parity import assert_bitwise, dev, host, rand_state

pytestmark = pytest.mark.gpu

FRW = dict(a1=1.0, H1=0.5, t1=1.0)
TOL = 1e-12


def close(g, r, what, tol=TOL):
    err = np.abs(g - r).max()
    scale = np.abs(r).max()
    assert err <= tol * scale, f"{what}: max err {err:.3e} vs scale {scale:.3e} (rel {err / scale:.2e})"


def r_spectrum(O, N, L, seed=5, amp=0.05):
    """A comoving-curvature half spectrum: FFT of a smooth Gaussian random field."""
    R = O.grf_from_normals(N, O.normals(seed, N ** 3), O.spectrum_table(O_SPEC_GAUSS, N, L, (0.8,)))
    R *= amp / np.sqrt(np.mean(R * R))
    return O.fft_r2c(N, R)


O_SPEC_GAUSS = 0


def setup(cfl, O, kind, N, L=None, M=0, seed=3):
    L = float(N) if L is None else L
    m = 1.1
    p = cfl.kg_params(N, L, m, 0.0, cfl.radiation(**FRW))
    cosmo = O.Cosmo.radiation(**FRW)
    rng = np.random.default_rng(seed)
    field, f_a = None, 1.0
    if kind == O.FIXED:
        field = 0.2 * rng.standard_normal(N ** 3)
    elif kind == O.SQRT:
        f_a = 0.7
    else:
        field = r_spectrum(O, N, L)
    v = cfl.Variant(p, kind, None if field is None else dev(field), f_a=f_a, M=M)
    ov = O.VariantEquation(kind, N, L, m, cosmo, field=field, f_a=f_a, M=M)
    return v, ov


CASES = [("fixed", 1, 16, 0), ("fixed", 1, 24, 0), ("sqrt", 2, 16, 0), ("sqrt", 2, 40, 0),
         ("comoving", 3, 16, 0), ("comoving", 3, 24, 0), ("approx", 4, 16, 8), ("approx", 4, 24, 8),
         ("approx", 4, 16, 16)]


@pytest.mark.parametrize("name,kind,N,M", CASES)
def test_rhs_vs_oracle(cfl, O, name, kind, N, M):
    print(f"case rhs {name} N={N} M={M}", flush=True)
    v, ov = setup(cfl, O, kind, N, M=M)
    x = rand_state(N, 7, 0.5)
    for t in (1.0, 1.37):
        g = host(v.rhs(dev(x), t))
        r = ov.rhs(x, t)
        if kind == O.SQRT:
            assert_bitwise(g, r, f"{name} rhs t={t}")
        else:
            close(g, r, f"{name} rhs t={t}")
    v.close()


@pytest.mark.parametrize("name,kind,N,M", [c for c in CASES if c[1] in (3, 4)])
def test_psi_vs_oracle(cfl, O, name, kind, N, M):
    v, ov = setup(cfl, O, kind, N, M=M)
    for t in (1.0, 2.5):
        gp, gd = v.psi(t)
        rp, rd = ov.psi(t)
        close(host(gp), rp, f"{name} Psi t={t}")
        close(host(gd), rd, f"{name} dPsi t={t}")
    v.close()


@pytest.mark.parametrize("N", [16, 32])
def test_transfer_vs_oracle(cfl, O, N):
    R = r_spectrum(O, N, float(N))
    f, H = 0.37, 0.41
    gp, gd = cfl.comoving_psi_fft(N, dev(R), f, H)
    rp, rd = O.comoving_psi_fft(N, R, f, H)
    close(host(gp), rp, "Psi^", 1e-14)
    close(host(gd), rd, "dPsi^", 1e-14)


@pytest.mark.parametrize("N,M", [(16, 8), (24, 8), (32, 32)])
def test_cutoff_bitwise(cfl, O, N, M):
    R = r_spectrum(O, N, float(N))
    assert_bitwise(host(cfl.cutoff_fouriers(N, M, dev(R))), O.cutoff_fouriers(N, M, R), "cutoff")


@pytest.mark.parametrize("name,kind,N,M", CASES)
def test_rk4_vs_oracle(cfl, O, name, kind, N, M):
    print(f"case rk4 {name} N={N} M={M}", flush=True)
    v, ov = setup(cfl, O, kind, N, M=M)
    x0 = rand_state(N, 11, 0.3)
    xd = dev(x0)
    t0, dt, n = 1.0, 0.05, 4
    v.rk4_run(xd, t0, 0, n, dt)
    r = ov.rk4_run(x0.copy(), t0, dt, n)
    if kind == O.SQRT:
        assert_bitwise(host(xd), r, f"{name} rk4")
    else:
        close(host(xd), r, f"{name} rk4")
    v.close()


@pytest.mark.parametrize("kind,M", [(1, 0), (2, 0), (3, 0), (4, 8)])
def test_fused_step_equals_unfused(cfl, O, kind, M):
    """One fused step == sys() + scale_sum2/scale_sum5 with the same library."""
    N, dt, t = 16, 0.1, 1.2
    v, _ = setup(cfl, O, kind, N, M=M)
    x = dev(rand_state(N, 4, 0.4))
    xf = x.clone()
    v.rk4_step(xf, t, dt)
    k = [torch.empty_like(x) for _ in range(4)]
    y = torch.empty_like(x)
    dh = dt / 2
    v.rhs(x, t, k[0])
    cfl.scale_sum(y, [x, k[0]], [1.0, dh])
    v.rhs(y, t + dh, k[1])
    cfl.scale_sum(y, [x, k[1]], [1.0, dh])
    v.rhs(y, t + dh, k[2])
    cfl.scale_sum(y, [x, k[2]], [1.0, dt])
    v.rhs(y, t + dt, k[3])
    out = torch.empty_like(x)
    cfl.scale_sum(out, [x] + k, [1.0, dt / 6, dt / 3, dt / 3, dt / 6])
    assert_bitwise(host(xf), host(out), f"kind {kind}: fused vs unfused")
    v.close()


@pytest.mark.parametrize("name,kind,N,M", [c for c in CASES if c[2] == 16])
def test_energy_density_vs_oracle(cfl, O, name, kind, N, M):
    v, ov = setup(cfl, O, kind, N, M=M)
    x = rand_state(N, 9, 0.5)
    for t in (1.0, 1.8):
        g = host(v.energy_density(dev(x), t))
        r = ov.energy_density(x, t)
        if kind == O.SQRT:
            assert_bitwise(g, r, f"{name} rho t={t}")
        else:
            close(g, r, f"{name} rho t={t}")
    v.close()


def test_bad_arguments(cfl):
    p = cfl.kg_params(16, 16.0, 1.0, 0.0, cfl.radiation(**FRW))
    with pytest.raises(cfl.CfError):
        cfl.Variant(p, 9)
    with pytest.raises(cfl.CfError):
        cfl.Variant(p, cfl.CF_VAR_FIXED_CURVATURE)  # needs Psi
    R = torch.zeros(2 * 16 * 16 * 9, dtype=torch.float64, device="cuda")
    with pytest.raises(cfl.CfError):
        cfl.Variant(p, cfl.CF_VAR_APPROX_COMOVING_CURVATURE, R, M=6)  # M must divide N
    with pytest.raises(cfl.CfError):
        cfl.Variant(p, cfl.CF_VAR_SQRT_POTENTIAL, f_a=0.0)


@pytest.mark.parametrize("name,kind,N,steps,chunk", [("fixed", 1, 16, 3, 0), ("fixed", 1, 40, 2, 7),
                                                      ("fixed", 1, 64, 1, 0), ("sqrt", 2, 16, 3, 0),
                                                      ("sqrt", 2, 40, 2, 5), ("sqrt", 2, 96, 1, 0)])
def test_rk4_tb_equals_stage_kernels(cfl, O, name, kind, N, steps, chunk):
    """The temporally blocked variant step gives the four-kernel step's bits (ragged
    tiles at N=40, a chunk that does not divide N, odd and even step counts); the
    sqrt potential's also the oracle's."""
    print(f"case rk4_tb {name} N={N} steps={steps} chunk={chunk}", flush=True)
    v, ov = setup(cfl, O, kind, N)
    x0 = rand_state(N, 12, 0.3)
    xa, xb = dev(x0), dev(x0)
    t0, dt = 1.0, 0.05
    v.rk4_run(xa, t0, 2, steps, dt)
    v.rk4_run_tb(xb, t0, 2, steps, dt, chunk=chunk)
    assert_bitwise(host(xb), host(xa), f"{name} tb vs stage kernels")
    xc, xo = dev(x0), torch.empty_like(xa)
    for s in range(steps):  # the out-of-place single step, buffers swapped
        v.rk4_step_tb(xc, xo, t0 + float(2 + s) * dt, dt, chunk=chunk)
        xc, xo = xo, xc
    assert_bitwise(host(xc), host(xa), f"{name} single tb steps vs stage kernels")
    if kind == O.SQRT and N <= 40:
        assert_bitwise(host(xb), ov.rk4_run(x0.copy(), t0, dt, steps, k0=2), f"{name} tb vs oracle")
    v.close()


@pytest.mark.parametrize("name,kind,N,M,steps", [("comoving", 3, 16, 0, 3), ("comoving", 3, 40, 0, 2),
                                                  ("approx", 4, 16, 8, 3), ("approx", 4, 48, 16, 2),
                                                  ("approx", 4, 24, 24, 1)])
def test_rk4_tb_comoving_equals_stage_kernels(cfl, O, name, kind, N, M, steps):
    """The temporally blocked comoving steps (Psi at the three stage times from the
    cache, read per stage at the stage's plane) give the stage kernels' bits, and
    stay within the variants' 1e-12 bar of the oracle."""
    print(f"case rk4_tb {name} N={N} M={M} steps={steps}", flush=True)
    v, ov = setup(cfl, O, kind, N, M=M)
    x0 = rand_state(N, 13, 0.3)
    xa, xb = dev(x0), dev(x0)
    t0, dt = 1.0, 0.05
    v.rk4_run(xa, t0, 1, steps, dt)
    v2, _ = setup(cfl, O, kind, N, M=M)  # a fresh Psi cache
    v2.rk4_run_tb(xb, t0, 1, steps, dt)
    assert_bitwise(host(xb), host(xa), f"{name} tb vs stage kernels")
    xc, xo = dev(x0), torch.empty_like(xa)
    for s in range(steps):  # out-of-place single steps, sharing v's cache with rk4_run
        v.rk4_step_tb(xc, xo, t0 + float(1 + s) * dt, dt)
        xc, xo = xo, xc
    assert_bitwise(host(xc), host(xa), f"{name} single tb steps vs stage kernels")
    close(host(xb), ov.rk4_run(x0.copy(), t0, dt, steps, k0=1), f"{name} tb vs oracle")
    v.close()
    v2.close()
