"""Pins of the oracle's curvature / potential variants (CPU only).

Each check is chosen so that a plausible slip in the restatement fails it:
* Psi = 0 reduces the comoving variant to the Klein-Gordon FRW equation bit for
  bit (same CPU operation order), and its RK4 to the oracle's RK4, which is pinned
  against the reference's own odeint build;
* the approximate variant at M = N and Psi = 0 equals the fixed-curvature variant
  at Psi = 0 bit for bit (same CUDA operation order), and for small Psi it stays
  within the Taylor remainder O(Psi^3) of the exact-exponential variant (a wrong
  Taylor coefficient leaves an O(Psi^2) gap);
* constant Psi is a rescaling of m^2 and of the gradient term;
* the sqrt potential on a homogeneous field has a closed form, and tends to the
  quadratic potential for f_a -> infinity;
* dPsi/dt from its transfer function equals the time derivative of Psi(t)
  (radiation era, where d(k eta/sqrt 3)/dt = x H), which fixes its sign and
  powers of x;
* the cut-off spectrum of a band-limited field is the spectrum of the field
  sampled every N/M sites, which fixes the negative-frequency index map.
"""
import numpy as np
import pytest

FRW = dict(a1=1.0, H1=0.5, t1=1.0)


def rand_state(N, seed, amp=0.5):
    rng = np.random.default_rng(seed)
    return np.concatenate([amp * rng.standard_normal(N ** 3), 0.3 * amp * rng.standard_normal(N ** 3)])


def r_spec(O, N, L, amp, seed=5):
    R = O.grf_from_normals(N, O.normals(seed, N ** 3), O.spectrum_table(0, N, L, (0.8,)))
    R *= amp / np.sqrt(np.mean(R * R))
    return R, O.fft_r2c(N, R)


@pytest.mark.parametrize("N", [8, 12])
def test_comoving_with_zero_curvature_is_kg(O, N):
    cosmo = O.Cosmo.radiation(**FRW)
    L, m = float(N), 1.3
    R0 = np.zeros(2 * N * N * (N // 2 + 1))
    v = O.VariantEquation(O.COMOVING, N, L, m, cosmo, field=R0)
    eq = O.Equation(N, L, m, 0.0, cosmo)
    x = rand_state(N, 1)
    for t in (1.0, 1.7):
        assert np.array_equal(v.rhs(x, t), eq.rhs(x, t))
        assert np.array_equal(v.energy_density(x, t), eq.energy_density(x, t))
    xr = x.copy()
    eq.rk4_run(xr, 1.0, 0.05, 3)
    assert np.array_equal(v.rk4_run(x.copy(), 1.0, 0.05, 3), xr)


def test_approx_at_full_grid_and_zero_curvature_is_fixed(O):
    N = 8
    cosmo = O.Cosmo.radiation(**FRW)
    R0 = np.zeros(2 * N * N * (N // 2 + 1))
    va = O.VariantEquation(O.APPROX, N, 8.0, 0.9, cosmo, field=R0, M=N)
    vf = O.VariantEquation(O.FIXED, N, 8.0, 0.9, cosmo, field=np.zeros(N ** 3))
    x = rand_state(N, 2)
    assert np.array_equal(va.rhs(x, 1.3), vf.rhs(x, 1.3))


def test_approx_taylor_remainder(O):
    """|approx - exact| / |exact force| scales like Psi^3 (a wrong Taylor
    coefficient would leave Psi^2)."""
    N = 12
    cosmo = O.Cosmo.radiation(**FRW)
    x = rand_state(N, 3)
    gaps = []
    for amp in (2e-2, 1e-2):
        _, Rk = r_spec(O, N, float(N), amp)
        ve = O.VariantEquation(O.COMOVING, N, float(N), 1.0, cosmo, field=Rk)
        va = O.VariantEquation(O.APPROX, N, float(N), 1.0, cosmo, field=Rk, M=N)
        t = 1.0
        ps, _ = ve.psi(t)
        fe, fa = ve.rhs(x, t)[N ** 3:], va.rhs(x, t)[N ** 3:]
        gaps.append((np.abs(fe - fa).max() / np.abs(fe).max(), np.abs(ps).max()))
    (g1, p1), (g2, p2) = gaps
    assert g1 < 10 * p1 ** 3 and g2 < 10 * p2 ** 3, gaps
    ratio = g1 / g2
    assert 5.0 < ratio < 11.0, f"remainder ratio {ratio:.2f} for a 2x change of Psi (expected ~8)"


def test_fixed_constant_psi_rescales(O):
    N, m, psi0 = 8, 0.8, 0.1
    cosmo = O.Cosmo.radiation(**FRW)
    v = O.VariantEquation(O.FIXED, N, 8.0, m, cosmo, field=np.full(N ** 3, psi0))
    x = rand_state(N, 4)
    t = 1.4
    H, inv = O.cosmo_H(cosmo, t), O.inv_ah_sqr(N, 8.0, O.cosmo_a(cosmo, t))
    phi = x[: N ** 3].reshape(N, N, N)
    lap = sum(np.roll(phi, 1, d) + np.roll(phi, -1, d) for d in range(3)) - 6 * phi
    want = -3 * H * x[N ** 3:] - np.exp(2 * psi0) * m * m * x[: N ** 3] + np.exp(4 * psi0) * inv * lap.ravel()
    got = v.rhs(x, t)
    assert np.array_equal(got[: N ** 3], x[N ** 3:])
    assert np.abs(got[N ** 3:] - want).max() <= 1e-13 * np.abs(want).max()
    # energy density: e^{-2Psi} pi^2/2 + e^{2Psi} (grad)^2/2 + m^2 phi^2/2 (central differences)
    g2 = sum(((np.roll(phi, -1, d) - np.roll(phi, 1, d)) / 2) ** 2 for d in range(3)).ravel()
    rho = (np.exp(-2 * psi0) * 0.5 * x[N ** 3:] ** 2 + np.exp(2 * psi0) * 0.5 * inv * g2
           + 0.5 * m * m * x[: N ** 3] ** 2)
    assert np.abs(v.energy_density(x, t) - rho).max() <= 1e-13 * rho.max()


def test_sqrt_potential_closed_forms(O):
    N, m, fa = 4, 1.2, 0.7
    cosmo = O.Cosmo.radiation(**FRW)
    v = O.VariantEquation(O.SQRT, N, 4.0, m, cosmo, f_a=fa)
    phi0, pi0 = 0.5, 0.25  # exact binary fractions: the stencil sum is exactly 0
    x = np.concatenate([np.full(N ** 3, phi0), np.full(N ** 3, pi0)])
    t = 1.3
    H = O.cosmo_H(cosmo, t)
    want = (-3.0 * H) * pi0 - (m * m * phi0) / np.sqrt(1.0 + phi0 * phi0 / (fa * fa))
    assert np.all(v.rhs(x, t)[N ** 3:] == want)
    rho = 0.5 * pi0 * pi0 + m * m * fa * fa * (np.sqrt(1 + phi0 ** 2 / fa ** 2) - 1)
    assert np.allclose(v.energy_density(x, t), rho, rtol=1e-15, atol=0)
    # f_a -> infinity: the quadratic potential
    vq = O.VariantEquation(O.SQRT, N, 4.0, m, cosmo, f_a=1e9)
    eq = O.Equation(N, 4.0, m, 0.0, cosmo)
    y = rand_state(N, 5)
    assert np.abs(vq.rhs(y, t) - eq.rhs(y, t)).max() <= 1e-13 * np.abs(eq.rhs(y, t)).max()


@pytest.mark.parametrize("t", [1.0, 2.3])
def test_dpsi_is_time_derivative_of_psi(O, t):
    N, L = 12, 12.0
    cosmo = O.Cosmo.radiation(**FRW)
    _, Rk = r_spec(O, N, L, 0.05)
    v = O.VariantEquation(O.COMOVING, N, L, 1.0, cosmo, field=Rk)
    h = 1e-5
    fd = (v.psi(t + h)[0] - v.psi(t - h)[0]) / (2 * h)
    d = v.psi(t)[1]
    assert np.abs(fd - d).max() <= 1e-7 * np.abs(d).max()


def test_transfer_small_k_limit(O):
    """(sin x - x cos x) * 2/x^3 -> 2/3 as x -> 0: Psi = (2/3) R for long modes."""
    N = 8
    Rk = np.zeros(2 * N * N * (N // 2 + 1))
    Rk[2 * 1] = 3.0  # mode (0,0,1)
    ps, _ = O.comoving_psi_fft(N, Rk, 1e-4, 0.5)
    assert abs(ps[2] - 2.0 / N ** 3) <= 1e-8 * 2.0 / N ** 3
    assert ps[0] == 0.0  # the k = 0 mode is dropped


@pytest.mark.parametrize("N,M", [(16, 8), (24, 8), (12, 12)])
def test_cutoff_of_band_limited_field_is_subsampling(O, N, M):
    rng = np.random.default_rng(0)
    Mh = M // 2
    # a real field with modes only below M/2 on every axis
    fk = np.zeros((N, N, N // 2 + 1), dtype=complex)
    for _ in range(20):
        a, b, c = rng.integers(-Mh + 1, Mh, size=3)
        c = abs(c)
        fk[a % N, b % N, c] += rng.standard_normal() + 1j * rng.standard_normal()
    f = np.fft.irfftn(fk, s=(N, N, N), axes=(0, 1, 2))
    spec = O.fft_r2c(N, f.ravel())
    cut = O.cutoff_fouriers(N, M, spec)
    sd = N // M
    sub = f[::sd, ::sd, ::sd].ravel()
    back = O.fft_c2r(M, cut) / M ** 3
    assert np.abs(back - sub).max() <= 1e-12 * np.abs(sub).max()
    if N == M:
        assert np.array_equal(cut, spec)
