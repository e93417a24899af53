"""Pins of the oracle that do not need the reference's output files.

Each test checks a fact the mathematics fixes, chosen so that a plausible slip
in the restatement (a dropped term, a wrong sign, a transposed index, a missing
periodic wrap, a mis-weighted half-spectrum column) fails it:

* the RK4 step against the exact amplification matrix of a lattice plane wave
  (src/initializer.hpp:247-269) along each axis;
* the homogeneous field against a scalar RK4 of phi'' + 3H phi' + m^2 phi = 0;
* the oracle RK4/integrate_const against the reference's own vendored
  Boost.odeint + eigen_operations build (oracle/_ref/odeint_ref), bit for bit;
* the FFT against a direct DFT and against numpy (pocketfft);
* the binning against the full-cube definition (fdm3d.hpp:27-40), Parseval,
  and the set of empty bins (sums of three bounded squares).
"""
import math
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

REF_BIN = os.path.join(ROOT, "oracle", "_ref", "odeint_ref")


def rk4_matrix(A, dt):
    I = np.eye(2)
    M = dt * A
    return I + M + M @ M / 2 + M @ M @ M / 6 + M @ M @ M @ M / 24


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("N,L,m", [(8, 8.0, 1.0), (16, 10.0, 0.5)])
def test_plane_wave_rk4_exact(O, axis, N, L, m):
    k = 1
    idx = np.indices((N, N, N))[axis]
    phi = np.cos(2 * np.pi * k * idx / N).ravel()
    x = np.concatenate([phi, np.zeros(N ** 3)])
    eq = O.Equation(N, L, m, 0.0, O.Cosmo.minkowski())
    dt, steps = 0.1, 20
    eq.rk4_run(x, 0.0, dt, steps)
    h = L / N
    w2 = m * m + (4 / (h * h)) * math.sin(math.pi * k / N) ** 2
    R = np.linalg.matrix_power(rk4_matrix(np.array([[0, 1], [-w2, 0]]), dt), steps)
    exp_phi = R[0, 0] * phi
    exp_pi = R[1, 0] * phi
    assert np.max(np.abs(x[: N ** 3] - exp_phi)) < 1e-13
    assert np.max(np.abs(x[N ** 3:] - exp_pi)) < 1e-12


def test_homogeneous_field_frw(O):
    """A constant field obeys phi'' = -3H phi' - m^2 phi; same RK4 on the scalar."""
    N, L, m, f, df = 8, 8.0, 1.3, 0.7, -0.2
    cosmo = O.Cosmo.radiation(1.0, 0.5, 1.0)
    eq = O.Equation(N, L, m, 0.0, cosmo)
    x = np.concatenate([np.full(N ** 3, f), np.full(N ** 3, df)])
    t0, dt, n = 1.0, 0.05, 40
    eq.rk4_run(x, t0, dt, n)

    def F(y, t):
        H = O.cosmo_H(cosmo, t)
        return np.array([y[1], -3.0 * H * y[1] - m * m * y[0]])

    y = np.array([f, df])
    for k in range(n):
        t = t0 + k * dt
        k1 = F(y, t)
        k2 = F(y + dt / 2 * k1, t + dt / 2)
        k3 = F(y + dt / 2 * k2, t + dt / 2)
        k4 = F(y + dt * k3, t + dt)
        y = y + dt / 6 * k1 + dt / 3 * k2 + dt / 3 * k3 + dt / 6 * k4
    assert np.max(np.abs(x[: N ** 3] - y[0])) < 1e-13
    assert np.max(np.abs(x[N ** 3:] - y[1])) < 1e-13


def test_cosmology_closed_form(O):
    """a = (t/t_i)^(1/2), H = 1/(2t) for a1=1, H1=1/2, t1=1 (radiation, p=1)."""
    c = O.Cosmo.radiation(1.0, 0.5, 1.0)
    for t in (1.0, 2.5, 51.0):
        assert abs(O.cosmo_a(c, t) - math.sqrt(t)) < 1e-15 * math.sqrt(t)
        assert abs(O.cosmo_H(c, t) - 0.5 / t) < 1e-16
        assert abs(O.cosmo_eta(c, t) - 2 * math.sqrt(t)) < 1e-14
    m = O.Cosmo.minkowski()
    assert O.cosmo_a(m, 7.0) == 1.0 and O.cosmo_H(m, 7.0) == 0.0


def test_lambda_term_and_energy_density(O):
    """rho integrates the force: with a=1, H=0 the lattice energy sum_x rho is
    conserved by the continuous flow; check d/dt sum rho = 0 through the RHS
    (central difference of the exact lattice energy along dx/dt)."""
    N, m, lam, inv = 8, 0.9, 0.3, 1.7
    rng = np.random.default_rng(5)
    x = np.concatenate([0.5 * rng.standard_normal(N ** 3), 0.5 * rng.standard_normal(N ** 3)])
    d = O.rhs(N, x, m, lam, 0.0, inv)

    def E(xx):
        phi = xx[: N ** 3].reshape(N, N, N)
        pi = xx[N ** 3:]
        grad = sum(((np.roll(phi, -1, ax) - phi) ** 2).sum() for ax in range(3))
        return 0.5 * (pi ** 2).sum() + 0.5 * inv * grad + 0.5 * m * m * (phi ** 2).sum() + 0.25 * lam * (phi ** 4).sum()

    eps = 1e-6
    dE = (E(x + eps * d) - E(x - eps * d)) / (2 * eps)
    assert abs(dE) < 1e-7 * np.abs(d).sum()
    # the observer's rho uses central differences (equations.cpp:106-138): its
    # sum equals the forward-difference energy built from next-nearest pairs
    rho = O.energy_density(N, x, m, lam, inv)
    phi = x[: N ** 3].reshape(N, N, N)
    pi = x[N ** 3:].reshape(N, N, N)
    cd = sum((np.roll(phi, -1, ax) - np.roll(phi, 1, ax)) ** 2 for ax in range(3))
    ref = 0.5 * pi ** 2 + 0.125 * inv * cd + 0.5 * m * m * phi ** 2 + 0.25 * lam * phi ** 4
    np.testing.assert_allclose(rho, ref.ravel(), rtol=1e-13, atol=1e-14)


@pytest.mark.skipif(not os.path.exists(REF_BIN), reason="reference tree not mounted here")
@pytest.mark.parametrize("lam,H1", [(0.0, 0.5), (0.2, 0.5), (0.0, 0.0)])
def test_oracle_equals_reference_odeint(O, tmp_path, lam, H1):
    """The reference's integrate_const + runge_kutta4_classic + eigen_operations
    (its own headers) around the oracle RHS give bit-identical states and the
    same observer times as the oracle's restated integrator."""
    N, L, m, t0, t1, dt = 12, 15.0, 1.0, 1.0, 2.2, 0.05
    x = O.initial_state(N, L, 7, O.SPEC_GAUSS, (0.5,), 0.6)
    x.tofile(str(tmp_path / "in.bin"))
    r = subprocess.run([REF_BIN, str(tmp_path / "in.bin"), str(tmp_path), str(N), repr(L), repr(m), repr(lam),
                        "1.0", repr(H1), "1.0", repr(t0), repr(t1), repr(dt)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    cosmo = O.Cosmo.radiation(1.0, H1, 1.0) if H1 else O.Cosmo.minkowski()
    eq = O.Equation(N, L, m, lam, cosmo)
    times = []
    n = O.integrate_const(eq.rk4_step, lambda xx, t: times.append(t), x, t0, t1, dt)
    assert int(r.stdout) == n
    assert np.array_equal(np.fromfile(str(tmp_path / "state.bin")), x)
    assert np.array_equal(np.fromfile(str(tmp_path / "times.bin")), np.array(times))
    cz = np.fromfile(str(tmp_path / "cosmo.bin")).reshape(16, 3)
    for k in range(16):
        t = t0 + k * (t1 - t0) / 15.0
        assert cz[k, 0] == O.cosmo_a(cosmo, t) and cz[k, 1] == O.cosmo_H(cosmo, t)
        if H1:
            assert cz[k, 2] == O.cosmo_eta(cosmo, t)


@pytest.mark.parametrize("N", [4, 6, 8])
def test_fft_vs_direct_dft(O, N):
    f = np.random.default_rng(N).standard_normal(N ** 3)
    a = O.fft_r2c(N, f)
    d = O.dft_r2c_direct(N, f)
    assert np.max(np.abs(a - d)) < 1e-13 * N ** 1.5


@pytest.mark.parametrize("N", [12, 16, 96])
def test_fft_vs_numpy_and_roundtrip(O, N):
    f = np.random.default_rng(N).standard_normal(N ** 3)
    a = O.fft_r2c(N, f).reshape(N, N, N // 2 + 1, 2)
    b = np.fft.rfftn(f.reshape(N, N, N))
    assert np.max(np.abs(a[..., 0] + 1j * a[..., 1] - b)) < 1e-14 * np.abs(b).max()
    back = O.fft_c2r(N, O.fft_r2c(N, f)) / N ** 3
    assert np.max(np.abs(back - f)) < 1e-14 * N


@pytest.mark.parametrize("N", [4, 8])
def test_binning_equals_full_cube_definition(O, N):
    """spectrum[s] = sum over the full reciprocal cube with |n|^2 = s of |f_n|^2
    (fdm3d.hpp:27-40), computed here from numpy's full complex FFT."""
    f = np.random.default_rng(3).standard_normal(N ** 3)
    full = np.fft.fftn(f.reshape(N, N, N))
    sh = np.minimum(np.arange(N), N - np.arange(N))
    s = sh[:, None, None] ** 2 + sh[None, :, None] ** 2 + sh[None, None, :] ** 2
    ref = np.bincount(s.ravel(), weights=(np.abs(full) ** 2).ravel(), minlength=O.nbins(N))
    np.testing.assert_allclose(O.power_spectrum(N, f), ref, rtol=1e-12, atol=1e-12 * ref.max())


@pytest.mark.parametrize("N", [8, 16, 32])
def test_parseval_and_empty_bins(O, N):
    f = np.random.default_rng(4).standard_normal(N ** 3)
    sp = O.power_spectrum(N, f)
    assert abs(sp.sum() - N ** 3 * (f ** 2).sum()) < 1e-12 * sp.sum()
    h = N // 2
    r = np.arange(h + 1) ** 2
    reachable = np.zeros(O.nbins(N), bool)
    reachable[(r[:, None, None] + r[None, :, None] + r[None, None, :]).ravel()] = True
    assert np.array_equal(sp == 0, ~reachable)
    # Legendre: 4^a(8b+7) is never a sum of three squares
    for s in range(sp.size):
        t = s
        while t and t % 4 == 0:
            t //= 4
        if s and t % 8 == 7:
            assert sp[s] == 0


def test_mode_spectrum_plane_wave(O):
    """cos(2 pi c/N), pi = 0: all power in bin 1, value N^6/2 (two modes of |N^3/2|^2)."""
    N = 8
    c = np.indices((N, N, N))[2]
    x = np.concatenate([np.cos(2 * np.pi * c / N).ravel(), np.zeros(N ** 3)])
    sp = O.mode_power_spectrum(N, 8.0, 1.0, 1.0, x)
    assert abs(sp[1] - N ** 6 / 2) < 1e-9 * N ** 6
    assert np.all(np.abs(np.delete(sp, 1)) < 1e-18 * N ** 6)


def test_radial_bins(O):
    N = 16
    sb = np.random.default_rng(1).random(O.nbins(N))
    rb = O.radial_bins(N, sb)
    s = np.arange(sb.size)
    j = np.floor(np.sqrt(s) + 1e-12).astype(int)
    ref = np.bincount(j[j < N // 2], weights=sb[j < N // 2], minlength=N // 2)
    np.testing.assert_allclose(rb, ref, rtol=1e-15)


def test_grf_spectrum_shape(O):
    """The GRF multiplies the white-noise modes by sqrt(P(|k|)) (random_field.cpp:60-73):
    a one-bin spectrum table leaves power in that bin only."""
    N, L = 16, 16.0
    P = np.zeros(O.nbins(N))
    P[5] = 1.0
    phi = O.grf_from_normals(N, O.normals(0, N ** 3), P)
    sp = O.power_spectrum(N, phi)
    assert sp[5] > 0 and np.all(np.abs(np.delete(sp, 5)) < 1e-20 * sp[5])


def test_normals_are_std_normal_distribution(O):
    """mt19937(seed) + normal_distribution<double>: deterministic per seed,
    standard normal in distribution."""
    a = O.normals(123, 200000)
    assert np.array_equal(a, O.normals(123, 200000))
    assert abs(a.mean()) < 0.01 and abs(a.std() - 1) < 0.01


def test_f32c_compensation_matters(O):
    """The compensated FP32 step stays within ~1e-12 of FP64 over many steps,
    while the same step with the compensation term dropped (a mistake planted
    in the oracle copy here) drifts to ~1e-7."""
    N, L = 16, 16.0
    x = O.initial_state(N, L, 9, O.SPEC_GAUSS, (0.5,), 0.5)
    eq = O.Equation(N, L, 1.0, 0.0, O.Cosmo.radiation(1.0, 0.5, 1.0))
    ref = x.copy()
    hi = x.astype(np.float32)
    lo = (x - hi.astype(np.float64)).astype(np.float32)
    hi2, lo2 = hi.copy(), np.zeros_like(lo)
    t0, dt = 1.0, 0.05
    for k in range(100):
        t = t0 + k * dt
        eq.rk4_step(ref, t, dt)
        eq.rk4_step_f32c(hi, lo, t, dt)
        eq.rk4_step_f32c(hi2, lo2, t, dt)
        lo2[:] = 0.0
    comp = hi.astype(np.float64) + lo.astype(np.float64)
    nocomp = hi2.astype(np.float64)
    e1 = np.linalg.norm(comp - ref) / np.linalg.norm(ref)
    e2 = np.linalg.norm(nocomp - ref) / np.linalg.norm(ref)
    assert e1 < 1e-6 and e2 > 5 * e1, (e1, e2)
