"""Pins of the oracle's scale-factor dependence (CPU only).

The 1/a^2 of the gradient term enters the Klein-Gordon equation in an expanding
background through inv_ah_sqr = 1/((L/N)^2 a^2) (src/equations.cpp:80) and the
mode spectrum through the continuum frequency m^2 + (2 pi/L)^2 s / a^2
(src/fdm3d.cpp:64). The reference's shipped outputs are all at a = 1 and the
homogeneous-field pin has no gradient, so these tests fix both factors at a != 1
by closed forms:

* a lattice plane wave cos(2 pi k x / N) along each axis stays a plane wave in FRW;
  its amplitude obeys the scalar ODE
      phi'' = -3 H phi' - (m^2 + 4 sin^2(pi k / N) / (h^2 a(t)^2)) phi,
  integrated here by the same classic RK4 (same stage times) on two numbers;
* the mode spectrum of phi = A cos(2 pi c/N), pi = B cos(4 pi b/N) has exactly two
  non-empty bins, s = 1 with A^2 N^6 / 2 and s = 4 with
  B^2 N^6 / 2 / (m^2 + (2 pi / L)^2 4 / a^2).

Each check is also run against a copy of the oracle with a 1/a slip planted in
place of 1/a^2 (in inv_ah_sqr and in the mode-spectrum denominator); the copy must
fail both, so the pins do discriminate that mistake.
"""
import importlib.util
import math
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

ORACLE_DIR = os.path.join(ROOT, "oracle")


@pytest.fixture(scope="module")
def O_planted(tmp_path_factory):
    """A second instance of the oracle module bound to a copy of cf_oracle.cpp in
    which a^2 is replaced by a in both scale-factor sites."""
    src = open(os.path.join(ORACLE_DIR, "cf_oracle.cpp")).read()
    inv_site = "return 1.0 / ((L / N) * (L / N)) / (a_t * a_t);"
    mode_site = "(m * m + (2 * pi / L) * (2 * pi / L) * s / (a_t * a_t))"
    assert src.count(inv_site) == 1 and src.count(mode_site) == 1, "scale-factor sites moved"
    src = src.replace(inv_site, "return 1.0 / ((L / N) * (L / N)) / a_t;")
    src = src.replace(mode_site, "(m * m + (2 * pi / L) * (2 * pi / L) * s / a_t)")
    d = tmp_path_factory.mktemp("planted")
    cpp, so = str(d / "planted.cpp"), str(d / "libplanted.so")
    with open(cpp, "w") as f:
        f.write(src)
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-march=x86-64-v3", "-ffp-contract=off", "-fno-fast-math",
                        "-fPIC", "-shared", "-pthread", "-w", "-o", so, cpp], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    spec = importlib.util.spec_from_file_location("cf_oracle_planted", os.path.join(ORACLE_DIR, "cf_oracle.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    mod._PATH = so
    return mod


def scalar_mode_rk4(f, g, t0, dt, n, m, w2_grad, H, a):
    """Classic RK4 (runge_kutta4_classic stage times t, t+dt/2, t+dt/2, t+dt) of
    f' = g, g' = -3 H(t) g - (m^2 + w2_grad / a(t)^2) f."""

    def F(y, t):
        return np.array([y[1], -3.0 * H(t) * y[1] - (m * m + w2_grad / (a(t) * a(t))) * y[0]])

    y = np.array([f, g], dtype=np.float64)
    for k in range(n):
        t = t0 + k * dt
        k1 = F(y, t)
        k2 = F(y + dt / 2 * k1, t + dt / 2)
        k3 = F(y + dt / 2 * k2, t + dt / 2)
        k4 = F(y + dt * k3, t + dt)
        y = y + dt / 6 * k1 + dt / 3 * k2 + dt / 3 * k3 + dt / 6 * k4
    return y


def frw_plane_wave_error(O, axis, N, L, m, k):
    """Max deviation of the oracle's RK4 run of a plane wave in a radiation-era
    background from the scalar amplitude ODE (t from 4 to 6: a from 2 to 2.45)."""
    cosmo = O.Cosmo.radiation(1.0, 0.5, 1.0)
    eq = O.Equation(N, L, m, 0.0, cosmo)
    pattern = np.cos(2 * np.pi * k * np.indices((N, N, N))[axis] / N).ravel()
    A, B = 0.8, -0.3
    x = np.concatenate([A * pattern, B * pattern])
    t0, dt, n = 4.0, 0.05, 40
    eq.rk4_run(x, t0, dt, n)
    h = L / N
    w2 = 4.0 * math.sin(math.pi * k / N) ** 2 / (h * h)
    y = scalar_mode_rk4(A, B, t0, dt, n, m, w2, lambda t: O.cosmo_H(cosmo, t), lambda t: O.cosmo_a(cosmo, t))
    return max(np.max(np.abs(x[: N ** 3] - y[0] * pattern)), np.max(np.abs(x[N ** 3:] - y[1] * pattern)))


def two_mode_spectrum_error(O, N, L, m, a_t):
    """Relative deviation of the oracle's mode spectrum of a two-mode state from
    the closed form (largest over all bins, relative to each expected bin)."""
    idx = np.indices((N, N, N))
    A, B = 0.7, 1.9
    phi = A * np.cos(2 * np.pi * idx[2] / N)      # s = 1
    pi_ = B * np.cos(2 * np.pi * 2 * idx[1] / N)  # s = 4
    x = np.concatenate([phi.ravel(), pi_.ravel()])
    sp = O.mode_power_spectrum(N, L, m, a_t, x)
    ref = np.zeros_like(sp)
    ref[1] = A * A * N ** 6 / 2
    ref[4] = B * B * N ** 6 / 2 / (m * m + (2 * math.pi / L) ** 2 * 4 / (a_t * a_t))
    return np.max(np.abs(sp - ref)) / ref[4]


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("N,L,m,k", [(8, 8.0, 0.5, 1), (12, 9.0, 0.3, 2)])
def test_frw_plane_wave_rk4(O, axis, N, L, m, k):
    assert frw_plane_wave_error(O, axis, N, L, m, k) < 1e-13


def test_frw_plane_wave_planted_slip_fails(O_planted):
    for axis in (0, 1, 2):
        assert frw_plane_wave_error(O_planted, axis, 8, 8.0, 0.5, 1) > 1e-4


def test_inv_ah_sqr_closed_form(O, O_planted):
    for N, L, a in ((8, 8.0, 2.0), (64, 48.0, 1.7), (512, 512.0, 5.3)):
        ref = (N / L) ** 2 / (a * a)
        assert abs(O.inv_ah_sqr(N, L, a) - ref) <= 4e-16 * ref
        assert abs(O_planted.inv_ah_sqr(N, L, a) - ref) > 0.1 * ref


@pytest.mark.parametrize("a_t", [2.5, 0.6])
def test_mode_spectrum_two_modes_at_a(O, a_t):
    assert two_mode_spectrum_error(O, 8, 8.0, 0.5, a_t) < 1e-12


def test_mode_spectrum_planted_slip_fails(O_planted):
    assert two_mode_spectrum_error(O_planted, 8, 8.0, 0.5, 2.5) > 1e-3


# The compensated storage with the single-precision stencil (orc_rk4_step_f32s): the
# same closed forms, to the precision of float forces on float stage inputs.
def frw_plane_wave_error_f32s(O, axis, N, L, m, k):
    cosmo = O.Cosmo.radiation(1.0, 0.5, 1.0)
    eq = O.Equation(N, L, m, 0.0, cosmo)
    pattern = np.cos(2 * np.pi * k * np.indices((N, N, N))[axis] / N).ravel()
    A, B = 0.8, -0.3
    x = np.concatenate([A * pattern, B * pattern])
    hi = x.astype(np.float32)
    lo = (x - hi.astype(np.float64)).astype(np.float32)
    t0, dt, n = 4.0, 0.05, 40
    for s in range(n):
        eq.rk4_step_f32s(hi, lo, t0 + s * dt, dt)
    xs = hi.astype(np.float64) + lo.astype(np.float64)
    h = L / N
    w2 = 4.0 * math.sin(math.pi * k / N) ** 2 / (h * h)
    y = scalar_mode_rk4(A, B, t0, dt, n, m, w2, lambda t: O.cosmo_H(cosmo, t), lambda t: O.cosmo_a(cosmo, t))
    return max(np.max(np.abs(xs[: N ** 3] - y[0] * pattern)), np.max(np.abs(xs[N ** 3:] - y[1] * pattern)))


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("N,L,m,k", [(8, 8.0, 0.5, 1), (12, 9.0, 0.3, 2)])
def test_f32s_frw_plane_wave_rk4(O, axis, N, L, m, k):
    err = frw_plane_wave_error_f32s(O, axis, N, L, m, k)
    assert err < 1e-7, err
    assert err > 0.0  # float forces: not the FP64 result


def test_f32s_frw_plane_wave_planted_slip_fails(O_planted):
    for axis in (0, 1, 2):
        assert frw_plane_wave_error_f32s(O_planted, axis, 8, 8.0, 0.5, 1) > 1e-4


def test_f32s_homogeneous_lambda(O):
    """A homogeneous field with lambda phi^4/4 in a radiation era obeys
    phi'' = -3 H phi' - m^2 phi - lambda phi^3 (no gradient); the f32s step follows the
    classic RK4 of that scalar ODE to float precision, and an equation with lambda
    dropped does not."""
    N, L, m, lam = 8, 8.0, 0.7, 0.1
    cosmo = O.Cosmo.radiation(1.0, 0.5, 1.0)
    H = lambda t: O.cosmo_H(cosmo, t)  # noqa: E731

    def scalar(f, g, t0, dt, n, lam_):
        def F(y, t):
            return np.array([y[1], -3.0 * H(t) * y[1] - m * m * y[0] - lam_ * y[0] ** 3])
        y = np.array([f, g])
        for s in range(n):
            t = t0 + s * dt
            k1 = F(y, t)
            k2 = F(y + dt / 2 * k1, t + dt / 2)
            k3 = F(y + dt / 2 * k2, t + dt / 2)
            k4 = F(y + dt * k3, t + dt)
            y = y + dt / 6 * k1 + dt / 3 * k2 + dt / 3 * k3 + dt / 6 * k4
        return y

    f0, g0, t0, dt, n = 1.3, -0.2, 2.0, 0.05, 60
    x = np.concatenate([np.full(N ** 3, f0), np.full(N ** 3, g0)])
    hi = x.astype(np.float32)
    lo = (x - hi.astype(np.float64)).astype(np.float32)
    eq = O.Equation(N, L, m, lam, cosmo)
    for s in range(n):
        eq.rk4_step_f32s(hi, lo, t0 + s * dt, dt)
    xs = hi.astype(np.float64) + lo.astype(np.float64)
    y = scalar(f0, g0, t0, dt, n, lam)
    y0 = scalar(f0, g0, t0, dt, n, 0.0)
    err = max(np.max(np.abs(xs[: N ** 3] - y[0])), np.max(np.abs(xs[N ** 3:] - y[1])))
    assert err < 1e-7, err
    assert max(abs(y[0] - y0[0]), abs(y[1] - y0[1])) > 1e-2  # the lambda term is visible


def test_f32s_tracks_fp64_random_field(O):
    """Over 100 FRW steps of a random field (lambda = 0.1) the f32s state stays within
    1e-6 of FP64 (relative L2) and within 2x the deviation of the FP64-arithmetic
    compensated step."""
    N, L = 16, 16.0
    x = O.initial_state(N, L, 9, O.SPEC_GAUSS, (0.5,), 0.5)
    eq = O.Equation(N, L, 1.0, 0.1, O.Cosmo.radiation(1.0, 0.5, 1.0))
    ref = x.copy()
    hi = x.astype(np.float32)
    lo = (x - hi.astype(np.float64)).astype(np.float32)
    hc, lc = hi.copy(), lo.copy()
    t0, dt = 1.0, 0.05
    for s in range(100):
        t = t0 + s * dt
        eq.rk4_step(ref, t, dt)
        eq.rk4_step_f32s(hi, lo, t, dt)
        eq.rk4_step_f32c(hc, lc, t, dt)
    es = np.linalg.norm(hi.astype(np.float64) + lo - ref) / np.linalg.norm(ref)
    ec = np.linalg.norm(hc.astype(np.float64) + lc - ref) / np.linalg.norm(ref)
    assert es < 1e-6 and es < 2 * ec, (es, ec)
