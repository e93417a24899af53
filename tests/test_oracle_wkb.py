"""Pins of the oracle's WKB extrapolation (wkb.cpp:5-121), CPU only.

* The closed-form phase is the integral of omega_eff(t) = sqrt(m^2 + k^2/a(t)^2 +
  3/(16 t^2)) with a(t)^2 = 2 a1^2 H1 t (radiation era, t1 = 1/(2 H1)): checked
  against adaptive quadrature for a spread of k (a sign or power slip in any of its
  five terms fails this);
* evaluating at t = t_i returns the initial state (phase 0; the a^1.5 and 1.5 H
  factors of the constructor and of evaluate_at cancel) to FFT round-off.
"""
import numpy as np
import pytest
from scipy.integrate import quad


@pytest.mark.parametrize("t_f", [12.0, 40.0, 130.0])
def test_phase_is_integral_of_omega(O, t_f):
    N, L, m, a1, H1 = 16, 12.8, 1.0, 1.0, 0.05
    t_i = 1.0 / (2 * H1)
    _, _, _, phase = O.wkb_tables(N, L, m, a1, H1, t_i, t_f)
    for s in (0, 1, 7, 40, O.nbins(N) - 1):
        k2 = s * (2 * np.pi / L) ** 2

        def omega(t):
            return np.sqrt(m * m + k2 / (2 * a1 * a1 * H1 * t) + 3 / (16 * t * t))

        ref, err = quad(omega, t_i, t_f, epsabs=0, epsrel=1e-13, limit=200)
        assert abs(phase[s] - ref) <= 1e-10 * abs(ref), (s, phase[s], ref)


def test_evaluate_at_start_returns_state(O):
    N, L, m = 16, 12.8, 1.0
    cosmo = O.Cosmo.radiation(1.0, 0.05, 10.0)
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.standard_normal(N ** 3), 0.2 * rng.standard_normal(N ** 3)])
    w = O.WKB(N, L, m, cosmo, x, 10.0)
    y = w.evaluate_at(10.0)
    assert np.abs(y - x).max() <= 1e-13 * np.abs(x).max()
    c, s, _, ph = O.wkb_tables(N, L, m, 1.0, 0.05, 10.0, 10.0)
    assert np.all(ph == 0.0) and np.all(c == 1.0) and np.all(s == 0.0)


def test_rotation_preserves_mode_invariant(O):
    """Per mode, evaluate_at rotates (phi^, pi^) by [[c, s/w], [-s w, c]] with
    w = sqrt(omega_eff(t_i)): a determinant-one map that keeps
    omega_eff(t_i) |phi^|^2 + |pi^|^2. Recovering the rotated modes from the output
    state (undoing the a^1.5 and 1.5 H factors) must give the same invariant per
    mode; a sign or factor slip in either row breaks it."""
    N, L, m = 12, 9.6, 1.0
    H1, t_i, t_f = 0.05, 10.0, 47.0
    cosmo = O.Cosmo.radiation(1.0, H1, t_i)
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.standard_normal(N ** 3), 0.3 * rng.standard_normal(N ** 3)])
    w = O.WKB(N, L, m, cosmo, x, t_i)
    y = w.evaluate_at(t_f)
    a15 = O.cosmo_a(cosmo, t_f) ** 1.5
    H = O.cosmo_H(cosmo, t_f)
    n3 = N ** 3
    p_new = O.fft_r2c(N, y[:n3] * a15)
    d_new = O.fft_r2c(N, y[n3:] * a15 + 1.5 * H * y[:n3] * a15)
    fs = 2 * N * N * (N // 2 + 1)
    p_old, d_old = w.ffts[:fs], w.ffts[fs:]
    _, _, sqw, _ = O.wkb_tables(N, L, m, 1.0, H1, t_i, t_f)
    a = np.arange(N)[:, None, None]
    b = np.arange(N)[None, :, None]
    c = np.arange(N // 2 + 1)[None, None, :]
    s = (np.minimum(a, N - a) ** 2 + np.minimum(b, N - b) ** 2 + c ** 2).ravel()
    om = sqw[s] ** 2

    def inv(p, d):
        pc = p.reshape(-1, 2)
        dc = d.reshape(-1, 2)
        return om * (pc ** 2).sum(1) + (dc ** 2).sum(1)

    q0, q1 = inv(p_old, d_old), inv(p_new, d_new)
    assert np.abs(q1 - q0).max() <= 1e-10 * q0.max()
