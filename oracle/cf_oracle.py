"""TEST INFRASTRUCTURE ONLY -- the CPU oracle for the Cosmic-Fields-Lite hot path.

ctypes front end of ``libcforacle.so`` (``cf_oracle.cpp``) plus a restatement of
the control flow around it:

* ``integrate_const`` with an explicit stepper
  (external/boost/numeric/odeint/integrate/detail/integrate_const.hpp:42-68 and
  util/detail/less_with_sign.hpp:46-53);
* ``ConstIntervalObserver::operator()`` (src/observer.hpp:87-124);
* the workload initial conditions (Gaussian random fields per
  src/random_field.cpp:36-89, rescaled to a target rms).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module; the product never does.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
# Library this module binds. Tests that plant a mistake in a copy of the oracle
# load that copy into a separate instance of this module by changing the path.
_PATH = os.path.join(_HERE, "libcforacle.so")

DBL_EPS = np.finfo(np.float64).eps


class Cosmo(C.Structure):
    """StaticEOSCosmology (src/physics.hpp:15-56)."""

    _fields_ = [("a1", C.c_double), ("H1", C.c_double), ("t1", C.c_double), ("p", C.c_double)]

    @staticmethod
    def minkowski():
        # physics.hpp:30-31, the default constructor
        return Cosmo(1.0, 0.0, 0.0, 1.0)

    @staticmethod
    def radiation(a1, H1, t1):
        # physics.hpp:25-27, p = 1
        return Cosmo(a1, H1, t1, 1.0)


class Eq(C.Structure):
    _fields_ = [("N", C.c_int64), ("L", C.c_double), ("m", C.c_double), ("lam", C.c_double),
                ("cosmo", Cosmo)]


_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)


def lib():
    global _LIB
    if _LIB is None:
        path = _PATH
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(path)
        L.orc_set_threads.argtypes = [C.c_int]
        for n in ("orc_cosmo_a", "orc_cosmo_H", "orc_cosmo_eta"):
            getattr(L, n).argtypes = [C.POINTER(Cosmo), C.c_double]
            getattr(L, n).restype = C.c_double
        L.orc_inv_ah_sqr.argtypes = [C.c_int64, C.c_double, C.c_double]
        L.orc_inv_ah_sqr.restype = C.c_double
        L.orc_rhs.argtypes = [C.c_int64, _dp, _dp, C.c_double, C.c_double, C.c_double, C.c_double]
        L.orc_energy_density.argtypes = [C.c_int64, _dp, _dp, C.c_double, C.c_double, C.c_double]
        L.orc_rk4_step.argtypes = [C.POINTER(Eq), _dp, C.c_double, C.c_double, _dp]
        L.orc_rk4_run.argtypes = [C.POINTER(Eq), _dp, C.c_double, C.c_double, C.c_int, _dp]
        L.orc_rk4_run.restype = C.c_int
        L.orc_rk4_step_f32c.argtypes = [C.POINTER(Eq), _fp, _fp, C.c_double, C.c_double, _dp, _fp]
        L.orc_rk4_step_f32s.argtypes = [C.POINTER(Eq), _fp, _fp, C.c_double, C.c_double, _dp, _fp]
        L.orc_fft_r2c.argtypes = [C.c_int64, _dp, _dp]
        L.orc_fft_c2r.argtypes = [C.c_int64, _dp, _dp]
        L.orc_dft_r2c_direct.argtypes = [C.c_int64, _dp, _dp]
        L.orc_bin_power.argtypes = [C.c_int64, _dp, _dp]
        L.orc_power_spectrum.argtypes = [C.c_int64, _dp, _dp]
        L.orc_bin_mode_power.argtypes = [C.c_int64, C.c_double, C.c_double, C.c_double, _dp, _dp, _dp]
        L.orc_mode_power_spectrum.argtypes = [C.c_int64, C.c_double, C.c_double, C.c_double, _dp, _dp]
        L.orc_normals.argtypes = [C.c_uint32, C.c_int64, _dp]
        L.orc_grf_from_normals.argtypes = [C.c_int64, _dp, _dp, _dp, _dp]
        L.orc_spectrum_table.argtypes = [C.c_int, C.c_int64, C.c_double, _dp, _dp]
        L.orc_rescale_rms.argtypes = [C.c_int64, _dp, C.c_double]
        L.orc_rescale_rms.restype = C.c_double
        L.orc_scaled_fourier_modes.argtypes = [C.c_int64, _dp, _dp, _dp]
        L.orc_comoving_psi_fft.argtypes = [C.c_int64, _dp, _dp, _dp, C.c_double, C.c_double]
        L.orc_curvature_energy_density.argtypes = [C.c_int64, _dp, _dp, _dp, C.c_double, C.c_double]
        L.orc_ci_pade.argtypes = [C.c_double]
        L.orc_rhs_fixed_curvature.argtypes = [C.c_int64, _dp, _dp, _dp, C.c_double, C.c_double, C.c_double]
        L.orc_rhs_sqrt_potential.argtypes = [C.c_int64, _dp, _dp, C.c_double, C.c_double, C.c_double,
                                             C.c_double]
        L.orc_energy_density_sqrt.argtypes = [C.c_int64, _dp, _dp, C.c_double, C.c_double, C.c_double]
        L.orc_rhs_variable_curvature.argtypes = [C.c_int64, _dp, _dp, _dp, _dp, C.c_double, C.c_double,
                                                 C.c_double]
        L.orc_rhs_approx_curvature.argtypes = [C.c_int64, C.c_int64, _dp, _dp, _dp, _dp, C.c_double,
                                               C.c_double, C.c_double]
        L.orc_cutoff_fouriers.argtypes = [C.c_int64, C.c_int64, _dp, _dp]
        L.orc_curvature_energy_density_cuda.argtypes = [C.c_int64, _dp, _dp, _dp, C.c_double, C.c_double]
        L.orc_wkb_tables.argtypes = [C.c_int64] + [C.c_double] * 6 + [_dp] * 4
        L.orc_wkb_init.argtypes = [C.c_int64, _dp, C.c_double, C.c_double, _dp]
        L.orc_wkb_evaluate.argtypes = [C.c_int64, _dp, _dp, _dp, _dp, C.c_double, C.c_double, _dp]
        L.orc_ci_pade.restype = C.c_double
        L.orc_set_threads(min(8, os.cpu_count() or 1))
        _LIB = L
    return _LIB


def _p(a, dtype=np.float64):
    assert a.dtype == dtype and a.flags.c_contiguous, (a.dtype, a.flags)
    return a.ctypes.data_as(_dp if dtype == np.float64 else _fp)


def set_threads(n: int):
    lib().orc_set_threads(int(n))


def nbins(N: int) -> int:
    """Spectrum length 3(N/2)^2+1 (src/fdm3d.cpp:11)."""
    return 3 * (N // 2) * (N // 2) + 1


# --------------------------------------------------------------------------- cosmology
def cosmo_a(c: Cosmo, t: float) -> float:
    return lib().orc_cosmo_a(C.byref(c), t)


def cosmo_H(c: Cosmo, t: float) -> float:
    return lib().orc_cosmo_H(C.byref(c), t)


def cosmo_eta(c: Cosmo, t: float) -> float:
    return lib().orc_cosmo_eta(C.byref(c), t)


def inv_ah_sqr(N: int, L: float, a_t: float) -> float:
    return lib().orc_inv_ah_sqr(N, L, a_t)


# --------------------------------------------------------------------------- RHS / rho
def rhs(N, x, m, lam, H, inv):
    out = np.empty_like(x)
    lib().orc_rhs(N, _p(x), _p(out), m, lam, H, inv)
    return out


def energy_density(N, x, m, lam, inv):
    out = np.empty(N ** 3)
    lib().orc_energy_density(N, _p(x), _p(out), m, lam, inv)
    return out


# --------------------------------------------------------------------------- RK4
class Equation:
    """The oracle's KleinGordonEquationInFRW / +lambda phi^4 (equations.cpp:71-104)."""

    def __init__(self, N, L, m, lam=0.0, cosmo: Cosmo | None = None):
        self.N, self.L, self.m, self.lam = int(N), float(L), float(m), float(lam)
        self.cosmo = cosmo if cosmo is not None else Cosmo.minkowski()
        self._eq = Eq(self.N, self.L, self.m, self.lam, self.cosmo)
        self._work = None

    def a(self, t):
        return cosmo_a(self.cosmo, t)

    def H(self, t):
        return cosmo_H(self.cosmo, t)

    def inv_ah2(self, t):
        return inv_ah_sqr(self.N, self.L, self.a(t))

    def rhs(self, x, t):
        return rhs(self.N, x, self.m, self.lam, self.H(t), self.inv_ah2(t))

    def energy_density(self, x, t):
        return energy_density(self.N, x, self.m, self.lam, self.inv_ah2(t))

    def _w(self, k):
        n = 2 * self.N ** 3
        if self._work is None or self._work.size < k * n:
            self._work = np.empty(k * n)
        return self._work

    def rk4_step(self, x, t, dt):
        lib().orc_rk4_step(C.byref(self._eq), _p(x), t, dt, _p(self._w(5)))

    def rk4_run(self, x, t0, dt, n_steps):
        return lib().orc_rk4_run(C.byref(self._eq), _p(x), t0, dt, int(n_steps), _p(self._w(5)))

    def rk4_step_f32c(self, hi, lo, t, dt):
        n = 2 * self.N ** 3
        fw = np.empty(2 * n, np.float32)
        lib().orc_rk4_step_f32c(C.byref(self._eq), _p(hi, np.float32), _p(lo, np.float32), t, dt,
                                _p(self._w(2)), _p(fw, np.float32))

    def rk4_step_f32s(self, hi, lo, t, dt):
        """The single-precision-stencil compensated step (cf_oracle.cpp orc_rk4_step_f32s)."""
        n = 2 * self.N ** 3
        fw = np.empty(2 * n, np.float32)
        lib().orc_rk4_step_f32s(C.byref(self._eq), _p(hi, np.float32), _p(lo, np.float32), t, dt,
                                _p(self._w(1)), _p(fw, np.float32))


def less_eq_with_sign(t1, t2, dt):
    """util/detail/less_with_sign.hpp:46-53."""
    if dt > 0:
        return (t1 - t2) <= DBL_EPS
    return (t2 - t1) <= DBL_EPS


def integrate_const(step, observer, x, t0, t1, dt):
    """detail/integrate_const.hpp:42-68 for an explicit stepper.

    ``step(x, t, dt)`` advances x in place; ``observer(x, t)`` is called before
    each step and once after the last. Returns the number of steps.
    """
    time = t0
    n = 0
    while less_eq_with_sign(time + dt, t1, dt):
        observer(x, time)
        step(x, time, dt)
        n += 1
        time = t0 + float(n) * dt
    observer(x, time)
    return n


class ConstIntervalObserver:
    """src/observer.hpp:64-126 (saves field + density spectra; slices optional)."""

    def __init__(self, eq: Equation, t_start, t_end, t_interval, save_density=False,
                 rho_fn=None):
        self.eq = eq
        self.t_start, self.t_end, self.t_interval = t_start, t_end, t_interval
        self.t_last = t_start
        self.save_density = save_density
        self.rho_fn = rho_fn
        self.t_list = []
        self.field_spectra = []
        self.rho_spectra = []
        self.rho_slices = []
        self.rho_axis_averages = []

    def __call__(self, x, t):
        if t >= self.t_last + self.t_interval or t == self.t_end or t == self.t_start:
            N = self.eq.N
            a_t = self.eq.a(t)
            self.field_spectra.append(mode_power_spectrum(N, self.eq.L, self.eq.m, a_t, x))
            rho = self.rho_fn(x, t) if self.rho_fn else self.eq.energy_density(x, t)
            self.rho_spectra.append(power_spectrum(N, rho))
            if self.save_density:
                self.rho_slices.append(rho[: N * N].copy())
                self.rho_axis_averages.append(rho.reshape(N, N * N).mean(axis=0))
            self.t_list.append(t)
            self.t_last = t


# --------------------------------------------------------------------------- FFT / spectra
def fft_r2c(N, f):
    out = np.empty(2 * N * N * (N // 2 + 1))
    lib().orc_fft_r2c(N, _p(np.ascontiguousarray(f, np.float64)), _p(out))
    return out


def fft_c2r(N, fk):
    out = np.empty(N ** 3)
    lib().orc_fft_c2r(N, _p(np.ascontiguousarray(fk, np.float64)), _p(out))
    return out


def dft_r2c_direct(N, f):
    out = np.empty(2 * N * N * (N // 2 + 1))
    lib().orc_dft_r2c_direct(N, _p(np.ascontiguousarray(f, np.float64)), _p(out))
    return out


def bin_power(N, fk):
    out = np.empty(nbins(N))
    lib().orc_bin_power(N, _p(fk), _p(out))
    return out


def power_spectrum(N, f):
    out = np.empty(nbins(N))
    lib().orc_power_spectrum(N, _p(np.ascontiguousarray(f, np.float64)), _p(out))
    return out


def mode_power_spectrum(N, L, m, a_t, state):
    out = np.empty(nbins(N))
    lib().orc_mode_power_spectrum(N, L, m, a_t, _p(np.ascontiguousarray(state, np.float64)), _p(out))
    return out


def radial_bins(N, s_spectrum):
    """Radial product of the s-bins: bin j = isqrt(s), j < N/2 (no reference
    counterpart; plot_util.py:54-74 rebins s offline)."""
    s = np.arange(s_spectrum.size)
    j = np.array([int(np.floor(np.sqrt(v))) for v in s])
    j = np.where((j + 1) ** 2 <= s, j + 1, j)
    j = np.where(j * j > s, j - 1, j)
    out = np.zeros(N // 2)
    keep = j < N // 2
    for jj, v in zip(j[keep], s_spectrum[keep]):
        out[jj] += v
    return out


# --------------------------------------------------------------------------- GRF
SPEC_GAUSS, SPEC_KM3_GAUSS, SPEC_SCALE_INV, SPEC_POWER_CUTOFF, SPEC_DERIV_POWER_CUTOFF = range(5)


def spectrum_table(kind, N, L, params=()):
    out = np.empty(nbins(N))
    prm = np.zeros(8)
    prm[: len(params)] = params
    lib().orc_spectrum_table(kind, N, L, _p(prm), _p(out))
    return out


def normals(seed, n):
    out = np.empty(n)
    lib().orc_normals(seed, n, _p(out))
    return out


def grf_from_normals(N, nrm, P_by_s, Psi=None):
    sq = np.sqrt(P_by_s)
    out = np.empty(N ** 3)
    lib().orc_grf_from_normals(N, _p(nrm), _p(sq), _p(Psi) if Psi is not None else None, _p(out))
    return out


def rescale_rms(x, target):
    lib().orc_rescale_rms(x.size, _p(x), target)
    return x


def initial_state(N, L, seed, kind, params, rms):
    """Workload initial condition: phi = GRF(seed, P) rescaled to `rms`, dphi/dt = 0."""
    nrm = normals(seed, N ** 3)
    phi = grf_from_normals(N, nrm, spectrum_table(kind, N, L, params))
    rescale_rms(phi, rms)
    return np.concatenate([phi, np.zeros(N ** 3)])


def scaled_fourier_modes(N, f, kernel_by_s):
    out = np.empty(N ** 3)
    lib().orc_scaled_fourier_modes(N, _p(np.ascontiguousarray(f)), _p(kernel_by_s), _p(out))
    return out


def ci_pade(x):
    return lib().orc_ci_pade(x)


def comoving_psi_fft(N, R_fft, factor, H_t):
    psi = np.empty_like(R_fft)
    dpsi = np.empty_like(R_fft)
    lib().orc_comoving_psi_fft(N, _p(R_fft), _p(psi), _p(dpsi), factor, H_t)
    return psi, dpsi


def curvature_energy_density(N, x, Psi, m, inv):
    out = np.empty(N ** 3)
    lib().orc_curvature_energy_density(N, _p(x), _p(Psi), _p(out), m, inv)
    return out


def curvature_energy_density_cuda(N, x, Psi, m, inv):
    out = np.empty(N ** 3)
    lib().orc_curvature_energy_density_cuda(N, _p(x), _p(Psi), _p(out), m, inv)
    return out


def cutoff_fouriers(N, M, fk):
    out = np.empty(2 * M * M * (M // 2 + 1))
    lib().orc_cutoff_fouriers(N, M, _p(fk), _p(out))
    return out


def comoving_factor(L, eta_t):
    """(2 pi / L) eta(t) / sqrt(3) -- src/equations.cpp:196."""
    return (2 * np.pi / L) * eta_t / np.sqrt(3.0)


def comoving_psi(G, R_fft, factor, H_t):
    """Psi, dPsi/dt in real space on a G^3 grid: transfer (equations.cpp:141-183)
    then the unnormalised C2R of each (equations.cpp:232-233)."""
    pk, dk = comoving_psi_fft(G, R_fft, factor, H_t)
    return fft_c2r(G, pk), fft_c2r(G, dk)


def wkb_tables(N, L, m, a1, H1, t_i, t_f):
    """cos(phase), sin(phase), sqrt(omega_eff_i), phase per s (wkb.cpp:40-96)."""
    nb = nbins(N)
    out = [np.empty(nb) for _ in range(4)]
    lib().orc_wkb_tables(N, L, m, a1, H1, t_i, t_f, *[_p(o) for o in out])
    return out


class WKB:
    """WKBSolutionForKleinGordonEquationInFRW (wkb.cpp:5-121) for a radiation-era
    cosmology (p = 1, t1 = 1/(2 H1))."""

    def __init__(self, N, L, m, cosmo, state, t_i):
        assert cosmo.p == 1.0 and cosmo.t1 == 1.0 / (2 * cosmo.H1)
        self.N, self.L, self.m, self.cosmo, self.t_i = N, L, m, cosmo, t_i
        self.ffts = np.empty(4 * N * N * (N // 2 + 1))
        lib().orc_wkb_init(N, _p(np.ascontiguousarray(state)), cosmo_a(cosmo, t_i), cosmo_H(cosmo, t_i),
                           _p(self.ffts))

    def evaluate_at(self, t):
        c, s, w, _ = wkb_tables(self.N, self.L, self.m, self.cosmo.a1, self.cosmo.H1, self.t_i, t)
        out = np.empty(2 * self.N ** 3)
        lib().orc_wkb_evaluate(self.N, _p(self.ffts), _p(c), _p(s), _p(w), cosmo_a(self.cosmo, t),
                               cosmo_H(self.cosmo, t), _p(out))
        return out


FIXED, SQRT, COMOVING, APPROX = 1, 2, 3, 4


class VariantEquation:
    """The curvature / potential variants with classic RK4 in the scale_sum order
    (runge_kutta4_classic.hpp:87-138): y = x + dh*k, and
    x <- (((x + dt6*k1) + dt3*k2) + dt3*k3) + dt6*k4, numpy ufuncs (no fused ops).

    kind FIXED: field = Psi (N^3); SQRT: f_a; COMOVING: field = R_fft (half
    spectrum of N^3); APPROX: field = R_fft and the coarse grid M."""

    def __init__(self, kind, N, L, m, cosmo, field=None, f_a=1.0, M=0):
        self.kind, self.N, self.L, self.m, self.cosmo = kind, N, L, m, cosmo
        self.field, self.f_a, self.M = field, f_a, M
        if kind == APPROX:
            self.R_M = cutoff_fouriers(N, M, field)

    def psi(self, t):
        """Psi(t), dPsi/dt(t) on the grid the right-hand side reads."""
        f = comoving_factor(self.L, cosmo_eta(self.cosmo, t))
        H = cosmo_H(self.cosmo, t)
        if self.kind == COMOVING:
            return comoving_psi(self.N, self.field, f, H)
        return comoving_psi(self.M, self.R_M, f, H)

    def rhs(self, x, t):
        N, L, m = self.N, self.L, self.m
        H = cosmo_H(self.cosmo, t)
        inv = inv_ah_sqr(N, L, cosmo_a(self.cosmo, t))
        out = np.empty_like(x)
        Lb = lib()
        if self.kind == FIXED:
            Lb.orc_rhs_fixed_curvature(N, _p(x), _p(out), _p(self.field), m, H, inv)
        elif self.kind == SQRT:
            Lb.orc_rhs_sqrt_potential(N, _p(x), _p(out), m, self.f_a, H, inv)
        elif self.kind == COMOVING:
            ps, dps = self.psi(t)
            Lb.orc_rhs_variable_curvature(N, _p(x), _p(out), _p(ps), _p(dps), m, H, inv)
        else:
            ps, dps = self.psi(t)
            Lb.orc_rhs_approx_curvature(N, self.M, _p(x), _p(out), _p(ps), _p(dps), m, H, inv)
        return out

    def rk4_step(self, x, t, dt):
        dh, dt6, dt3 = dt / 2.0, dt / 6.0, dt / 3.0
        th = t + dh
        k1 = self.rhs(x, t)
        acc = x + dt6 * k1
        k2 = self.rhs(x + dh * k1, th)
        acc = acc + dt3 * k2
        k3 = self.rhs(x + dh * k2, th)
        acc = acc + dt3 * k3
        k4 = self.rhs(x + dt * k3, t + dt)
        return acc + dt6 * k4

    def rk4_run(self, x, t0, dt, n_steps, k0=0):
        for k in range(n_steps):
            x = self.rk4_step(x, t0 + float(k0 + k) * dt, dt)
        return x

    def energy_density(self, x, t):
        N, m = self.N, self.m
        inv = inv_ah_sqr(N, self.L, cosmo_a(self.cosmo, t))
        if self.kind == SQRT:
            out = np.empty(N ** 3)
            lib().orc_energy_density_sqrt(N, _p(x), _p(out), m, self.f_a, inv)
            return out
        if self.kind == FIXED:
            return curvature_energy_density_cuda(N, x, self.field, m, inv)
        f = comoving_factor(self.L, cosmo_eta(self.cosmo, t))
        ps = fft_c2r(N, comoving_psi_fft(N, self.field, f, cosmo_H(self.cosmo, t))[0])
        return curvature_energy_density(N, x, ps, m, inv)


if __name__ == "__main__":  # pragma: no cover
    print("oracle library:", lib(), file=sys.stderr)
