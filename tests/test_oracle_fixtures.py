"""The oracle against the reference's own shipped outputs (tests/golden/).

The reference run (src/main.cpp:82-196, N=384) wrote spectra, density slices and
the save times at t_start. Recomputing them with the oracle pins the Gaussian
random field stream (std::mt19937 + std::normal_distribution, seed 0), the
spectrum factories, the Ci Pade A_s formula, the FFT conventions, both binning
functions, the comoving-curvature energy density and the observer cadence.

Agreement is to FFT rounding (the files came from FFTW under -ffast-math):
per bin |oracle - file| <= 1e-10*|file| + 1e-18*max(file). The DC bin of a random
field with P(0) = 0 holds roundoff only, hence the absolute floor.
"""
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.slow


def load(name):
    return np.fromfile(os.path.join(GOLDEN, name), dtype="<f8")


def read_param():
    """param.dat: POD bytes of MyParam (main.cpp:36-64), types per paramTypes.txt."""
    names = open(os.path.join(GOLDEN, "paramNames.txt")).read().split()
    types = open(os.path.join(GOLDEN, "paramTypes.txt")).read().split()
    raw = open(os.path.join(GOLDEN, "param.dat"), "rb").read()
    assert len(raw) == 8 * len(names) == 8 * len(types)
    out = {}
    for i, (n, t) in enumerate(zip(names, types)):
        dt = "<i8" if t == "Integer64" else "<f8"
        out[n] = np.frombuffer(raw[8 * i: 8 * i + 8], dtype=dt)[0].item()
    return out


def close(ours, ref, what, tol=1e-10, floor=1e-18):
    assert ours.shape == ref.shape, what
    err = np.abs(ours - ref)
    bound = tol * np.abs(ref) + floor * np.abs(ref).max()
    bad = err > bound
    nz = ref != 0
    rel = np.abs(ours[nz] - ref[nz]) / np.abs(ref[nz])
    assert not bad.any(), f"{what}: {bad.sum()} bins off; median rel {np.median(rel):.2e}"
    return float(np.median(rel))


def oracle_initial_run(O, P):
    """Restates the reference run's initialisation (initializer.hpp:150-206) and
    its first observation (observer.hpp:87-124 at t = t_start) for the parameter
    dict P (the fields of main.cpp:36-64); seed 0."""
    N, L = P["N"], P["L"]
    pi = math.pi
    cosmo = O.Cosmo.radiation(P["a1"], P["H1"], P["t1"])
    t0 = P["t_start"]
    eta_i = O.cosmo_eta(cosmo, t0)
    # A_s, initializer.hpp:156-167 (same operation order)
    p, sq3, Ls, sd = math.pow, math.sqrt(3), L, P["Psi_std_dev"]
    xs = (4 * pi * eta_i) / (sq3 * Ls)
    A_s = (-576 * p(pi, 6) * p(eta_i, 6) * p(sd, 2)) / (
        -81 * p(Ls, 4) * (p(Ls, 2) + 2 * p(pi, 2) * p(eta_i, 2))
        + Ls * (81 * p(Ls, 5) - 54 * p(Ls, 3) * p(pi, 2) * p(eta_i, 2) + 48 * Ls * p(pi, 4) * p(eta_i, 4))
        * math.cos(xs)
        + 256 * p(pi, 6) * p(eta_i, 6) * O.ci_pade(xs)
        + 4 * sq3 * Ls * pi * eta_i * (27 * p(Ls, 4) + 6 * p(Ls, 2) * p(pi, 2) * p(eta_i, 2)
                                       - 16 * p(pi, 4) * p(eta_i, 4)) * math.sin(xs))
    n3 = N ** 3
    nrm = O.normals(0, 3 * n3)  # one generator stream: R, varphi, dt_varphi
    R = O.grf_from_normals(N, nrm[:n3], O.spectrum_table(O.SPEC_SCALE_INV, N, L, (A_s,)))

    # Psi kernel, initializer.hpp:177-179, evaluated at k = sqrt(s)*(2*pi/L) (fdm3d.cpp:94)
    def kernel(k):
        if k == 0.0:
            return 0.0
        return (6 * sq3 * (-((k * eta_i * math.cos((k * eta_i) / sq3)) / sq3) + math.sin((k * eta_i) / sq3))) / (
            math.pow(k, 3) * math.pow(eta_i, 3))

    kern = np.array([kernel(math.sqrt(float(s)) * (2 * pi / L)) for s in range(O.nbins(N))])
    Psi = O.scaled_fourier_modes(N, R, kern)
    v = P["k_ast"] / (P["a1"] * P["m"])
    den = 12 * p(v, 4) + 50 * p(v, 2) + 50
    alpha_phi = 0.5 * (-3 * p(4 * p(v, 2) + 5, 2)) / den
    alpha_dphi = 0.5 * (25 - 20 * p(v, 2)) / den
    Pf = O.spectrum_table(O.SPEC_POWER_CUTOFF, N, L, (P["varphi_std_dev"], P["k_ast"], 0.0))
    Pdf = O.spectrum_table(O.SPEC_DERIV_POWER_CUTOFF, N, L, (P["varphi_std_dev"], P["k_ast"], 0.0, P["m"]))
    phi = O.grf_from_normals(N, nrm[n3: 2 * n3], Pf, alpha_phi * Psi)
    dphi = O.grf_from_normals(N, nrm[2 * n3:], Pdf, alpha_dphi * Psi)
    del nrm
    R_fft = O.fft_r2c(N, R)
    out = {"P": P, "N": N, "L": L, "cosmo": cosmo, "eta_i": eta_i, "R_fft": R_fft}
    # main.cpp:164-176: R round-tripped through z2d/N^3, Psi recomputed from it
    R2 = O.fft_c2r(N, R_fft) / float(N ** 3)
    out["R_spectrum"] = O.power_spectrum(N, R2)
    out["Psi_spectrum"] = O.power_spectrum(N, O.scaled_fourier_modes(N, R2, kern))
    state = np.concatenate([phi, dphi])
    del phi, dphi, R, R2, Psi
    out["state"] = state
    a_t = O.cosmo_a(cosmo, t0)
    out["field_spectrum"] = O.mode_power_spectrum(N, L, P["m"], a_t, state)
    # ComovingCurvatureEquationInFRW::compute_energy_density (equations.cpp:262-322)
    H_t, eta_t = O.cosmo_H(cosmo, t0), O.cosmo_eta(cosmo, t0)
    factor = (2 * pi / L) * eta_t / math.sqrt(3)
    psi_fft, _ = O.comoving_psi_fft(N, R_fft, factor, H_t)
    Psi_t = O.fft_c2r(N, psi_fft)
    rho = O.curvature_energy_density(N, state, Psi_t, P["m"], O.inv_ah_sqr(N, L, a_t))
    out["rho_spectrum"] = O.power_spectrum(N, rho)
    out["rho_slice"] = rho[: N * N].copy()
    out["rho_axis_average"] = rho.reshape(N, N * N).mean(axis=0)
    return out


@pytest.fixture(scope="module")
def run(O):
    return oracle_initial_run(O, read_param())


def test_param_file(run):
    P = run["P"]
    assert P["N"] == 384 and P["L"] == 384 * 0.8 and P["t_start"] == 10.0 and P["M"] == 128
    assert run["eta_i"] == 20.0


def test_R_spectrum(run):
    med = close(run["R_spectrum"], load("initial_R_spectrum.dat"), "initial_R_spectrum")
    assert med < 1e-12


def test_Psi_spectrum(run):
    close(run["Psi_spectrum"], load("initial_Psi_spectrum.dat"), "initial_Psi_spectrum")


def test_field_spectrum_t0(run):
    close(run["field_spectrum"], load("varphi_plus_spectrum_0.dat"), "varphi_plus_spectrum_0")


def test_rho_spectrum_t0(run):
    close(run["rho_spectrum"], load("rho_spectrum_0.dat"), "rho_spectrum_0")


def test_rho_slice_and_axis_average_t0(run):
    np.testing.assert_allclose(run["rho_slice"], load("rho_slice_0.dat"), rtol=1e-10)
    np.testing.assert_allclose(run["rho_axis_average"], load("rho_axis_average_0.dat"), rtol=1e-10)


def test_observer_cadence_matches_t_list(O):
    """integrate_const + ConstIntervalObserver save times (t_list.dat)."""
    P = read_param()
    times = []

    class Obs:
        t_last = P["t_start"]

        def __call__(self, x, t):
            # observer.hpp:89
            if t >= self.t_last + P["t_interval"] or t == P["t_end"] or t == P["t_start"]:
                times.append(t)
                self.t_last = t

    n = O.integrate_const(lambda x, t, dt: None, Obs(), None, P["t_start"], P["t_end"], P["delta_t"])
    assert n == 225
    np.testing.assert_array_equal(np.array(times), load("t_list.dat"))
