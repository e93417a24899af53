"""The HIP path against the reference's own shipped outputs (tests/golden/, N=384).

The reference run (src/main.cpp:82-196) initialised the comoving curvature R,
its Psi, and a perturbed field from one std::mt19937 stream (seed 0;
initializer.hpp:150-206) and observed at t_start. Here the whole initialisation
and observation runs in libcfhip.so -- Gaussian random fields (cf_grf), the
scaled-mode transform (cf_scaled_fourier_modes), the mode and power spectra, and
the comoving-curvature energy density (cf_variant_energy_density with
CF_VAR_COMOVING_CURVATURE) -- and each spectrum file is matched per bin to
|gpu - file| <= 1e-10*|file| + 1e-18*max(file) (the files came from FFTW under
-ffast-math; the floor covers the DC bins, which hold roundoff only).
"""
import math
import os

import numpy as np
import pytest
import torch

from conftest import GOLDEN
from test_oracle_fixtures import close, load, read_param

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu_run(cfl):
    P = read_param()
    N, L = P["N"], P["L"]
    pi = math.pi
    cosmo_c = cfl.radiation(P["a1"], P["H1"], P["t1"])
    t0 = P["t_start"]
    eta_i = cfl.cosmo_eta(cosmo_c, t0)
    # A_s, initializer.hpp:156-167 (Ci by its Pade approximant, special_function.hpp:17-26)
    p, sq3, Ls, sd = math.pow, math.sqrt(3), L, P["Psi_std_dev"]
    xs = (4 * pi * eta_i) / (sq3 * Ls)
    ci = math.log(xs) + (0.5772156649015329 - 0.24231497614160186 * p(xs, 2) + 0.007139183039136621 * p(xs, 4)
                         - 0.00011466618094101764 * p(xs, 6) + 8.443734405201243e-7 * p(xs, 8)
                         - 3.060472574705558e-9 * p(xs, 10) + 4.328624073851291e-12 * p(xs, 12)) / (
        1. + 0.013313955815300189 * p(xs, 2) + 0.00008836441800952094 * p(xs, 4) + 3.800404484365274e-7 * p(xs, 6)
        + 1.1376490214488613e-9 * p(xs, 8) + 2.297129602871981e-12 * p(xs, 10) + 2.510407760855278e-15 * p(xs, 12))
    A_s = (-576 * p(pi, 6) * p(eta_i, 6) * p(sd, 2)) / (
        -81 * p(Ls, 4) * (p(Ls, 2) + 2 * p(pi, 2) * p(eta_i, 2))
        + Ls * (81 * p(Ls, 5) - 54 * p(Ls, 3) * p(pi, 2) * p(eta_i, 2) + 48 * Ls * p(pi, 4) * p(eta_i, 4))
        * math.cos(xs)
        + 256 * p(pi, 6) * p(eta_i, 6) * ci
        + 4 * sq3 * Ls * pi * eta_i * (27 * p(Ls, 4) + 6 * p(Ls, 2) * p(pi, 2) * p(eta_i, 2)
                                       - 16 * p(pi, 4) * p(eta_i, 4)) * math.sin(xs))
    n3 = N ** 3
    print(f"case fixtures N={N}: normals", flush=True)
    nrm = cfl.normals_host(0, 3 * n3)  # one generator stream: R, varphi, dt_varphi
    scale = 1.0 / float(n3)            # random_field.cpp:44-46

    def white(i):
        return torch.from_numpy(scale * nrm[i * n3:(i + 1) * n3]).cuda()

    def sqrtP(kind, params):
        return torch.from_numpy(np.sqrt(cfl.spectrum_table_host(kind, N, L, params))).cuda()

    plan = cfl.SpectrumPlan(N)
    R = plan.grf(white(0), sqrtP(cfl.CF_SPEC_SCALE_INV, (A_s,)))

    def kernel(k):  # initializer.hpp:177-179
        if k == 0.0:
            return 0.0
        return (6 * sq3 * (-((k * eta_i * math.cos((k * eta_i) / sq3)) / sq3) + math.sin((k * eta_i) / sq3))) / (
            math.pow(k, 3) * math.pow(eta_i, 3))

    kern = torch.tensor([kernel(math.sqrt(float(s)) * (2 * pi / L)) for s in range(cfl.nbins(N))],
                        dtype=torch.float64, device="cuda")
    Psi = plan.scaled_fourier_modes(R, kern)
    v = P["k_ast"] / (P["a1"] * P["m"])
    den = 12 * p(v, 4) + 50 * p(v, 2) + 50
    alpha_phi = 0.5 * (-3 * p(4 * p(v, 2) + 5, 2)) / den
    alpha_dphi = 0.5 * (25 - 20 * p(v, 2)) / den
    state = torch.empty(2 * n3, dtype=torch.float64, device="cuda")
    plan.grf(white(1), sqrtP(cfl.CF_SPEC_POWER_CUTOFF, (P["varphi_std_dev"], P["k_ast"], 0.0)),
             alpha_phi * Psi, out=state[:n3])
    plan.grf(white(2), sqrtP(cfl.CF_SPEC_DERIV_POWER_CUTOFF, (P["varphi_std_dev"], P["k_ast"], 0.0, P["m"])),
             alpha_dphi * Psi, out=state[n3:])
    del nrm
    out = {}
    R_fft = plan.d2z(R)
    # main.cpp:164-176: R round-tripped through z2d / N^3, Psi recomputed from it
    R2 = plan.z2d(R_fft.clone()) / float(n3)
    out["R_spectrum"] = plan.power_spectrum(R2).cpu().numpy()
    out["Psi_spectrum"] = plan.power_spectrum(plan.scaled_fourier_modes(R2, kern)).cpu().numpy()
    a_t = cfl.cosmo_a(cosmo_c, t0)
    out["field_spectrum"] = plan.mode_power_spectrum(state, L, P["m"], a_t).cpu().numpy()
    prm = cfl.kg_params(N, L, P["m"], 0.0, cosmo_c)
    var = cfl.Variant(prm, cfl.CF_VAR_COMOVING_CURVATURE, R_fft)
    rho = var.energy_density(state, t0)
    out["rho_spectrum"] = plan.power_spectrum(rho).cpu().numpy()
    out["rho_slice"] = rho[: N * N].cpu().numpy()
    var.close()
    torch.cuda.synchronize()
    print("case fixtures: computed", flush=True)
    return out


def test_R_spectrum(gpu_run):
    close(gpu_run["R_spectrum"], load("initial_R_spectrum.dat"), "R spectrum")


def test_Psi_spectrum(gpu_run):
    close(gpu_run["Psi_spectrum"], load("initial_Psi_spectrum.dat"), "Psi spectrum")


def test_field_spectrum_t0(gpu_run):
    close(gpu_run["field_spectrum"], load("varphi_plus_spectrum_0.dat"), "field spectrum t0")


def test_rho_spectrum_t0(gpu_run):
    close(gpu_run["rho_spectrum"], load("rho_spectrum_0.dat"), "rho spectrum t0")


def test_rho_slice_t0(gpu_run):
    ref = load("rho_slice_0.dat")
    assert np.abs(gpu_run["rho_slice"] - ref).max() <= 1e-10 * np.abs(ref).max()
