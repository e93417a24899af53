"""The C++ host layer (cosmic-fields-lite_amd/host/cfl) on the GPU.

kg_run drives WorkspaceGeneric + HIP equation + ConstIntervalObserver +
integrate_const exactly as the reference's main.cpp does; its outputs are
checked against the oracle's restatement of the same run: final state bit for
bit, save times exactly, spectra to FFT rounding. odeint_hip_dropin runs the HIP
equation under the reference's own Boost.odeint (built only where the reference
tree is mounted).
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

KG_RUN = os.path.join(ROOT, "cosmic-fields-lite_amd", "bin", "kg_run")
DROPIN = os.path.join(ROOT, "oracle", "_ref", "odeint_hip_dropin")


def spec_close(g, r, tol=1e-10, floor=1e-18):
    assert np.all(g[r == 0] == 0)
    assert np.all(np.abs(g - r) <= tol * np.abs(r) + floor * np.abs(r).max())


def run_kg(tmp_path, stepper, eq, N=32, L=40.0, m=1.0, lam=0.0, t0=1.0, t1=3.0, interval=0.5, dt=0.05,
           seed=1, kind=0, sp=0.5, rms=0.5):
    out = str(tmp_path) + "/"
    a1, H1, tt1 = (1.0, 0.5, 1.0) if eq != "kg" else (1.0, 0.0, 0.0)
    args = [KG_RUN, out, stepper, eq, str(N), repr(L), repr(m), repr(lam), repr(a1), repr(H1), repr(tt1),
            repr(t0), repr(t1), repr(interval), repr(dt), str(seed), str(kind), repr(sp), repr(rms)]
    print("case", " ".join(args[2:]), flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return out, dict(N=N, L=L, m=m, lam=lam, t0=t0, t1=t1, interval=interval, dt=dt, a1=a1, H1=H1, t1c=tt1,
                     eq=eq)


def oracle_run(O, out, p, f32c=False, f32s=False):
    N = p["N"]
    x = np.fromfile(out + "state0.dat")
    cosmo = O.Cosmo.radiation(p["a1"], p["H1"], p["t1c"]) if p["H1"] != 0 else O.Cosmo.minkowski()
    if p.get("eq") == "sqrt":
        # the sqrt-potential variant (the lambda argument is f_a); the KG equation
        # only supplies a(t) and the lattice constants to the observer
        eq = O.Equation(N, p["L"], p["m"], 0.0, cosmo)
        v = O.VariantEquation(O.SQRT, N, p["L"], p["m"], cosmo, f_a=p["lam"])
        obs = O.ConstIntervalObserver(eq, p["t0"], p["t1"], p["interval"], save_density=True,
                                      rho_fn=v.energy_density)

        def vstep(xx, t, dt):
            xx[:] = v.rk4_step(xx, t, dt)

        O.integrate_const(vstep, obs, x, p["t0"], p["t1"], p["dt"])
        return x, obs
    eq = O.Equation(N, p["L"], p["m"], p["lam"], cosmo)
    obs = O.ConstIntervalObserver(eq, p["t0"], p["t1"], p["interval"], save_density=True)
    if f32c or f32s:
        # the driver splits the state once and observes hi + lo from then on,
        # including at t_start
        hi = x.astype(np.float32)
        lo = (x - hi.astype(np.float64)).astype(np.float32)
        x[:] = hi.astype(np.float64) + lo.astype(np.float64)
        comp_step = eq.rk4_step_f32c if f32c else eq.rk4_step_f32s

        def step(_, t, dt):
            comp_step(hi, lo, t, dt)
            x[:] = hi.astype(np.float64) + lo.astype(np.float64)

        O.integrate_const(step, obs, x, p["t0"], p["t1"], p["dt"])
    else:
        O.integrate_const(eq.rk4_step, obs, x, p["t0"], p["t1"], p["dt"])
    return x, obs


def test_wkb_extrapolation_of_final_state(O, tmp_path):
    """cfl::WKBSolutionForKleinGordonEquationInFRW from the final state of an FRW run,
    evaluated at t_end + 10, against the oracle's WKB from the same state (FFT
    rounding apart: <= 1e-12 of the field's maximum)."""
    out, p = run_kg(tmp_path, "fused", "frw")
    x = np.fromfile(out + "state.dat")
    w = O.WKB(p["N"], p["L"], p["m"], O.Cosmo.radiation(p["a1"], p["H1"], p["t1c"]), x, p["t1"])
    ref = w.evaluate_at(p["t1"] + 10.0)
    g = np.fromfile(out + "wkb_state.dat")
    assert np.abs(g - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("stepper,eq,lam", [("classic", "frw", 0.0), ("fused", "frw", 0.0),
                                            ("fused", "lambda", 0.1), ("fused", "kg", 0.0),
                                            ("f32c", "frw", 0.0), ("f32s", "frw", 0.0), ("f32s", "lambda", 0.1),
                                            ("classic", "sqrt", 0.7), ("fused", "sqrt", 0.7)])
def test_kg_run_matches_oracle(O, tmp_path, stepper, eq, lam):
    out, p = run_kg(tmp_path, stepper, eq, lam=lam)
    x, obs = oracle_run(O, out, p, f32c=(stepper == "f32c"), f32s=(stepper == "f32s"))
    g = np.fromfile(out + "state.dat")
    assert np.array_equal(g, x), f"final state differs ({np.count_nonzero(g != x)} entries)"
    np.testing.assert_array_equal(np.fromfile(out + "t_list.dat"), np.array(obs.t_list))
    for i in range(len(obs.t_list)):
        spec_close(np.fromfile(out + f"varphi_plus_spectrum_{i}.dat"), obs.field_spectra[i])
        spec_close(np.fromfile(out + f"rho_spectrum_{i}.dat"), obs.rho_spectra[i])
        check_density_files(out, obs, i)
        check_radial_files(out, obs, i, p["N"])


def check_radial_files(out, obs, i, N):
    """N/2 radial bins of each saved spectrum: sums of the s-bins in order, so they
    follow the s-bins' own tolerance."""
    spec_close(np.fromfile(out + f"varphi_plus_radial_{i}.dat"), O_radial(N, obs.field_spectra[i]))
    spec_close(np.fromfile(out + f"rho_radial_{i}.dat"), O_radial(N, obs.rho_spectra[i]))


def O_radial(N, s):
    import cf_oracle
    return cf_oracle.radial_bins(N, s)


def check_density_files(out, obs, i):
    """rho slice at a = 0 and the average over a: bit for bit (the device average
    sums over a in order and divides once, as the reference's rowwise mean)."""
    assert np.array_equal(np.fromfile(out + f"rho_slice_{i}.dat"), obs.rho_slices[i])
    assert np.array_equal(np.fromfile(out + f"rho_axis_average_{i}.dat"), obs.rho_axis_averages[i])


@pytest.mark.skipif(not os.path.exists(DROPIN), reason="built only where the reference tree is mounted")
@pytest.mark.parametrize("lam", [0.0, 0.1])
def test_boost_odeint_dropin(O, tmp_path, lam):
    N, L, m, t0, t1, dt = 24, 30.0, 1.0, 1.0, 2.0, 0.05
    x = O.initial_state(N, L, 4, O.SPEC_GAUSS, (0.5,), 0.4)
    x.tofile(str(tmp_path / "in.bin"))
    args = [DROPIN, str(tmp_path / "in.bin"), str(tmp_path), str(N), repr(L), repr(m), repr(lam), "1.0", "0.5",
            "1.0", repr(t0), repr(t1), repr(dt)]
    print("case dropin lam", lam, flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    eq = O.Equation(N, L, m, lam, O.Cosmo.radiation(1.0, 0.5, 1.0))
    times = []
    n = O.integrate_const(eq.rk4_step, lambda xx, t: times.append(t), x, t0, t1, dt)
    assert int(r.stdout.strip()) == n
    assert np.array_equal(np.fromfile(str(tmp_path / "state.bin")), x)
    np.testing.assert_array_equal(np.fromfile(str(tmp_path / "times.bin")), np.array(times))


def _check_run(O, out, p, n_saves):
    x, obs = oracle_run(O, out, p)
    g = np.fromfile(out + "state.dat")
    n = p["N"] ** 3
    assert np.array_equal(g, x), (f"final state differs in {np.count_nonzero(g != x)} entries; rel L2 phi "
                                  f"{np.linalg.norm(g[:n] - x[:n]) / np.linalg.norm(x[:n]):.3e}")
    np.testing.assert_array_equal(np.fromfile(out + "t_list.dat"), np.array(obs.t_list))
    assert len(obs.t_list) == n_saves
    for i in range(len(obs.t_list)):
        spec_close(np.fromfile(out + f"varphi_plus_spectrum_{i}.dat"), obs.field_spectra[i])
        spec_close(np.fromfile(out + f"rho_spectrum_{i}.dat"), obs.rho_spectra[i])
        check_density_files(out, obs, i)
        check_radial_files(out, obs, i, p["N"])


def test_config_kg_n64_minkowski(O, tmp_path):
    """Plain Klein-Gordon, N=64, L=64, m=1, seed 0, P ~ exp(-k^2/(2*0.5^2)), rms 1,
    dt=0.1 from t=0 to 10, spectra every 1.0 (11 saves): the fused HIP run equals
    the oracle's restatement of the reference CPU run bit for bit."""
    out, p = run_kg(tmp_path, "fused", "kg", N=64, L=64.0, m=1.0, t0=0.0, t1=10.0, interval=0.999, dt=0.1,
                    seed=0, kind=0, sp=0.5, rms=1.0)
    _check_run(O, out, p, 11)


def test_config_frw_n256_oracle(O, tmp_path):
    """KG in FRW (radiation era), N=256, L=256, m=1, seed 1, P ~ k^-3 exp(-k^2/2),
    rms 0.1, dt=0.05 from t=1: the FP64 HIP drop-in run of the first 40 steps with
    saves every 20 equals the oracle bit for bit (state) and to 1e-10 (spectra).
    The oracle needs ~0.4 s per step at 256^3, so the full 2000 steps are checked
    on the GPU by the test below."""
    out, p = run_kg(tmp_path, "fused", "frw", N=256, L=256.0, m=1.0, t0=1.0, t1=3.0, interval=0.999,
                    dt=0.05, seed=1, kind=1, sp=0.0, rms=0.1)
    print("GPU run done; oracle run (40 steps at 256^3) follows", flush=True)
    _check_run(O, out, p, 3)


def test_config_frw_n256_2000_steps(cfl):
    """The same workload for all 2000 steps, states compared every 100 steps between
    three independent GPU evaluations of the reference's step: the odeint flow
    (sys() + scale_sum2/scale_sum5, as the reference's GPU path runs it), the four
    fused stage kernels, and the temporally blocked one-kernel step."""
    import torch
    N, L, dt, t0 = 256, 256.0, 0.05, 1.0
    cosmo = cfl.radiation(1.0, 0.5, 1.0)
    p = cfl.kg_params(N, L, 1.0, 0.0, cosmo)
    plan = cfl.SpectrumPlan(N)
    x0 = plan.init_grf_state(1, cfl.CF_SPEC_KM3_GAUSS, (), L, 0.1)
    xf, xt, xc = x0.clone(), x0.clone(), x0.clone()
    wf, wt = cfl.alloc_work(N), cfl.alloc_tb_work(N)
    k = [torch.empty_like(x0) for _ in range(4)]
    y = torch.empty_like(x0)

    def f(xx, tt, out):
        a = cfl.cosmo_a(cosmo, tt)
        cfl.rhs_f64(N, xx, out, 1.0, 0.0, cfl.cosmo_H(cosmo, tt), cfl.inv_ah_sqr(N, L, a))

    for block in range(20):
        for j in range(100):
            t = t0 + float(block * 100 + j) * dt
            dh = dt / 2.0
            f(xc, t, k[0])
            cfl.scale_sum(y, [xc, k[0]], [1.0, dh])
            f(y, t + dh, k[1])
            cfl.scale_sum(y, [xc, k[1]], [1.0, dh])
            f(y, t + dh, k[2])
            cfl.scale_sum(y, [xc, k[2]], [1.0, dt])
            f(y, t + dt, k[3])
            cfl.scale_sum(xc, [xc] + k, [1.0, dt / 6, dt / 3, dt / 3, dt / 6])
        cfl.rk4_run_f64(p, xf, t0, block * 100, 100, dt, wf)
        cfl.rk4_run_tb_f64(p, xt, t0, block * 100, 100, dt, wt)
        torch.cuda.synchronize()
        assert torch.equal(xf, xc), f"fused vs odeint flow differ after {(block + 1) * 100} steps"
        assert torch.equal(xt, xc), f"temporally blocked vs odeint flow differ after {(block + 1) * 100} steps"
        print(f"step {(block + 1) * 100}: identical", flush=True)
    assert torch.isfinite(xc).all()


REF_STYLE = os.path.join(ROOT, "cosmic-fields-lite_amd", "bin", "reference_style_equation")


@pytest.mark.parametrize("N,L", [(32, 40.0), (48, 38.4)])
def test_reference_style_comoving_equation(N, L):
    """An equation written the reference's way against workspace.fft_wrapper
    (execute_inplace_z2d + PADDED_IDX_OF, equations_cuda.cu:678-731) gives the
    library's comoving-curvature dx/dt to 1e-12 of its maximum."""
    print(f"case reference_style_equation N={N}", flush=True)
    r = subprocess.run([REF_STYLE, str(N), repr(L), "1.0", "1.0", "0.05", "0.02"], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr


FINITE = os.path.join(ROOT, "cosmic-fields-lite_amd", "bin", "observer_finite_check")


def test_observer_finite_check(tmp_path):
    """The observer's opt-in finite check: a clean state saves; planted inf / NaN /
    -inf values raise cfl::NonFiniteState with their count, also out of
    integrate_const."""
    print("case observer_finite_check", flush=True)
    r = subprocess.run([FINITE, str(tmp_path) + "/"], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("ok ") == 5


def test_count_nonfinite(cfl):
    import torch
    x = torch.zeros(1 << 20, dtype=torch.float64, device="cuda")
    assert cfl.count_nonfinite(x) == 0
    x[0] = float("nan")
    x[-1] = float("inf")
    x[12345] = -float("inf")
    x[7] = 1e308
    assert cfl.count_nonfinite(x) == 3
