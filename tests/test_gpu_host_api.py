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


def oracle_run(O, out, p, f32c=False):
    N = p["N"]
    x = np.fromfile(out + "state0.dat")
    cosmo = O.Cosmo.radiation(p["a1"], p["H1"], p["t1c"]) if p["H1"] != 0 else O.Cosmo.minkowski()
    if p.get("eq") == "sqrt":
        # the sqrt-potential variant (the lambda argument is f_a); the KG equation
        # only supplies a(t) and the lattice constants to the observer
        eq = O.Equation(N, p["L"], p["m"], 0.0, cosmo)
        v = O.VariantEquation(O.SQRT, N, p["L"], p["m"], cosmo, f_a=p["lam"])
        obs = O.ConstIntervalObserver(eq, p["t0"], p["t1"], p["interval"], rho_fn=v.energy_density)

        def vstep(xx, t, dt):
            xx[:] = v.rk4_step(xx, t, dt)

        O.integrate_const(vstep, obs, x, p["t0"], p["t1"], p["dt"])
        return x, obs
    eq = O.Equation(N, p["L"], p["m"], p["lam"], cosmo)
    obs = O.ConstIntervalObserver(eq, p["t0"], p["t1"], p["interval"])
    if f32c:
        hi = x.astype(np.float32)
        lo = (x - hi.astype(np.float64)).astype(np.float32)

        def step(_, t, dt):
            eq.rk4_step_f32c(hi, lo, t, dt)
            x[:] = hi.astype(np.float64) + lo.astype(np.float64)

        O.integrate_const(step, obs, x, p["t0"], p["t1"], p["dt"])
    else:
        O.integrate_const(eq.rk4_step, obs, x, p["t0"], p["t1"], p["dt"])
    return x, obs


@pytest.mark.parametrize("stepper,eq,lam", [("classic", "frw", 0.0), ("fused", "frw", 0.0),
                                            ("fused", "lambda", 0.1), ("fused", "kg", 0.0),
                                            ("f32c", "frw", 0.0), ("classic", "sqrt", 0.7),
                                            ("fused", "sqrt", 0.7)])
def test_kg_run_matches_oracle(O, tmp_path, stepper, eq, lam):
    out, p = run_kg(tmp_path, stepper, eq, lam=lam)
    x, obs = oracle_run(O, out, p, f32c=(stepper == "f32c"))
    g = np.fromfile(out + "state.dat")
    assert np.array_equal(g, x), f"final state differs ({np.count_nonzero(g != x)} entries)"
    np.testing.assert_array_equal(np.fromfile(out + "t_list.dat"), np.array(obs.t_list))
    for i in range(len(obs.t_list)):
        spec_close(np.fromfile(out + f"varphi_plus_spectrum_{i}.dat"), obs.field_spectra[i])
        spec_close(np.fromfile(out + f"rho_spectrum_{i}.dat"), obs.rho_spectra[i])


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
