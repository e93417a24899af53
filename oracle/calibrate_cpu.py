"""CPU calibration of the oracle port against the reference's own integrator.

TEST INFRASTRUCTURE ONLY. Times, on one core, FP64 RK4 steps of KG-FRW
(m=1, L=N, radiation era a1=1, H1=0.5, t1=1, dt=0.05) at N=64 and N=128:
  * reference: oracle/_ref/odeint_ref -- the reference's vendored Boost.odeint
    integrate_const + runge_kutta4_classic and src/odeint_eigen/eigen_operations.hpp
    (its scale_sum passes) around the oracle's right-hand side; the step loop is
    timed inside the binary with std::chrono::steady_clock;
  * port: the oracle's own RK4 (cf_oracle.Equation.rk4_run), timed around the call.
Prints GSite*RK4-step/s of both and their ratio. Needs `make -C oracle ref`.

    python oracle/calibrate_cpu.py [N ...]
"""
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import cf_oracle as O  # noqa: E402

REF_BIN = os.path.join(HERE, "_ref", "odeint_ref")
STEPS = {64: 200, 128: 40}


def main(sizes):
    if not os.path.exists(REF_BIN):
        sys.exit(f"{REF_BIN} missing: build it with `make -C oracle ref`")
    O.set_threads(1)
    for N in sizes:
        steps = STEPS.get(N, 2)
        rng = np.random.default_rng(7)
        x = np.concatenate([rng.standard_normal(N ** 3), np.zeros(N ** 3)])
        t0, dt = 1.0, 0.05
        with tempfile.TemporaryDirectory() as d:
            x.tofile(os.path.join(d, "in.bin"))
            r = subprocess.run([REF_BIN, os.path.join(d, "in.bin"), d, str(N), repr(float(N)), "1.0", "0.0",
                                "1.0", "0.5", "1.0", repr(t0), repr(t0 + steps * dt), repr(dt)],
                               capture_output=True, text=True, check=True, env=dict(os.environ, OMP_NUM_THREADS="1"))
            el_ref = float(r.stderr.split("elapsed_s")[1].split()[0])
            n_ref = int(r.stdout)
            ref_state = np.fromfile(os.path.join(d, "state.bin"))
        eq = O.Equation(N, float(N), 1.0, 0.0, O.Cosmo.radiation(1.0, 0.5, 1.0))
        eq._w(5).fill(0.0)
        xp = x.copy()
        t = time.perf_counter()
        eq.rk4_run(xp, t0, dt, n_ref)
        el_port = time.perf_counter() - t
        same = np.array_equal(xp.view(np.uint64), ref_state.view(np.uint64))
        g_ref = N ** 3 * n_ref / el_ref / 1e9
        g_port = N ** 3 * n_ref / el_port / 1e9
        print(f"N={N} steps={n_ref}: reference {g_ref:.4f} GSite*step/s ({el_ref:.2f} s), "
              f"port {g_port:.4f} ({el_port:.2f} s), port/reference {g_port / g_ref:.2f}, "
              f"final states bit-identical: {same}", flush=True)


if __name__ == "__main__":
    main([int(a) for a in sys.argv[1:]] or [64, 128])
