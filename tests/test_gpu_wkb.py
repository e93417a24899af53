"""WKB extrapolation (cf_wkb_*) against the oracle's restatement of
WKBSolutionForKleinGordonEquationInFRW (wkb.cpp:5-121).

The per-s tables come from the library's host code in the reference's operation
order and must equal the oracle's bit for bit; the extrapolated state differs
from the oracle only by FFT rounding (hipFFT vs the oracle's transform):
|gpu - oracle| <= 1e-12 max|oracle|. Evaluating at t_i returns the state."""
import ctypes as C

import numpy as np
import pytest
import torch

from test_gpu_parity import dev, host

pytestmark = pytest.mark.gpu


def test_tables_bitwise(cfl, O):
    N, L, m, a1, H1, t_i = 24, 19.2, 1.0, 1.0, 0.05, 10.0
    for t_f in (10.0, 37.5, 160.0):
        nb = cfl.nbins(N)
        g = [np.empty(nb) for _ in range(4)]
        cfl.check(cfl.lib.cf_wkb_tables_host(N, L, m, a1, H1, t_i, t_f,
                                             *[x.ctypes.data_as(C.c_void_p) for x in g]))
        r = O.wkb_tables(N, L, m, a1, H1, t_i, t_f)
        for a, b in zip(g, r):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("N", [16, 24])
def test_evaluate_vs_oracle(cfl, O, N):
    L, m, H1 = 0.8 * N, 1.0, 0.05
    t_i = 1.0 / (2 * H1)
    rng = np.random.default_rng(N)
    x = np.concatenate([rng.standard_normal(N ** 3), 0.3 * rng.standard_normal(N ** 3)])
    p = cfl.kg_params(N, L, m, 0.0, cfl.radiation(1.0, H1, t_i))
    w = cfl.WKB(p, dev(x), t_i)
    ow = O.WKB(N, L, m, O.Cosmo.radiation(1.0, H1, t_i), x, t_i)
    back = host(w.evaluate(t_i))
    assert np.abs(back - x).max() <= 1e-13 * np.abs(x).max()
    for t in (12.0, 55.0, 300.0):
        g = host(w.evaluate(t))
        r = ow.evaluate_at(t)
        err = np.abs(g - r).max()
        assert err <= 1e-12 * np.abs(r).max(), (t, err, np.abs(r).max())
    w.close()


def test_rejects_non_radiation(cfl):
    N = 8
    x = torch.zeros(2 * N ** 3, dtype=torch.float64, device="cuda")
    p = cfl.kg_params(N, 8.0, 1.0, 0.0, cfl.radiation(1.0, 0.5, 3.0))  # t1 != 1/(2 H1)
    with pytest.raises(cfl.CfError):
        cfl.WKB(p, x, 3.0)
