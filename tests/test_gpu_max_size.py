"""The largest lattice the spectrum plans accept (N=1024: 2.1e9 state values, past
32-bit indexing), checked through properties that do not need an oracle run:

* Parseval: the power spectrum sums |f_k|^2 over every mode of the unnormalised
  transform, so sum_s P(s) = N^3 sum_x f(x)^2 (to FFT round-off, 1e-10 relative);
* one Fourier mode cos(2 pi k.x / N) puts all its power, 2 (N^3/2)^2 = N^6/2, into the
  single bin s = |k|^2 (k chosen with large, unequal components so that a wrong fold
  of the half spectrum or a wrong bin index moves it);
* one temporally blocked step equals one step of the four stage kernels bit for bit,
  FP64 and FP32-comp.
"""
import math

import pytest
import torch

pytestmark = pytest.mark.gpu

N = 1024
FRW = dict(a1=1.0, H1=0.5, t1=1.0)


@pytest.fixture(scope="module")
def state(cfl):
    print(f"case max size N={N}: GRF state", flush=True)
    plan = cfl.SpectrumPlan(N)
    x0 = plan.init_grf_state(4, cfl.CF_SPEC_GAUSS, (0.25,), float(N), 1.0)
    yield plan, x0
    del plan, x0
    torch.cuda.empty_cache()


def test_parseval(cfl, state):
    plan, x0 = state
    phi = x0[: N ** 3]
    P = plan.power_spectrum(phi)
    lhs = P.sum().item()
    rhs = float(N) ** 3 * (phi * phi).sum().item()
    assert abs(lhs - rhs) <= 1e-10 * rhs, (lhs, rhs)


def test_single_mode_bin(cfl, state):
    plan, _ = state
    k = (511, 300, 7)
    i = torch.arange(N, dtype=torch.int64, device="cuda")
    m = (k[0] * i.view(N, 1, 1) + k[1] * i.view(1, N, 1) + k[2] * i.view(1, 1, N)) % N  # exact phase index
    f = torch.cos((2 * math.pi / N) * m.to(torch.float64)).reshape(-1)
    del m
    P = plan.power_spectrum(f)
    s = k[0] ** 2 + k[1] ** 2 + k[2] ** 2
    expect = float(N) ** 6 / 2
    assert abs(P[s].item() - expect) <= 1e-9 * expect
    P[s] = 0.0
    assert P.abs().max().item() <= 1e-9 * expect  # round-off elsewhere only


def test_tb_equals_stage_kernels(cfl, state):
    _, x0 = state
    print(f"case max size N={N}: TB vs stage kernels", flush=True)
    p = cfl.kg_params(N, float(N), 1.0, 0.0, cfl.radiation(**FRW))
    a = x0.clone()
    w = cfl.alloc_work(N)
    cfl.rk4_run_f64(p, a, 1.0, 0, 1, 0.05, w)
    del w
    b = x0.clone()
    w = cfl.alloc_tb_work(N)
    cfl.rk4_run_tb_f64(p, b, 1.0, 0, 1, 0.05, w)
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    del a, b, w
    hi, lo = cfl.split_f64(x0)
    h2, l2 = hi.clone(), lo.clone()
    w = cfl.alloc_work(N, cfl.CF_PREC_F32C)
    cfl.rk4_run_f32c(p, hi, lo, 1.0, 0, 1, 0.05, w)
    del w
    w = cfl.alloc_tb_work(N, cfl.CF_PREC_F32C)
    cfl.rk4_run_tb_f32c(p, h2, l2, 1.0, 0, 1, 0.05, w)
    torch.cuda.synchronize()
    assert torch.equal(hi, h2) and torch.equal(lo, l2)
