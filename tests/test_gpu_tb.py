"""Temporally blocked RK4 (one kernel per step) against the oracle and the
four-kernel path: bit-identical states for every lattice size class (N smaller
than the tile, partial edge tiles, several a-chunks, odd step counts that end
with the ping-pong copy, the default chunk of the launch cost model -- N=128: 19
planes, a partial last chunk) and both precisions."""
import numpy as np
import pytest
import torch

from test_gpu_parity import assert_bitwise, dev, host, rand_state

pytestmark = pytest.mark.gpu

FRW = dict(a1=1.0, H1=0.5, t1=1.0)


@pytest.mark.parametrize("N,lam,steps,chunk", [(4, 0.1, 2, 0), (6, 0.0, 3, 0), (8, 0.0, 3, 0), (16, 0.1, 2, 0),
                                               (24, 0.0, 1, 0), (128, 0.1, 2, 0),
                                               (40, 0.1, 3, 7), (64, 0.0, 2, 16), (96, 0.1, 3, 0)])
def test_tb_f64_bitwise_oracle(cfl, O, N, lam, steps, chunk):
    print(f"case tb f64 N={N} lam={lam} steps={steps} chunk={chunk}", flush=True)
    L = 1.25 * N
    eq = O.Equation(N, L, 1.0, lam, O.Cosmo.radiation(**FRW))
    x = rand_state(N, 21, amp=0.5)
    xd = dev(x)
    p = cfl.kg_params(N, L, 1.0, lam, cfl.radiation(**FRW))
    cfl.rk4_run_tb_f64(p, xd, 1.0, 3, steps, 0.05, cfl.alloc_tb_work(N), chunk=chunk)
    for k in range(steps):
        eq.rk4_step(x, 1.0 + (3 + k) * 0.05, 0.05)
    assert_bitwise(host(xd), x, "tb state")


@pytest.mark.parametrize("N,lam,steps", [(4, 0.0, 2), (6, 0.1, 1), (16, 0.0, 3), (64, 0.1, 2), (96, 0.0, 1),
                                         (128, 0.0, 2)])
def test_tb_f32c_bitwise_oracle(cfl, O, N, lam, steps):
    print(f"case tb f32c N={N} lam={lam} steps={steps}", flush=True)
    L = float(N)
    eq = O.Equation(N, L, 1.0, lam, O.Cosmo.radiation(**FRW))
    x = rand_state(N, 22, amp=0.5)
    hi = x.astype(np.float32)
    lo = (x - hi.astype(np.float64)).astype(np.float32)
    hd, ld = dev(hi.copy()), dev(lo.copy())
    p = cfl.kg_params(N, L, 1.0, lam, cfl.radiation(**FRW))
    cfl.rk4_run_tb_f32c(p, hd, ld, 1.0, 0, steps, 0.05, cfl.alloc_tb_work(N, cfl.CF_PREC_F32C))
    for k in range(steps):
        eq.rk4_step_f32c(hi, lo, 1.0 + k * 0.05, 0.05)
    assert_bitwise(host(hd), hi, "tb f32c hi")
    assert_bitwise(host(ld), lo, "tb f32c lo")


@pytest.mark.parametrize("N", [256, 512])
def test_tb_equals_stage_kernels_large(cfl, N):
    """Size-independent property at full size: both schedules give the same bits."""
    print(f"case tb vs stages N={N}", flush=True)
    plan = cfl.SpectrumPlan(N)
    x0 = plan.init_grf_state(2, cfl.CF_SPEC_GAUSS, (0.25,), float(N), 1.0)
    del plan
    p = cfl.kg_params(N, float(N), 1.0, 0.0, cfl.radiation(**FRW))
    a = x0.clone()
    cfl.rk4_run_f64(p, a, 1.0, 0, 2, 0.05, cfl.alloc_work(N))
    b = x0.clone()
    cfl.rk4_run_tb_f64(p, b, 1.0, 0, 2, 0.05, cfl.alloc_tb_work(N))
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    del a, b
    hi, lo = cfl.split_f64(x0)
    h2, l2 = hi.clone(), lo.clone()
    cfl.rk4_run_f32c(p, hi, lo, 1.0, 0, 2, 0.05, cfl.alloc_work(N, cfl.CF_PREC_F32C))
    cfl.rk4_run_tb_f32c(p, h2, l2, 1.0, 0, 2, 0.05, cfl.alloc_tb_work(N, cfl.CF_PREC_F32C))
    torch.cuda.synchronize()
    assert torch.equal(hi, h2) and torch.equal(lo, l2)


@pytest.mark.parametrize("N,lam", [(16, 0.1), (40, 0.0)])
def test_tb_single_step_out_of_place(cfl, O, N, lam):
    """cf_rk4_step_tb_f64 / _f32c (input left untouched, result in the second
    buffer) give the in-place step's bits; chained by swapping buffers they give
    the oracle's three steps."""
    print(f"case tb single step N={N} lam={lam}", flush=True)
    L = float(N)
    eq = O.Equation(N, L, 1.0, lam, O.Cosmo.radiation(**FRW))
    p = cfl.kg_params(N, L, 1.0, lam, cfl.radiation(**FRW))
    x = rand_state(N, 23, amp=0.5)
    a, b = dev(x), torch.empty(x.size, dtype=torch.float64, device="cuda")
    for k in range(3):
        cfl.rk4_step_tb_f64(p, a, b, 1.0 + k * 0.05, 0.05)
        if k == 0:
            assert_bitwise(host(a), x, "input untouched")
        a, b = b, a
        eq.rk4_step(x, 1.0 + k * 0.05, 0.05)
    assert_bitwise(host(a), x, "three swapped steps")
    y = rand_state(N, 24, amp=0.5)
    hi = y.astype(np.float32)
    lo = (y - hi.astype(np.float64)).astype(np.float32)
    hd, ld = dev(hi.copy()), dev(lo.copy())
    ho, lo_o = torch.empty_like(hd), torch.empty_like(ld)
    cfl.rk4_step_tb_f32c(p, hd, ld, ho, lo_o, 1.0, 0.05)
    eq.rk4_step_f32c(hi, lo, 1.0, 0.05)
    assert_bitwise(host(ho), hi, "f32c hi")
    assert_bitwise(host(lo_o), lo, "f32c lo")


def test_tb_overlap_checks(cfl):
    """Overlapping buffers are refused; adjacent ones of unequal length (hi, then
    the work buffer of cf_rk4_tb_work_bytes, then lo, in one allocation) are accepted."""
    N = 16
    p = cfl.kg_params(N, 16.0, 1.0, 0.0, cfl.radiation(**FRW))
    x = torch.zeros(4 * N ** 3, dtype=torch.float64, device="cuda")
    with pytest.raises(cfl.CfError):
        cfl.rk4_step_tb_f64(p, x, x[N ** 3:], 1.0, 0.05)  # half overlap
    with pytest.raises(cfl.CfError):
        cfl.rk4_run_tb_f64(p, x, 1.0, 0, 1, 0.05, x[N ** 3:])
    n = 2 * N ** 3
    y = rand_state(N, 25, amp=0.5)
    hi = y.astype(np.float32)
    lo = (y - hi.astype(np.float64)).astype(np.float32)
    w = cfl.tb_work_bytes(N, cfl.CF_PREC_F32C) // 4  # work buffer, in floats
    big = torch.empty(2 * n + w, dtype=torch.float32, device="cuda")
    big[:n] = dev(hi)
    big[n + w:] = dev(lo)
    with pytest.raises(cfl.CfError):  # a work buffer one float short of the documented size
        cfl.rk4_run_tb_f32c(p, big[:n], big[n + w - 1:2 * n + w - 1], 1.0, 0, 1, 0.05, big[n:n + w])
    cfl.rk4_run_tb_f32c(p, big[:n], big[n + w:], 1.0, 0, 1, 0.05, big[n:n + w])
    h2, l2 = dev(hi), dev(lo)
    cfl.rk4_run_f32c(p, h2, l2, 1.0, 0, 1, 0.05, cfl.alloc_work(N, cfl.CF_PREC_F32C))
    torch.cuda.synchronize()
    assert torch.equal(big[:n], h2) and torch.equal(big[n + w:], l2)


@pytest.mark.parametrize("steps", [3, 4])
def test_tb_f32c_work_placement_keeps_bits(cfl, steps):
    """hi, lo and the work buffer carved from one allocation at offsets that are not
    multiples of 4 KB (the run places its second hi and lo inside work at the caller's
    residues modulo 4 KB) give the bits of separately allocated buffers; an odd step
    count ends with the copy back from the work buffer."""
    N = 64
    n = 2 * N ** 3
    print(f"case f32c work placement steps={steps}", flush=True)
    p = cfl.kg_params(N, float(N), 1.0, 0.1, cfl.radiation(**FRW))
    y = rand_state(N, 26, amp=0.5)
    hi = y.astype(np.float32)
    lo = (y - hi.astype(np.float64)).astype(np.float32)
    wb = cfl.tb_work_bytes(N, cfl.CF_PREC_F32C)
    o_lo, o_w = 4 * n + 256, 8 * n + 256 + 1280  # byte offsets of lo and work
    raw = torch.empty(o_w + wb, dtype=torch.uint8, device="cuda")
    hv = raw[:4 * n].view(torch.float32)
    lv = raw[o_lo:o_lo + 4 * n].view(torch.float32)
    wv = raw[o_w:o_w + wb]
    hv.copy_(dev(hi))
    lv.copy_(dev(lo))
    cfl.rk4_run_tb_f32c(p, hv, lv, 1.0, 0, steps, 0.05, wv)
    h2, l2 = dev(hi), dev(lo)
    cfl.rk4_run_tb_f32c(p, h2, l2, 1.0, 0, steps, 0.05, cfl.alloc_tb_work(N, cfl.CF_PREC_F32C))
    torch.cuda.synchronize()
    assert torch.equal(hv, h2) and torch.equal(lv, l2)
