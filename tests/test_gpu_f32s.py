"""The compensated storage with the single-precision stencil (cf_rk4_run_tb_f32s,
cf_rk4_step_tb_f32s) against the oracle's orc_rk4_step_f32s: bit-identical hi and lo
for N below the tile, partial edge tiles, explicit and default chunks, lambda on and
off, and odd step counts (which end with the ping-pong copy). Paced and unpaced
launches give the same bits, and the pace tuner leaves the state alone. On a 256-CU GPU
the launch cost model runs N=192 over 32 x 32 regions and every other size here over
24 x 40, so both instantiations of the kernel are checked."""
import numpy as np
import pytest
import torch

from test_gpu_parity import assert_bitwise, dev, host, rand_state

pytestmark = pytest.mark.gpu

FRW = dict(a1=1.0, H1=0.5, t1=1.0)


def split(x):
    hi = x.astype(np.float32)
    return hi, (x - hi.astype(np.float64)).astype(np.float32)


@pytest.mark.parametrize("N,lam,steps,chunk", [(4, 0.0, 2, 0), (6, 0.1, 1, 0), (16, 0.0, 3, 0), (40, 0.1, 3, 7),
                                               (64, 0.1, 2, 0), (96, 0.0, 1, 0), (128, 0.0, 2, 0),
                                               (192, 0.1, 1, 0), (192, 0.0, 2, 24)])
def test_f32s_bitwise_oracle(cfl, O, N, lam, steps, chunk):
    print(f"case tb f32s N={N} lam={lam} steps={steps} chunk={chunk}", flush=True)
    L = float(N)
    eq = O.Equation(N, L, 1.0, lam, O.Cosmo.radiation(**FRW))
    hi, lo = split(rand_state(N, 23, amp=0.5))
    hd, ld = dev(hi.copy()), dev(lo.copy())
    p = cfl.kg_params(N, L, 1.0, lam, cfl.radiation(**FRW))
    cfl.rk4_run_tb_f32s(p, hd, ld, 1.0, 2, steps, 0.05, cfl.alloc_tb_work(N, cfl.CF_PREC_F32S), chunk=chunk)
    for k in range(steps):
        eq.rk4_step_f32s(hi, lo, 1.0 + (2 + k) * 0.05, 0.05)
    assert_bitwise(host(hd), hi, "tb f32s hi")
    assert_bitwise(host(ld), lo, "tb f32s lo")


def test_f32s_region_choice(cfl):
    """The region shape follows the launch cost model: 32 x 32 where its tiles fill the
    CUs in fewer rounds (N=192, 384 on 256 CUs), 24 x 40 elsewhere."""
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    print(f"case f32s region choice cus={cus}", flush=True)
    for N in (4, 16, 100, 512):
        assert cfl.rk4_tb_f32s_region(N) in ((24, 40), (32, 32))
    if cus == 256:
        got = {N: cfl.rk4_tb_f32s_region(N) for N in (128, 192, 256, 384, 512, 768, 1024)}
        assert got == {128: (24, 40), 192: (32, 32), 256: (24, 40), 384: (32, 32), 512: (24, 40),
                       768: (24, 40), 1024: (24, 40)}, got


def test_f32s_steps_equal_run(cfl):
    """Three out-of-place steps equal a three-step run."""
    N = 64
    print(f"case f32s steps vs run N={N}", flush=True)
    p = cfl.kg_params(N, float(N), 1.0, 0.1, cfl.radiation(**FRW))
    hi, lo = split(rand_state(N, 24, amp=0.5))
    bufs = [(dev(hi.copy()), dev(lo.copy())), (dev(hi.copy()), dev(lo.copy()))]
    for k in range(3):
        i, o = bufs[k & 1], bufs[(k + 1) & 1]
        cfl.rk4_step_tb_f32s(p, i[0], i[1], o[0], o[1], 1.0 + k * 0.05, 0.05)
    hd, ld = dev(hi.copy()), dev(lo.copy())
    cfl.rk4_run_tb_f32s(p, hd, ld, 1.0, 0, 3, 0.05, cfl.alloc_tb_work(N, cfl.CF_PREC_F32S))
    assert torch.equal(bufs[1][0], hd) and torch.equal(bufs[1][1], ld)


@pytest.mark.parametrize("N", [96, 192])
def test_f32s_paced_same_bits(cfl, N):
    """Pacing (here far slower than the kernel's own pace) changes timing only."""
    print(f"case f32s paced N={N}", flush=True)
    p = cfl.kg_params(N, float(N), 1.0, 0.0, cfl.radiation(**FRW))
    hi, lo = split(rand_state(N, 25, amp=0.5))
    work = cfl.alloc_tb_work(N, cfl.CF_PREC_F32S)
    runs = []
    for pace in (0, 160, 400):
        cfl.rk4_tb_set_pace(N, pace)
        assert cfl.rk4_tb_get_pace(N) == pace
        hd, ld = dev(hi.copy()), dev(lo.copy())
        cfl.rk4_run_tb_f32s(p, hd, ld, 1.0, 0, 3, 0.05, work)
        runs.append((hd, ld))
    cfl.rk4_tb_set_pace(N, 0)
    for hd, ld in runs[1:]:
        assert torch.equal(hd, runs[0][0]) and torch.equal(ld, runs[0][1])


def test_f32s_tune_pace_keeps_state(cfl, O):
    N = 128
    print(f"case f32s tune pace N={N}", flush=True)
    p = cfl.kg_params(N, float(N), 1.0, 0.0, cfl.radiation(**FRW))
    hi, lo = split(rand_state(N, 26, amp=0.5))
    hd, ld = dev(hi.copy()), dev(lo.copy())
    work = cfl.alloc_tb_work(N, cfl.CF_PREC_F32S)
    model_chunk = cfl.rk4_tb_get_chunk(N)
    pace = cfl.rk4_tb_tune_pace(p, hd, ld, 1.0, 0.05, work)
    chunk = cfl.rk4_tb_get_chunk(N)
    print(f"tuned pace N={N}: {pace}, chunk {chunk} (cost model {model_chunk})", flush=True)
    assert 0 <= pace <= 332 and cfl.rk4_tb_get_pace(N) == pace and cfl.rk4_tb_get_pace(N, chunk) == pace
    assert 1 <= chunk <= N
    assert np.array_equal(host(hd), hi) and np.array_equal(host(ld), lo)
    cfl.rk4_run_tb_f32s(p, hd, ld, 1.0, 0, 2, 0.05, work)  # with the tuned chunk and pace
    cfl.rk4_tb_set_pace(N, 0)
    cfl.rk4_tb_set_chunk(N, 0)
    assert cfl.rk4_tb_get_chunk(N) == model_chunk
    eq = O.Equation(N, float(N), 1.0, 0.0, O.Cosmo.radiation(**FRW))
    for k in range(2):
        eq.rk4_step_f32s(hi, lo, 1.0 + k * 0.05, 0.05)
    assert_bitwise(host(hd), hi, "tuned f32s hi")
    assert_bitwise(host(ld), lo, "tuned f32s lo")
