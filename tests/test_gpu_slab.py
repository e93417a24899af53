"""The blocked FP64 step on a slab of the lattice (the share of one GPU in a slab
decomposition along a): with its four ghost planes on each side copied from the
periodic neighbours, every own site equals the whole-lattice step bit for bit, and
the ghost planes of the output are left alone."""
import numpy as np
import pytest
import torch

from test_gpu_parity import assert_bitwise, host

pytestmark = pytest.mark.gpu

FRW = dict(a1=1.0, H1=0.5, t1=1.0)


def state(cfl, N, seed):
    rng = np.random.default_rng(seed)
    x = np.concatenate([rng.standard_normal(N ** 3), 0.3 * rng.standard_normal(N ** 3)])
    return torch.from_numpy(x).cuda()


def slab_of(x, N, a0, n):
    """Planes a0 - 4 .. a0 + n + 3 (periodic) of phi, then of pi."""
    idx = torch.tensor([(a0 - 4 + i) % N for i in range(n + 8)], device=x.device)
    phi = x[: N ** 3].view(N, N * N)[idx]
    pi = x[N ** 3:].view(N, N * N)[idx]
    return torch.cat([phi.reshape(-1), pi.reshape(-1)])


def check_split(cfl, N, sizes, lam=0.0, chunk=0, t=1.3, dt=0.05):
    print(f"case slab N={N} sizes={sizes} lam={lam} chunk={chunk}", flush=True)
    assert sum(sizes) == N
    p = cfl.kg_params(N, float(N), 1.0, lam, cfl.radiation(**FRW))
    x = state(cfl, N, N + len(sizes))
    whole = torch.empty_like(x)
    cfl.rk4_step_tb_f64(p, x, whole, t, dt)
    ref = host(whole)
    a0 = 0
    for n in sizes:
        xin = slab_of(x, N, a0, n)
        xout = torch.full_like(xin, float("nan"))
        cfl.rk4_step_tb_f64_slab(p, n, xin, xout, t, dt, chunk)
        out = host(xout).reshape(2, n + 8, N * N)
        for comp in range(2):
            own = out[comp, 4:4 + n].reshape(-1)
            exp = ref[comp * N ** 3:(comp + 1) * N ** 3].reshape(N, N * N)[a0:a0 + n].reshape(-1)
            assert_bitwise(own, exp, f"slab a0={a0} n={n} comp={comp}")
            assert np.isnan(out[comp, :4]).all() and np.isnan(out[comp, 4 + n:]).all(), "ghost planes written"
        a0 += n


def test_slab_even_split(cfl):
    check_split(cfl, 64, [16, 16, 16, 16])


def test_slab_uneven_split_partial_tiles_lambda(cfl):
    # N=40: one 24-column tile and one of 16 per side; slabs of 13, 1, 26 planes
    check_split(cfl, 40, [13, 1, 26], lam=0.1)


def test_slab_explicit_chunk(cfl):
    check_split(cfl, 32, [20, 12], chunk=3)


def test_slab_one_gpu_share_n512(cfl):
    """N=512 split for 8 GPUs: the last slab's upper ghost planes wrap to planes 0..3."""
    N, n = 512, 64
    print("case slab N=512 share of 8", flush=True)
    p = cfl.kg_params(N, float(N), 1.0, 0.0, cfl.radiation(**FRW))
    x = state(cfl, N, 5)
    whole = torch.empty_like(x)
    cfl.rk4_step_tb_f64(p, x, whole, 1.0, 0.05)
    for a0 in (0, N - n):
        xin = slab_of(x, N, a0, n)
        xout = torch.full_like(xin, float("nan"))
        cfl.rk4_step_tb_f64_slab(p, n, xin, xout, 1.0, 0.05)
        for comp in range(2):
            own = xout.view(2, n + 8, N * N)[comp, 4:4 + n]
            exp = whole[comp * N ** 3:(comp + 1) * N ** 3].view(N, N * N)[a0:a0 + n]
            assert torch.equal(own, exp), f"a0={a0} comp={comp}"
        del xin, xout
    torch.cuda.synchronize()


@pytest.mark.parametrize("mode", ["f32c", "f32s"])
@pytest.mark.parametrize("N,sizes,lam", [(64, [16, 16, 16, 16], 0.0), (40, [13, 1, 26], 0.1)])
def test_slab_f32c(cfl, N, sizes, lam, mode):
    """Compensated storage (FP64 arithmetic, or float forces paced): hi and lo slabs, own
    sites equal the whole-lattice step of the same form."""
    print(f"case slab {mode} N={N} sizes={sizes} lam={lam}", flush=True)
    p = cfl.kg_params(N, float(N), 1.0, lam, cfl.radiation(**FRW))
    hi, lo = cfl.split_f64(state(cfl, N, 7))
    whi, wlo = torch.empty_like(hi), torch.empty_like(lo)
    whole = cfl.rk4_step_tb_f32c if mode == "f32c" else cfl.rk4_step_tb_f32s
    slab_step = cfl.rk4_step_tb_f32c_slab if mode == "f32c" else cfl.rk4_step_tb_f32s_slab
    whole(p, hi, lo, whi, wlo, 1.3, 0.05)
    chunk = 5
    if mode == "f32s":
        cfl.rk4_tb_set_pace(N, 170, chunk)  # paced slabs (timing only)
    a0 = 0
    for n in sizes:
        shi, slo = slab_of(hi, N, a0, n), slab_of(lo, N, a0, n)
        ohi = torch.full_like(shi, float("nan"))
        olo = torch.full_like(slo, float("nan"))
        slab_step(p, n, shi, slo, ohi, olo, 1.3, 0.05, chunk)
        for got, whole, name in ((ohi, whi, "hi"), (olo, wlo, "lo")):
            g = got.view(2, n + 8, N * N)
            w = whole.view(2, N, N * N)
            assert torch.equal(g[:, 4:4 + n], w[:, a0:a0 + n]), f"{name} a0={a0} n={n}"
            assert torch.isnan(g[:, :4]).all() and torch.isnan(g[:, 4 + n:]).all(), "ghost planes written"
        a0 += n
    if mode == "f32s":
        cfl.rk4_tb_set_pace(N, 0, chunk)


def test_slab_rejects_bad_sizes(cfl):
    p = cfl.kg_params(16, 16.0, 1.0, 0.0, cfl.radiation(**FRW))
    buf = torch.zeros(2 * 9 * 256, dtype=torch.float64, device="cuda")
    out = torch.zeros_like(buf)
    with pytest.raises(cfl.CfError):
        cfl.rk4_step_tb_f64_slab(p, 0, buf, out, 1.0, 0.05)
    with pytest.raises(cfl.CfError):
        cfl.rk4_step_tb_f64_slab(p, 17, buf, out, 1.0, 0.05)
    with pytest.raises(cfl.CfError):
        cfl.rk4_step_tb_f64_slab(p, 1, buf, buf, 1.0, 0.05)
