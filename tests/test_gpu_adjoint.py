"""Hand-written discrete adjoint (cf_adjoint_*) against torch float64 autograd on
the CPU (the same computation differentiated automatically) and against
central finite differences of the GPU loss (the config: N=128, 1024 steps,
checkpoints every 32, epsilon = 1e-6, 16 random unit directions and m)."""
import math

import numpy as np
import pytest
import torch

from test_gpu_parity import host

pytestmark = pytest.mark.gpu

FRW = dict(a1=1.0, H1=0.5, t1=1.0)


def shift_sq(N):
    sh = np.minimum(np.arange(N), N - np.arange(N))
    return sh[:, None, None] ** 2 + sh[None, :, None] ** 2 + sh[None, None, :] ** 2


def t_rhs(x, N, m, lam, H, inv):
    phi = x[: N ** 3].reshape(N, N, N)
    pi = x[N ** 3:].reshape(N, N, N)
    lap = sum(torch.roll(phi, 1, d) + torch.roll(phi, -1, d) for d in range(3)) - 6 * phi
    dpi = -3 * H * pi - m * m * phi - lam * phi ** 3 + inv * lap
    return torch.cat([pi.reshape(-1), dpi.reshape(-1)])


def t_step(x, N, L, m, lam, cosmo, t, dt, O):
    def f(y, tt):
        a = O.cosmo_a(cosmo, tt)
        return t_rhs(y, N, m, lam, O.cosmo_H(cosmo, tt), O.inv_ah_sqr(N, L, a))
    k1 = f(x, t)
    k2 = f(x + dt / 2 * k1, t + dt / 2)
    k3 = f(x + dt / 2 * k2, t + dt / 2)
    k4 = f(x + dt * k3, t + dt)
    return x + dt / 6 * k1 + dt / 3 * k2 + dt / 3 * k3 + dt / 6 * k4


def t_loss(x, N, m, lam, inv, W):
    phi = x[: N ** 3].reshape(N, N, N)
    pi = x[N ** 3:].reshape(N, N, N)
    cd = sum((torch.roll(phi, -1, d) - torch.roll(phi, 1, d)) ** 2 for d in range(3))
    rho = 0.5 * pi ** 2 + 0.125 * inv * cd + 0.5 * m * m * phi ** 2 + 0.25 * lam * phi ** 4
    F = torch.fft.fftn(rho)
    return (W * (F.real ** 2 + F.imag ** 2)).sum()


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


@pytest.mark.parametrize("lam", [0.0, 0.2])
def test_loss_adjoint_vs_torch(cfl, O, lam):
    from cfl_amd.adjoint import AdjointProblem, spectral_weights
    N, L, m, t = 12, 15.0, 1.1, 2.0
    cosmo_o = O.Cosmo.radiation(**FRW)
    rng = np.random.default_rng(1)
    x = np.concatenate([0.5 * rng.standard_normal(N ** 3), 0.3 * rng.standard_normal(N ** 3)])
    prob = AdjointProblem(N, L, lam, cfl.radiation(**FRW), 1.0, 0.05, 20, 4)
    Lg, gx, dm = prob.loss_at(torch.from_numpy(x).cuda(), m)
    w = spectral_weights(N, L)
    W = torch.from_numpy(w[shift_sq(N)])
    xt = torch.tensor(x, requires_grad=True)
    mt = torch.tensor(m, dtype=torch.float64, requires_grad=True)
    inv = O.inv_ah_sqr(N, L, O.cosmo_a(cosmo_o, t))
    Lt = t_loss(xt, N, mt, lam, inv, W)
    gxt, gmt = torch.autograd.grad(Lt, (xt, mt))
    assert abs(Lg - Lt.item()) <= 1e-12 * abs(Lt.item())
    assert rel(host(gx), gxt.numpy()) < 1e-12
    assert abs(dm - gmt.item()) <= 1e-11 * abs(gmt.item())


@pytest.mark.parametrize("lam", [0.0, 0.3])
def test_step_adjoint_vs_torch(cfl, O, lam):
    import ctypes as C
    N, L, m, t, dt = 10, 12.0, 0.9, 1.5, 0.1
    cosmo_o = O.Cosmo.radiation(**FRW)
    rng = np.random.default_rng(2)
    x = np.concatenate([0.6 * rng.standard_normal(N ** 3), 0.4 * rng.standard_normal(N ** 3)])
    lamv = rng.standard_normal(2 * N ** 3)
    p = cfl.kg_params(N, L, m, lam, cfl.radiation(**FRW))
    ld = torch.from_numpy(lamv.copy()).cuda()
    part = torch.zeros(4 * 512, dtype=torch.float64, device="cuda")
    work = torch.empty(7 * 2 * N ** 3, dtype=torch.float64, device="cuda")
    cfl.check(cfl.lib.cf_rk4_step_adjoint_f64(C.byref(p), cfl._ptr(torch.from_numpy(x).cuda()), t, dt, cfl._ptr(ld),
                                              cfl._ptr(part), cfl._ptr(work), cfl._stream()))
    xt = torch.tensor(x, requires_grad=True)
    mt = torch.tensor(m, dtype=torch.float64, requires_grad=True)
    y = t_step(xt, N, L, mt, lam, cosmo_o, t, dt, O)
    gx, gm = torch.autograd.grad(y, (xt, mt), grad_outputs=torch.from_numpy(lamv))
    assert rel(host(ld), gx.numpy()) < 1e-13
    assert abs(host(part).sum() - gm.item()) <= 1e-12 * abs(gm.item())


@pytest.mark.parametrize("N,lam", [(70, 0.25), (34, 0.0), (4, 0.1), (128, 0.0)])
def test_step_adjoint_partial_tiles_vs_torch(cfl, O, N, lam):
    """One step back where the 28 x 28 output tiles of the backward kernel do not
    divide the lattice: N = 70 (tiles of 28, 28, 14 per side), N = 34 (28, 6), N = 4
    (one tile; the 32 x 32 region wraps the lattice eight times per side) and N = 128
    (28 x 4 + 16, ten chunks along a). The kernel skips the columns outside the
    lattice and wraps the neighbours; the work buffer is exactly the documented 2N^3
    doubles, and every one of the 2048 dL/dm partial slots is written (they start as
    NaN)."""
    import ctypes as C
    print(f"case step adjoint N={N} lam={lam}", flush=True)
    L, m, t, dt = float(N), 1.2, 1.4, 0.07
    cosmo_o = O.Cosmo.radiation(**FRW)
    rng = np.random.default_rng(N)
    x = np.concatenate([0.5 * rng.standard_normal(N ** 3), 0.3 * rng.standard_normal(N ** 3)])
    lamv = rng.standard_normal(2 * N ** 3)
    p = cfl.kg_params(N, L, m, lam, cfl.radiation(**FRW))
    ld = torch.from_numpy(lamv.copy()).cuda()
    part = torch.full((4 * 512,), float("nan"), dtype=torch.float64, device="cuda")
    work = torch.empty(2 * N ** 3, dtype=torch.float64, device="cuda")
    cfl.check(cfl.lib.cf_rk4_step_adjoint_f64(C.byref(p), cfl._ptr(torch.from_numpy(x).cuda()), t, dt, cfl._ptr(ld),
                                              cfl._ptr(part), cfl._ptr(work), cfl._stream()))
    xt = torch.tensor(x, requires_grad=True)
    mt = torch.tensor(m, dtype=torch.float64, requires_grad=True)
    y = t_step(xt, N, L, mt, lam, cosmo_o, t, dt, O)
    gx, gm = torch.autograd.grad(y, (xt, mt), grad_outputs=torch.from_numpy(lamv))
    assert rel(host(ld), gx.numpy()) < 1e-13
    assert abs(host(part).sum() - gm.item()) <= 1e-12 * abs(gm.item())


@pytest.mark.parametrize("n_steps,ckpt,lam", [(24, 8, 0.0), (23, 5, 0.2), (7, 16, 0.0), (1, 1, 0.2)])
def test_full_chain_vs_torch(cfl, O, n_steps, ckpt, lam):
    """Checkpointed forward + backward through n_steps equals torch autograd of the
    same RK4 chain + loss (float64, CPU)."""
    from cfl_amd.adjoint import AdjointProblem, spectral_weights
    print(f"case chain n={n_steps} ckpt={ckpt} lam={lam}", flush=True)
    N, L, m, t0, dt = 8, 10.0, 1.0, 1.0, 0.05
    cosmo_o = O.Cosmo.radiation(**FRW)
    phi0 = O.initial_state(N, L, 3, O.SPEC_GAUSS, (0.5,), 0.3)[: N ** 3]
    prob = AdjointProblem(N, L, lam, cfl.radiation(**FRW), t0, dt, n_steps, ckpt)
    Lg, gphi, gm = prob.loss_and_grads(torch.from_numpy(phi0).cuda(), m)
    pt = torch.tensor(phi0, requires_grad=True)
    mt = torch.tensor(m, dtype=torch.float64, requires_grad=True)
    x = torch.cat([pt, torch.zeros(N ** 3, dtype=torch.float64)])
    for k in range(n_steps):
        x = t_step(x, N, L, mt, lam, cosmo_o, t0 + k * dt, dt, O)
    t_end = t0 + n_steps * dt
    W = torch.from_numpy(spectral_weights(N, L)[shift_sq(N)])
    Lt = t_loss(x, N, mt, lam, O.inv_ah_sqr(N, L, O.cosmo_a(cosmo_o, t_end)), W)
    gpt, gmt = torch.autograd.grad(Lt, (pt, mt))
    assert abs(Lg - Lt.item()) <= 1e-11 * abs(Lt.item())
    assert rel(host(gphi), gpt.numpy()) < 1e-11
    assert abs(gm - gmt.item()) <= 1e-10 * abs(gmt.item())


def test_config5_finite_differences(cfl, O):
    """N=128, L=128, m=1, FRW, seed 3, rms 0.3 (P(k) ~ exp(-k^2/(2*0.5^2))), dt 0.05,
    1024 steps, checkpoints every 32: adjoint vs central differences (eps 1e-6) on 16
    random directions in phi0 and on m.

    The directions have unit RMS per site (|d|_2 = sqrt(N^3)). With unit-L2
    directions a step of 1e-6 moves each site by ~1e-9, and the change of the loss
    (dominated by its k=0 bin, the squared total energy) is then close to the
    round-off of a 1024-step evaluation: those differences carried 7e-5 .. 5e-3
    relative noise. Unit-RMS directions lift the signal 1448-fold; what remains is
    an absolute floor of ~1e-3 (seen on a direction whose derivative is 27 against
    a typical 2600), so each difference is measured against the larger of
    |adjoint| and |dL/dphi0|_2."""
    from cfl_amd.adjoint import AdjointProblem, SpectralLoss
    N, L, m, t0, dt, K = 128, 128.0, 1.0, 1.0, 0.05, 1024
    plan = cfl.SpectrumPlan(N)
    x0 = plan.init_grf_state(3, cfl.CF_SPEC_GAUSS, (0.5,), L, 0.3)
    del plan
    phi0 = x0[: N ** 3].clone()
    prob = AdjointProblem(N, L, 0.0, cfl.radiation(**FRW), t0, dt, K, 32)
    Lg, gphi, gm = prob.loss_and_grads(phi0, m)
    print(f"loss {Lg:.12e} dL/dm {gm:.12e} |dL/dphi0| {float(torch.linalg.norm(gphi)):.6e}", flush=True)
    # the autograd wrapper returns the same gradients
    pt = phi0.clone().requires_grad_(True)
    mt = torch.tensor(m, dtype=torch.float64, device="cuda", requires_grad=True)
    loss = SpectralLoss.apply(pt, mt, prob)
    loss.backward()
    assert loss.item() == Lg and torch.equal(pt.grad, gphi) and mt.grad.item() == gm
    eps = 1e-6
    gnorm = float(torch.linalg.norm(gphi))
    g = torch.Generator(device="cpu").manual_seed(0)
    errs = []
    for i in range(16):
        d = torch.randn(N ** 3, generator=g, dtype=torch.float64).cuda()
        d *= math.sqrt(N ** 3) / torch.linalg.norm(d)
        lp = prob.loss(phi0 + eps * d, m)[0]
        lm = prob.loss(phi0 - eps * d, m)[0]
        fd = (lp - lm) / (2 * eps)
        ad = float(torch.dot(gphi, d))
        # relative to the larger of |adjoint| and the typical size |g|_2 of a
        # directional derivative along a unit-RMS direction: the difference's
        # round-off floor is absolute (~eps * loss / epsilon)
        errs.append(abs(fd - ad) / max(abs(ad), gnorm))
        print(f"dir {i}: adjoint {ad:.10e} fd {fd:.10e} rel {abs(fd - ad) / abs(ad):.2e} "
              f"scaled {errs[-1]:.2e}", flush=True)
    fdm = (prob.loss(phi0, m + eps)[0] - prob.loss(phi0, m - eps)[0]) / (2 * eps)
    print(f"m: adjoint {gm:.12e} fd {fdm:.12e} rel {abs(fdm - gm) / abs(gm):.2e}", flush=True)
    assert max(errs) < 1e-5
    assert abs(fdm - gm) / abs(gm) < 1e-6
