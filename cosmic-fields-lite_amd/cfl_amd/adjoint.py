"""Gradients of the spectral loss L = sum_s w_s P_rho(s, t_end) through the RK4
lattice evolution, on the GPU (cf_adjoint_* in libcfhip.so), and a
torch.autograd.Function around them.

The forward model is the library's FP64 evolution (KleinGordonEquationInFRW /
+ lambda phi^4, classic RK4, t_k = t0 + k dt) from phi0 with dphi/dt(t0) = 0;
P_rho is compute_power_spectrum (fdm3d.cpp:7-34) of the energy density
(equations.cpp:106-138) at t_end. The gradients are those of the discrete,
computed quantities (discrete adjoint), with checkpoints every `ckpt` steps.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from . import SpectrumPlan, _ptr, _stream, check, kg_params, lib, nbins


def spectral_weights(N, L):
    """w_s = 1 / (1 + k_s^2), k_s = 2 pi sqrt(s) / L."""
    s = np.arange(nbins(N), dtype=np.float64)
    k = np.sqrt(s) * 2 * math.pi / L
    return 1.0 / (1.0 + k * k)


class AdjointProblem:
    """Holds the plan, weights and work buffers of one (N, n_steps, ckpt) problem."""

    def __init__(self, N, L, lam, cosmo, t0, dt, n_steps, ckpt, weights=None):
        self.N, self.L, self.lam, self.cosmo = int(N), float(L), float(lam), cosmo
        self.t0, self.dt, self.n_steps, self.ckpt = float(t0), float(dt), int(n_steps), int(ckpt)
        self.plan = SpectrumPlan(self.N)
        w = spectral_weights(self.N, self.L) if weights is None else np.asarray(weights, np.float64)
        self.weights = torch.from_numpy(w).cuda()
        nbytes = lib.cf_adjoint_work_bytes(self.N, self.n_steps, self.ckpt)
        self.work = torch.empty(int(nbytes), dtype=torch.uint8, device="cuda")
        self.rk4_work = torch.empty(int(lib.cf_rk4_tb_work_bytes(self.N, 0)), dtype=torch.uint8, device="cuda")

    def params(self, m):
        return kg_params(self.N, self.L, float(m), self.lam, self.cosmo)

    def state0(self, phi0):
        return torch.cat([phi0.reshape(-1).to(torch.float64), torch.zeros_like(phi0.reshape(-1), dtype=torch.float64)])

    def loss(self, phi0, m):
        """Forward only: L(phi0, m) (the temporally blocked step: the adjoint's forward bits)."""
        x = self.state0(phi0).contiguous()
        p = self.params(m)
        check(lib.cf_rk4_run_tb_f64(C.byref(p), _ptr(x), self.t0, 0, self.n_steps, self.dt, 0, _ptr(self.rk4_work),
                                    _stream()))
        return self.loss_at(x, m)

    def loss_at(self, x, m):
        """L at a state x at t_end."""
        p = self.params(m)
        t_end = self.t0 + self.n_steps * self.dt
        g = torch.empty_like(x)
        scratch = torch.empty(self.N ** 3, dtype=torch.float64, device="cuda")
        Lo, dm = C.c_double(), C.c_double()
        check(lib.cf_loss_spectrum_adjoint(self.plan.h, C.byref(p), _ptr(x), t_end, _ptr(self.weights), _ptr(g),
                                           _ptr(scratch), C.byref(Lo), C.byref(dm), _stream()))
        return Lo.value, g, dm.value

    def loss_and_grads(self, phi0, m):
        """(L, dL/dphi0 (N^3), dL/dm)."""
        x0 = self.state0(phi0).contiguous()
        p = self.params(m)
        g = torch.empty_like(x0)
        Lo, dm = C.c_double(), C.c_double()
        check(lib.cf_adjoint_spectrum_loss(self.plan.h, C.byref(p), _ptr(x0), self.t0, self.dt, self.n_steps,
                                           self.ckpt, _ptr(self.weights), _ptr(g), C.byref(Lo), C.byref(dm),
                                           _ptr(self.work), _stream()))
        return Lo.value, g[: self.N ** 3].clone(), dm.value


class SpectralLoss(torch.autograd.Function):
    """loss = SpectralLoss.apply(phi0, m, problem): phi0 (N^3 float64 CUDA tensor),
    m (0-d float64 tensor). Backward returns grad_out * dL/dphi0 and grad_out * dL/dm
    from the hand-written discrete adjoint."""

    @staticmethod
    def forward(ctx, phi0, m, problem):
        L, g_phi, g_m = problem.loss_and_grads(phi0.detach(), float(m))
        ctx.save_for_backward(g_phi.reshape(phi0.shape), torch.tensor(g_m, dtype=torch.float64, device=phi0.device))
        return torch.tensor(L, dtype=torch.float64, device=phi0.device)

    @staticmethod
    def backward(ctx, grad_out):
        g_phi, g_m = ctx.saved_tensors
        return grad_out * g_phi, grad_out * g_m, None
