"""Python front end of libcfhip.so for tests, the benchmark and torch autograd.

Device buffers are torch tensors on ``cuda:0`` (torch supplies device memory and
the stream); every computation runs in the HIP kernels of libcfhip.so. Importing
this package fails if the library has not been built -- there is no fallback.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import (CF_ENS_GRID_FLAT, CF_ENS_REGION_24X40, CF_ENS_REGION_32X32, CF_PREC_F32C, CF_PREC_F32S, CF_PREC_F64, CF_SPEC_DERIV_POWER_CUTOFF, CF_SPEC_GAUSS,  # noqa: F401
                   CF_SPEC_KM3_GAUSS, CF_SPEC_POWER_CUTOFF, CF_SPEC_SCALE_INV,
                   CF_VAR_APPROX_COMOVING_CURVATURE, CF_VAR_COMOVING_CURVATURE, CF_VAR_FIXED_CURVATURE,
                   CF_VAR_SQRT_POTENTIAL, CfError, Cosmology, KGParams, check)

lib = _lib.load()
LIB_PATH = _lib.LIB_PATH


def _ptr(t):
    if t is None:
        return None
    if isinstance(t, torch.Tensor):
        assert t.is_cuda and t.is_contiguous(), "device tensors must be contiguous CUDA tensors"
        return C.c_void_p(t.data_ptr())
    return C.c_void_p(int(t))


def _stream(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def minkowski():
    """StaticEOSCosmology() (src/physics.hpp:30-31)."""
    return Cosmology(1.0, 0.0, 0.0, 1.0)


def radiation(a1, H1, t1):
    """StaticEOSCosmology(param) with p = 1 (src/physics.hpp:25-27)."""
    return Cosmology(a1, H1, t1, 1.0)


def kg_params(N, L, m, lam=0.0, cosmo=None):
    return KGParams(int(N), float(L), float(m), float(lam), cosmo if cosmo is not None else minkowski())


def cosmo_a(c, t):
    return lib.cf_cosmo_a(C.byref(c), t)


def cosmo_H(c, t):
    return lib.cf_cosmo_H(C.byref(c), t)


def cosmo_eta(c, t):
    return lib.cf_cosmo_eta(C.byref(c), t)


def inv_ah_sqr(N, L, a_t):
    return lib.cf_inv_ah_sqr(N, L, a_t)


def nbins(N):
    return int(lib.cf_spectrum_bins(N))


def sync(stream=None):
    check(lib.cf_sync(_stream(stream)))


# ---------------------------------------------------------------- odeint system / ops
def rhs_f64(N, x, dxdt, m, lam, H, inv, stream=None):
    check(lib.cf_rhs_f64(N, _ptr(x), _ptr(dxdt), m, lam, H, inv, _stream(stream)))


def scale_sum(v0, vs, alphas, stream=None):
    K = len(vs)
    arr = (C.c_void_p * K)(*[t.data_ptr() for t in vs])
    al = (C.c_double * K)(*alphas)
    check(lib.cf_scale_sum(K, _ptr(v0), arr, al, v0.numel(), _stream(stream)))


def norm_inf(v, stream=None):
    out = C.c_double()
    check(lib.cf_norm_inf(_ptr(v), v.numel(), C.byref(out), _stream(stream)))
    return out.value


def count_nonfinite(v, stream=None):
    """Number of NaN or infinite entries of a float64 device tensor (blocking)."""
    k = C.c_int64(0)
    check(lib.cf_count_nonfinite(_ptr(v), v.numel(), C.byref(k), _stream(stream)))
    return k.value


def energy_density_f64(N, x, rho, m, lam, inv, stream=None):
    check(lib.cf_energy_density_f64(N, _ptr(x), _ptr(rho), m, lam, inv, _stream(stream)))


def axis_average(N, rho, out=None, stream=None):
    if out is None:
        out = torch.empty(N * N, dtype=torch.float64, device=rho.device)
    check(lib.cf_axis_average(N, _ptr(rho), _ptr(out), _stream(stream)))
    return out


# ---------------------------------------------------------------- fused RK4
def work_bytes(N, precision=CF_PREC_F64):
    return int(lib.cf_rk4_work_bytes(N, precision))


def alloc_work(N, precision=CF_PREC_F64, device="cuda"):
    return torch.empty(work_bytes(N, precision), dtype=torch.uint8, device=device)


def rk4_step_f64(p, x, t, dt, work, stream=None):
    check(lib.cf_rk4_step_f64(C.byref(p), _ptr(x), t, dt, _ptr(work), _stream(stream)))


def rk4_run_f64(p, x, t0, k0, n_steps, dt, work, stream=None):
    check(lib.cf_rk4_run_f64(C.byref(p), _ptr(x), t0, k0, n_steps, dt, _ptr(work), _stream(stream)))


def rk4_step_f32c(p, hi, lo, t, dt, work, stream=None):
    check(lib.cf_rk4_step_f32c(C.byref(p), _ptr(hi), _ptr(lo), t, dt, _ptr(work), _stream(stream)))


def rk4_run_f32c(p, hi, lo, t0, k0, n_steps, dt, work, stream=None):
    check(lib.cf_rk4_run_f32c(C.byref(p), _ptr(hi), _ptr(lo), t0, k0, n_steps, dt, _ptr(work),
                              _stream(stream)))


def tb_work_bytes(N, precision=CF_PREC_F64):
    return int(lib.cf_rk4_tb_work_bytes(N, precision))


def alloc_tb_work(N, precision=CF_PREC_F64, device="cuda"):
    return torch.empty(tb_work_bytes(N, precision), dtype=torch.uint8, device=device)


def rk4_run_tb_f64(p, x, t0, k0, n_steps, dt, work, chunk=0, stream=None):
    """Temporally blocked RK4 (one kernel per step); same bits as rk4_run_f64."""
    check(lib.cf_rk4_run_tb_f64(C.byref(p), _ptr(x), t0, k0, n_steps, dt, chunk, _ptr(work), _stream(stream)))


def rk4_run_tb_f32c(p, hi, lo, t0, k0, n_steps, dt, work, chunk=0, stream=None):
    check(lib.cf_rk4_run_tb_f32c(C.byref(p), _ptr(hi), _ptr(lo), t0, k0, n_steps, dt, chunk, _ptr(work),
                                 _stream(stream)))


def rk4_run_tb_f32s(p, hi, lo, t0, k0, n_steps, dt, work, chunk=0, stream=None):
    """Compensated storage with the single-precision stencil (float forces, double
    increment); work from alloc_tb_work(N, CF_PREC_F32S)."""
    check(lib.cf_rk4_run_tb_f32s(C.byref(p), _ptr(hi), _ptr(lo), t0, k0, n_steps, dt, chunk, _ptr(work),
                                 _stream(stream)))


def rk4_tb_tune_pace(p, hi, lo, t, dt, work, chunk=0, stream=None):
    """Time the single-precision-stencil step of this lattice at a range of paces and
    keep the best for its later runs and steps (state unchanged); returns the pace. With
    chunk 0 the near-best chunk lengths are tried too and the fastest becomes N's
    default chunk (rk4_tb_get_chunk)."""
    out = C.c_int(0)
    check(lib.cf_rk4_tb_tune_pace(C.byref(p), _ptr(hi), _ptr(lo), t, dt, chunk, _ptr(work), _stream(stream),
                                  C.byref(out)))
    return out.value


def rk4_tb_set_pace(N, pace, chunk=0):
    check(lib.cf_rk4_tb_set_pace(N, chunk, pace))


def rk4_tb_get_pace(N, chunk=0):
    return lib.cf_rk4_tb_get_pace(N, chunk)


def rk4_tb_set_chunk(N, chunk):
    """Default chunk of the f32s step for N on this device (0: the cost model's)."""
    check(lib.cf_rk4_tb_set_chunk(N, chunk))


def rk4_tb_get_chunk(N):
    return lib.cf_rk4_tb_get_chunk(N)


def rk4_tb_f32s_region(N):
    """(rows, columns) of the regions the f32s step covers an N^3 lattice with."""
    rows, cols = C.c_int(0), C.c_int(0)
    check(lib.cf_rk4_tb_f32s_region(N, C.byref(rows), C.byref(cols)))
    return rows.value, cols.value


def rk4_step_tb_f64(p, x_in, x_out, t, dt, chunk=0, stream=None):
    """One temporally blocked step out of place: x_out = step(x_in)."""
    check(lib.cf_rk4_step_tb_f64(C.byref(p), _ptr(x_in), _ptr(x_out), t, dt, chunk, _stream(stream)))


def rk4_step_tb_f64_slab(p, n_planes, x_in, x_out, t, dt, chunk=0, stream=None):
    """One temporally blocked FP64 step of a slab of n_planes planes with four ghost
    planes on each side ([phi (n_planes+8) N^2 | pi (n_planes+8) N^2]); writes the
    slab's own planes of x_out."""
    check(lib.cf_rk4_step_tb_f64_slab(C.byref(p), n_planes, _ptr(x_in), _ptr(x_out), t, dt, chunk,
                                      _stream(stream)))


def rk4_step_tb_f32c_slab(p, n_planes, hi_in, lo_in, hi_out, lo_out, t, dt, chunk=0, stream=None):
    """The FP32-comp form of rk4_step_tb_f64_slab (hi and lo laid out like its slab)."""
    check(lib.cf_rk4_step_tb_f32c_slab(C.byref(p), n_planes, _ptr(hi_in), _ptr(lo_in), _ptr(hi_out), _ptr(lo_out),
                                       t, dt, chunk, _stream(stream)))


def rk4_step_tb_f32s_slab(p, n_planes, hi_in, lo_in, hi_out, lo_out, t, dt, chunk=0, stream=None):
    """The float-force form of rk4_step_tb_f32c_slab (paced with the pace kept for N and
    the slab's chunk)."""
    check(lib.cf_rk4_step_tb_f32s_slab(C.byref(p), n_planes, _ptr(hi_in), _ptr(lo_in), _ptr(hi_out), _ptr(lo_out),
                                       t, dt, chunk, _stream(stream)))


def rk4_step_tb_f32c(p, hi_in, lo_in, hi_out, lo_out, t, dt, chunk=0, stream=None):
    check(lib.cf_rk4_step_tb_f32c(C.byref(p), _ptr(hi_in), _ptr(lo_in), _ptr(hi_out), _ptr(lo_out), t, dt, chunk,
                                  _stream(stream)))


def rk4_step_tb_f32s(p, hi_in, lo_in, hi_out, lo_out, t, dt, chunk=0, stream=None):
    check(lib.cf_rk4_step_tb_f32s(C.byref(p), _ptr(hi_in), _ptr(lo_in), _ptr(hi_out), _ptr(lo_out), t, dt, chunk,
                                  _stream(stream)))


def split_f64(x, stream=None):
    hi = torch.empty(x.numel(), dtype=torch.float32, device=x.device)
    lo = torch.empty_like(hi)
    check(lib.cf_split_f64(_ptr(x), _ptr(hi), _ptr(lo), x.numel(), _stream(stream)))
    return hi, lo


def join_f64(hi, lo, stream=None):
    x = torch.empty(hi.numel(), dtype=torch.float64, device=hi.device)
    check(lib.cf_join_f64(_ptr(hi), _ptr(lo), _ptr(x), hi.numel(), _stream(stream)))
    return x


# ---------------------------------------------------------------- spectra
class SpectrumPlan:
    """hipFFT plans + bin table for an N^3 lattice (cufftWrapperNoBatching analogue)."""

    def __init__(self, N):
        self.N = int(N)
        h = C.c_void_p()
        check(lib.cf_spectrum_plan_create(self.N, C.byref(h)))
        self.h = h
        self.nb = nbins(self.N)

    def close(self):
        if self.h:
            lib.cf_spectrum_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _bins(self, out):
        if out is None:
            out = torch.empty(self.nb, dtype=torch.float64, device="cuda")
        return out

    def d2z(self, f, out=None, stream=None):
        N = self.N
        if out is None:
            out = torch.empty(2 * N * N * (N // 2 + 1), dtype=torch.float64, device="cuda")
        check(lib.cf_fft_d2z(self.h, _ptr(f), _ptr(out), _stream(stream)))
        return out

    def z2d(self, fk, out=None, stream=None):
        """Unnormalised C2R; fk is destroyed."""
        if out is None:
            out = torch.empty(self.N ** 3, dtype=torch.float64, device="cuda")
        check(lib.cf_fft_z2d(self.h, _ptr(fk), _ptr(out), _stream(stream)))
        return out

    def z2d_inplace(self, fk, stream=None):
        """In-place C2R: fk (2 N^2 (N/2+1) doubles) receives the unnormalised
        transform in the FFTW padded layout (rows of 2 (N/2+1) doubles)."""
        check(lib.cf_fft_z2d_inplace(self.h, _ptr(fk), _stream(stream)))
        return fk

    def power_spectrum(self, f, out=None, stream=None, consume=False):
        """Bins of |f_k|^2. consume=True hands f to the transform without the copy
        that preserves it: f's contents are undefined afterwards."""
        out = self._bins(out)
        fn = lib.cf_power_spectrum_consume if consume else lib.cf_power_spectrum
        check(fn(self.h, _ptr(f), _ptr(out), _stream(stream)))
        return out

    def mode_power_spectrum(self, state, L, m, a_t, out=None, stream=None):
        out = self._bins(out)
        check(lib.cf_mode_power_spectrum(self.h, _ptr(state), L, m, a_t, _ptr(out), _stream(stream)))
        return out

    def bin_power(self, fk, out=None, stream=None):
        out = self._bins(out)
        check(lib.cf_bin_power(self.h, _ptr(fk), _ptr(out), _stream(stream)))
        return out

    def bin_mode_power(self, phik, dphik, L, m, a_t, out=None, stream=None):
        out = self._bins(out)
        check(lib.cf_bin_mode_power(self.h, _ptr(phik), _ptr(dphik), L, m, a_t, _ptr(out),
                                    _stream(stream)))
        return out

    def radial_bins(self, sbins, out=None, stream=None):
        if out is None:
            out = torch.empty(self.N // 2, dtype=torch.float64, device="cuda")
        check(lib.cf_radial_bins(self.h, _ptr(sbins), _ptr(out), _stream(stream)))
        return out

    def grf(self, w, sqrtP, Psi=None, out=None, stream=None):
        if out is None:
            out = torch.empty(self.N ** 3, dtype=torch.float64, device="cuda")
        check(lib.cf_grf(self.h, _ptr(w), _ptr(sqrtP), _ptr(Psi), _ptr(out), _stream(stream)))
        return out

    def scaled_fourier_modes(self, f, kernel_by_s, out=None, stream=None):
        if out is None:
            out = torch.empty(self.N ** 3, dtype=torch.float64, device="cuda")
        check(lib.cf_scaled_fourier_modes(self.h, _ptr(f), _ptr(kernel_by_s), _ptr(out), _stream(stream)))
        return out

    def init_grf_state(self, seed, kind, params, L, rms, state=None, stream=None):
        if state is None:
            state = torch.empty(2 * self.N ** 3, dtype=torch.float64, device="cuda")
        prm = np.zeros(8)
        prm[: len(params)] = params
        check(lib.cf_init_grf_state(self.h, int(seed), int(kind), prm.ctypes.data_as(C.c_void_p), L, rms,
                                    _ptr(state), _stream(stream)))
        return state


def normals_host(seed, n):
    out = np.empty(n)
    check(lib.cf_normals_host(int(seed), int(n), out.ctypes.data_as(C.c_void_p)))
    return out


def spectrum_table_host(kind, N, L, params=()):
    out = np.empty(nbins(N))
    prm = np.zeros(8)
    prm[: len(params)] = params
    check(lib.cf_spectrum_table_host(int(kind), int(N), float(L), prm.ctypes.data_as(C.c_void_p),
                                     out.ctypes.data_as(C.c_void_p)))
    return out


def rescale_rms(x, target, stream=None):
    check(lib.cf_rescale_rms(_ptr(x), x.numel(), target, _stream(stream)))
    return x


# ---------------------------------------------------------------- ensembles
class Ensemble:
    """Realisations of uneven N advanced by one persistent kernel per step
    (cf_ensemble_*); runs can be captured into a hipGraph."""

    def __init__(self, Ns, Ls, m, lam, cosmo, precision=CF_PREC_F32C):
        n = len(Ns)
        Na = (C.c_int64 * n)(*[int(v) for v in Ns])
        La = (C.c_double * n)(*[float(v) for v in Ls])
        h = C.c_void_p()
        check(lib.cf_ensemble_create(n, Na, La, m, lam, C.byref(cosmo), precision, C.byref(h)))
        self.h, self.n, self.Ns, self.Ls, self.precision = h, n, list(Ns), list(Ls), precision

    def close(self):
        if self.h:
            lib.cf_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def sites(self):
        return int(lib.cf_ensemble_sites(self.h))

    def init_grf(self, seeds, kind, params, rms, stream=None):
        sa = (C.c_uint32 * self.n)(*[int(s) for s in seeds])
        prm = np.zeros(8)
        prm[: len(params)] = params
        check(lib.cf_ensemble_init_grf(self.h, sa, kind, prm.ctypes.data_as(C.c_void_p), rms, _stream(stream)))

    def state_ptrs(self, r):
        x, lo = C.c_void_p(), C.c_void_p()
        check(lib.cf_ensemble_state(self.h, r, C.byref(x), C.byref(lo)))
        return x.value, lo.value

    def state(self, r):
        """Copy of realisation r's state as a float64 torch tensor (joined for F32C)."""
        N = self.Ns[r]
        n = 2 * N ** 3
        x, lo = self.state_ptrs(r)
        out = torch.empty(n, dtype=torch.float64, device="cuda")
        if self.precision in (CF_PREC_F32C, CF_PREC_F32S):
            check(lib.cf_join_f64(C.c_void_p(x), C.c_void_p(lo), _ptr(out), n, _stream()))
        else:
            check(lib.cf_d2d(_ptr(out), C.c_void_p(x), n * 8, _stream()))
        return out

    def set_plan(self, N=0, chunk=0, grid=0):
        check(lib.cf_ensemble_set_plan(self.h, N, chunk, grid))

    def get_plan(self, N):
        c, g = C.c_int(), C.c_int()
        check(lib.cf_ensemble_get_plan(self.h, N, C.byref(c), C.byref(g)))
        return c.value, g.value

    def set_region(self, N, region):
        check(lib.cf_ensemble_set_region(self.h, N, region))

    def get_region(self, N):
        r = C.c_int()
        check(lib.cf_ensemble_get_region(self.h, N, C.byref(r)))
        return r.value

    def set_pace(self, pace):
        check(lib.cf_ensemble_set_pace(self.h, pace))

    def get_pace(self):
        return lib.cf_ensemble_get_pace(self.h)

    def tune(self, dt, reps=3):
        g = C.c_int()
        check(lib.cf_ensemble_tune(self.h, dt, reps, C.byref(g)))
        return g.value

    def run(self, t0, k0, n_steps, dt, use_graph=False, stream=None):
        check(lib.cf_ensemble_run(self.h, t0, k0, n_steps, dt, int(bool(use_graph)), _stream(stream)))

    def density_spectrum(self, r, t, out=None, stream=None):
        if out is None:
            out = torch.empty(nbins(self.Ns[r]), dtype=torch.float64, device="cuda")
        check(lib.cf_ensemble_density_spectrum(self.h, r, t, _ptr(out), _stream(stream)))
        return out


# ---------------------------------------------------------------- equation variants
class Variant:
    """A curvature / potential variant of the lattice equation (cf_variant_*):
    kind CF_VAR_FIXED_CURVATURE (field = Psi, N^3), CF_VAR_SQRT_POTENTIAL (f_a),
    CF_VAR_COMOVING_CURVATURE (field = R_fft, the half spectrum of the comoving
    curvature) or CF_VAR_APPROX_COMOVING_CURVATURE (R_fft and the coarse grid M)."""

    def __init__(self, p, kind, field=None, f_a=1.0, M=0, stream=None):
        h = C.c_void_p()
        check(lib.cf_variant_create(C.byref(p), kind, f_a, M, _ptr(field), C.byref(h), _stream(stream)))
        self.h, self.p, self.kind, self.M = h, p, kind, M
        self.N = p.N
        self._work = None
        self._tb_work = None

    def close(self):
        if self.h:
            lib.cf_variant_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def rhs(self, x, t, out=None, stream=None):
        if out is None:
            out = torch.empty_like(x)
        check(lib.cf_variant_rhs(self.h, _ptr(x), _ptr(out), t, _stream(stream)))
        return out

    def _w(self):
        if self._work is None:
            self._work = alloc_work(self.N, CF_PREC_F64)
        return self._work

    def rk4_step(self, x, t, dt, stream=None):
        check(lib.cf_variant_rk4_step(self.h, _ptr(x), t, dt, _ptr(self._w()), _stream(stream)))

    def rk4_run(self, x, t0, k0, n_steps, dt, stream=None):
        check(lib.cf_variant_rk4_run(self.h, _ptr(x), t0, k0, n_steps, dt, _ptr(self._w()), _stream(stream)))

    def rk4_run_tb(self, x, t0, k0, n_steps, dt, chunk=0, stream=None):
        """Temporally blocked steps (every variant); same bits as rk4_run."""
        if self._tb_work is None:
            self._tb_work = alloc_tb_work(self.N, CF_PREC_F64)
        check(lib.cf_variant_rk4_run_tb(self.h, _ptr(x), t0, k0, n_steps, dt, chunk, _ptr(self._tb_work),
                                        _stream(stream)))

    def rk4_step_tb(self, x_in, x_out, t, dt, chunk=0, stream=None):
        check(lib.cf_variant_rk4_step_tb(self.h, _ptr(x_in), _ptr(x_out), t, dt, chunk, _stream(stream)))

    def energy_density(self, x, t, out=None, stream=None):
        if out is None:
            out = torch.empty(self.N ** 3, dtype=torch.float64, device="cuda")
        check(lib.cf_variant_energy_density(self.h, _ptr(x), _ptr(out), t, _stream(stream)))
        return out

    def psi(self, t, stream=None):
        G = self.M if self.kind == CF_VAR_APPROX_COMOVING_CURVATURE else self.N
        ps = torch.empty(G ** 3, dtype=torch.float64, device="cuda")
        dps = torch.empty_like(ps)
        check(lib.cf_variant_psi(self.h, t, _ptr(ps), _ptr(dps), _stream(stream)))
        return ps, dps


def comoving_psi_fft(G, R_fft, factor, H, stream=None):
    ps = torch.empty_like(R_fft)
    dps = torch.empty_like(R_fft)
    check(lib.cf_comoving_psi_fft(G, _ptr(R_fft), _ptr(ps), _ptr(dps), factor, H, _stream(stream)))
    return ps, dps


def cutoff_fouriers(N, M, fk, stream=None):
    out = torch.empty(2 * M * M * (M // 2 + 1), dtype=torch.float64, device="cuda")
    check(lib.cf_cutoff_fouriers(N, M, _ptr(fk), _ptr(out), _stream(stream)))
    return out


# ---------------------------------------------------------------- WKB extrapolation
class WKB:
    """WKB extrapolation of the state at t_i to later times (cf_wkb_*)."""

    def __init__(self, p, state, t_i, stream=None):
        h = C.c_void_p()
        check(lib.cf_wkb_create(C.byref(p), _ptr(state), t_i, C.byref(h), _stream(stream)))
        self.h, self.N = h, p.N

    def close(self):
        if self.h:
            lib.cf_wkb_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def evaluate(self, t, out=None, stream=None):
        if out is None:
            out = torch.empty(2 * self.N ** 3, dtype=torch.float64, device="cuda")
        check(lib.cf_wkb_evaluate(self.h, t, _ptr(out), _stream(stream)))
        return out
