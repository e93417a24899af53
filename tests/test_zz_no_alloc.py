"""cf_rk4_run_tb_f64_noalloc: the temporally blocked FP64 run with the library's
static ping-pong buffer instead of a caller work buffer (no device allocation).

On the GPU it must equal cf_rk4_run_tb_f64 bit for bit, on consecutive calls that
share the buffer (the second with a smaller lattice, odd step counts so that the
result is copied back out of the buffer). On the CPU: sizes beyond the buffer are
rejected with CF_E_INVALID before any device work.
"""
import ctypes as C

import numpy as np
import pytest

FRW = dict(a1=1.0, H1=0.5, t1=1.0)


def _state(N, seed):
    rng = np.random.default_rng(seed)
    return np.concatenate([rng.standard_normal(N ** 3), 0.3 * rng.standard_normal(N ** 3)])


@pytest.mark.gpu
def test_noalloc_equals_work_buffer_run(cfl):
    import torch
    lib = cfl.lib
    for case, (N, k0, steps, lam) in enumerate(((128, 3, 5, 0.0), (64, 0, 3, 0.1), (128, 7, 4, 0.0))):
        print(f"case noalloc {case}: N={N} k0={k0} steps={steps} lambda={lam}", flush=True)
        p = cfl.kg_params(N, float(N), 1.0, lam, cfl.radiation(**FRW))
        x0 = torch.from_numpy(_state(N, 40 + case)).cuda()
        ref = x0.clone()
        cfl.rk4_run_tb_f64(p, ref, 1.0, k0, steps, 0.05, cfl.alloc_tb_work(N))
        got = x0.clone()
        cfl.check(lib.cf_rk4_run_tb_f64_noalloc(C.byref(p), cfl._ptr(got), 1.0, k0, steps, 0.05, 0, cfl._stream()))
        torch.cuda.synchronize()
        assert torch.equal(got, ref), f"case {case}: {int((got != ref).sum())} values differ"


def test_noalloc_rejects_large_lattice():
    import cfl_amd._lib as L
    lib = L.load()
    p = L.KGParams(L.CF_NOALLOC_MAX_N + 2, 130.0, 1.0, 0.0, L.Cosmology(1.0, 0.5, 1.0, 1.0))
    rc = lib.cf_rk4_run_tb_f64_noalloc(C.byref(p), C.c_void_p(0x1000), 1.0, 0, 1, 0.05, 0, None)
    assert rc == L.CF_E_INVALID
    msg = lib.cf_last_error().decode()
    assert "exceeds the preallocated buffer" in msg and str(L.CF_NOALLOC_MAX_N) in msg, msg
