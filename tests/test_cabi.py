"""The C ABI of libcfhip.so without a GPU: it loads, exports every function that
include/cfhip.h declares, its host-side functions equal the oracle bit for bit,
and bad arguments are rejected with CF_E_INVALID before any device work."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "cfhip.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cf_[A-Za-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    import cfl_amd._lib as L
    return L


def test_library_exports_every_declared_symbol(lib):
    L = lib.load()
    names = declared()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # the Python binding declares the same set
    assert sorted(lib.EXPORTED) == names


def test_cosmology_host_functions_equal_oracle(lib, O):
    L = lib.load()
    for a1, H1, t1 in ((1.0, 0.5, 1.0), (1.0, 0.05, 10.0), (2.0, 0.3, 0.5)):
        cg = lib.Cosmology(a1, H1, t1, 1.0)
        co = O.Cosmo.radiation(a1, H1, t1)
        for t in (1.0, 3.7, 51.0, 122.5):
            assert L.cf_cosmo_a(C.byref(cg), t) == O.cosmo_a(co, t)
            assert L.cf_cosmo_H(C.byref(cg), t) == O.cosmo_H(co, t)
            assert L.cf_cosmo_eta(C.byref(cg), t) == O.cosmo_eta(co, t)
    for N, Lb, a in ((64, 64.0, 1.0), (384, 307.2, 2.5), (96, 96.0, 7.1)):
        assert L.cf_inv_ah_sqr(N, Lb, a) == O.inv_ah_sqr(N, Lb, a)


def test_normals_equal_oracle(lib, O):
    L = lib.load()
    out = np.empty(10001)
    lib.check(L.cf_normals_host(77, out.size, out.ctypes.data_as(C.c_void_p)))
    assert np.array_equal(out, O.normals(77, out.size))


@pytest.mark.parametrize("kind,params", [(0, (0.5,)), (1, ()), (2, (4.6e-4,)), (3, (1.0, 1.0, 0.0)),
                                         (3, (0.02, 1.0, -3.0)), (4, (1.0, 1.0, 0.0, 1.0))])
def test_spectrum_tables_equal_oracle(lib, O, kind, params):
    L = lib.load()
    N, Lb = 64, 51.2
    prm = np.zeros(8)
    prm[: len(params)] = params
    out = np.empty(O.nbins(N))
    lib.check(L.cf_spectrum_table_host(kind, N, Lb, prm.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p)))
    assert np.array_equal(out, O.spectrum_table(kind, N, Lb, params))


def test_sizes(lib, O):
    L = lib.load()
    for N in (4, 64, 96, 512):
        assert L.cf_spectrum_bins(N) == O.nbins(N)
    assert L.cf_rk4_work_bytes(512, 0) == 3 * 2 * 512 ** 3 * 8
    assert L.cf_rk4_work_bytes(512, 1) == 3 * 2 * 512 ** 3 * 4
    assert L.cf_rk4_work_bytes(7, 0) == 0  # odd N unsupported
    # temporally blocked steps ping-pong with one more state (FP32-comp: hi and lo)
    assert L.cf_rk4_tb_work_bytes(512, 0) == 2 * 512 ** 3 * 8
    assert L.cf_rk4_tb_work_bytes(512, 1) == 2 * 2 * 512 ** 3 * 4 + 2 * 4096  # + slack to align the second pair
    assert L.cf_rk4_tb_work_bytes(512, lib.CF_PREC_F32S) == L.cf_rk4_tb_work_bytes(512, 1)  # same storage
    # adjoint: checkpoints + one segment + step work (2) + lambda (2) + forward work
    # (3) state vectors, and 4 dL/dm partials per block (512 blocks) per step plus one
    v = 2 * 128 ** 3 * 8
    for n_steps, ckpt in ((1024, 32), (23, 5), (0, 4)):
        nck = -(-n_steps // ckpt)
        assert L.cf_adjoint_work_bytes(128, n_steps, ckpt) == (nck + ckpt + 5) * v + (4 * n_steps + 1) * 512 * 8
    assert L.cf_adjoint_work_bytes(128, 10, 0) == 0 and L.cf_adjoint_work_bytes(7, 10, 2) == 0


def test_bad_arguments_rejected_before_device_work(lib):
    L = lib.load()
    fake = C.c_void_p(0x1000)
    assert L.cf_rhs_f64(7, fake, C.c_void_p(0x2000), 1.0, 0.0, 0.0, 1.0, None) == lib.CF_E_INVALID
    assert b"unsupported" in L.cf_last_error()
    assert L.cf_rhs_f64(8, fake, fake, 1.0, 0.0, 0.0, 1.0, None) == lib.CF_E_INVALID  # aliased
    p = lib.KGParams(3, 1.0, 1.0, 0.0, lib.Cosmology(1, 0, 0, 1))
    assert L.cf_rk4_step_f64(C.byref(p), fake, 0.0, 0.1, fake, None) == lib.CF_E_INVALID
    p.N, p.L = 8, -1.0
    assert L.cf_rk4_step_f64(C.byref(p), fake, 0.0, 0.1, fake, None) == lib.CF_E_INVALID
    assert L.cf_scale_sum(15, fake, None, None, 10, None) == lib.CF_E_INVALID
    assert L.cf_spectrum_plan_create(10, None) == lib.CF_E_INVALID
    assert L.cf_spectrum_table_host(9, 8, 1.0, None, C.c_void_p(0)) == lib.CF_E_INVALID
    with pytest.raises(lib.CfError):
        lib.check(L.cf_rhs_f64(5, fake, fake, 1, 0, 0, 1, None))
    # equation variants, WKB, ensembles, cut-off, observer average
    p = lib.KGParams(16, 16.0, 1.0, 0.0, lib.Cosmology(1.0, 0.5, 1.0, 1.0))
    h = C.c_void_p()
    assert L.cf_variant_create(C.byref(p), 9, 1.0, 0, fake, C.byref(h), None) == lib.CF_E_INVALID
    assert L.cf_variant_create(C.byref(p), lib.CF_VAR_FIXED_CURVATURE, 1.0, 0, None, C.byref(h), None) == \
        lib.CF_E_INVALID  # needs Psi
    assert L.cf_variant_create(C.byref(p), lib.CF_VAR_SQRT_POTENTIAL, 0.0, 0, None, C.byref(h), None) == \
        lib.CF_E_INVALID  # f_a must be positive
    assert L.cf_variant_create(C.byref(p), lib.CF_VAR_APPROX_COMOVING_CURVATURE, 1.0, 6, fake, C.byref(h),
                               None) == lib.CF_E_INVALID  # M must divide N
    assert h.value is None
    assert L.cf_variant_rk4_run_tb(None, fake, 1.0, 0, 1, 0.1, 0, fake, None) == lib.CF_E_INVALID  # null handle
    assert L.cf_variant_rk4_step_tb(None, fake, fake, 1.0, 0.1, 0, None) == lib.CF_E_INVALID
    x0 = C.c_void_p(0x100000)
    assert L.cf_rk4_step_tb_f64(C.byref(p), x0, C.c_void_p(0x100000 + 8), 1.0, 0.1, 0, None) == \
        lib.CF_E_INVALID  # overlapping in/out
    assert L.cf_rk4_step_tb_f32c(C.byref(p), x0, C.c_void_p(0x900000), x0, C.c_void_p(0xA00000), 1.0, 0.1, 0,
                                 None) == lib.CF_E_INVALID
    assert L.cf_rk4_step_tb_f32s(C.byref(p), x0, C.c_void_p(0x900000), x0, C.c_void_p(0xA00000), 1.0, 0.1, 0,
                                 None) == lib.CF_E_INVALID
    assert b"cf_rk4_step_tb_f32s" in L.cf_last_error()
    assert L.cf_rk4_run_tb_f32s(C.byref(p), x0, None, 1.0, 0, 1, 0.1, 0, fake, None) == lib.CF_E_INVALID
    # pacing of the single-precision-stencil step: bad lattice, bad pace, null buffers
    assert L.cf_rk4_tb_set_pace(15, 0, 100) == lib.CF_E_INVALID
    assert L.cf_rk4_tb_set_pace(16, 0, -1) == lib.CF_E_INVALID
    assert L.cf_rk4_tb_get_pace(15, 0) == -1
    rows, cols = C.c_int(0), C.c_int(0)
    assert L.cf_rk4_tb_f32s_region(15, C.byref(rows), C.byref(cols)) == lib.CF_E_INVALID
    assert L.cf_rk4_tb_f32s_region(16, None, C.byref(cols)) == lib.CF_E_INVALID
    assert L.cf_rk4_tb_set_chunk(15, 4) == lib.CF_E_INVALID
    assert L.cf_rk4_tb_set_chunk(16, 17) == lib.CF_E_INVALID
    assert L.cf_rk4_tb_set_chunk(16, -1) == lib.CF_E_INVALID
    assert L.cf_rk4_tb_get_chunk(15) == -1
    pace = C.c_int(7)
    assert L.cf_rk4_tb_tune_pace(C.byref(p), None, x0, 1.0, 0.1, 0, fake, None, C.byref(pace)) == lib.CF_E_INVALID
    assert pace.value == 7
    pw = lib.KGParams(16, 16.0, 1.0, 0.0, lib.Cosmology(1.0, 0.5, 3.0, 1.0))  # t1 != 1/(2 H1)
    assert L.cf_wkb_create(C.byref(pw), fake, 3.0, C.byref(h), None) == lib.CF_E_INVALID
    assert b"radiation" in L.cf_last_error()
    assert L.cf_cutoff_fouriers(16, 32, fake, C.c_void_p(0x2000), None) == lib.CF_E_INVALID  # M > N
    assert L.cf_axis_average(16, fake, fake, None) == lib.CF_E_INVALID  # aliased
    Ns = (C.c_int64 * 2)(16, 7)
    Ls = (C.c_double * 2)(16.0, 7.0)
    assert L.cf_ensemble_create(2, Ns, Ls, 1.0, 0.0, C.byref(lib.Cosmology(1, 0, 0, 1)), 0, C.byref(h)) == \
        lib.CF_E_INVALID  # odd N
    r = C.c_int()
    assert L.cf_ensemble_set_region(None, 16, lib.CF_ENS_REGION_24X40) == lib.CF_E_INVALID  # null handle
    assert L.cf_ensemble_get_region(None, 16, C.byref(r)) == lib.CF_E_INVALID
    # slab step: plane count outside 1..N, overlapping slabs (before any device work)
    assert L.cf_rk4_step_tb_f64_slab(C.byref(p), 0, x0, C.c_void_p(0x40000000), 1.0, 0.1, 0, None) == \
        lib.CF_E_INVALID
    assert L.cf_rk4_step_tb_f64_slab(C.byref(p), 17, x0, C.c_void_p(0x40000000), 1.0, 0.1, 0, None) == \
        lib.CF_E_INVALID
    assert b"n_planes" in L.cf_last_error()
    assert L.cf_rk4_step_tb_f64_slab(C.byref(p), 4, x0, C.c_void_p(0x100000 + 8 * 2 * 12 * 256 - 8), 1.0, 0.1, 0,
                                     None) == lib.CF_E_INVALID  # the slabs overlap by one double
    assert b"overlapping" in L.cf_last_error()


@pytest.mark.parametrize("t_f", [10.0, 37.5, 160.0])
def test_wkb_tables_equal_oracle(lib, O, t_f):
    """cf_wkb_tables_host (the phase and omega tables of the WKB extrapolation, host
    code) equals the oracle's restatement bit for bit."""
    L = lib.load()
    N, Lb, m, a1, H1, t_i = 24, 19.2, 1.0, 1.0, 0.05, 10.0
    nb = 3 * (N // 2) ** 2 + 1
    g = [np.empty(nb) for _ in range(4)]
    lib.check(L.cf_wkb_tables_host(N, Lb, m, a1, H1, t_i, t_f, *[x.ctypes.data_as(C.c_void_p) for x in g]))
    for a, b in zip(g, O.wkb_tables(N, Lb, m, a1, H1, t_i, t_f)):
        assert np.array_equal(a, b)
