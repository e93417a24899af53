"""ctypes declarations of libcfhip.so (include/cfhip.h).

The library is loaded from ``cosmic-fields-lite_amd/lib/libcfhip.so``; if it is
missing this module raises at import -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

_PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG_ROOT, "lib", "libcfhip.so")

CF_OK = 0
CF_E_INVALID = 1
CF_E_HIP = 2
CF_E_FFT = 3
CF_E_NOMEM = 4
CF_E_UNSUPPORTED = 5

CF_PREC_F64 = 0
CF_PREC_F32C = 1
CF_PREC_F32S = 2

CF_NOALLOC_MAX_N = 128  # largest N of cf_rk4_run_tb_f64_noalloc

CF_ENS_GRID_FLAT = -1  # ensemble plan: one block per work unit
CF_ENS_REGION_32X32 = 0  # ensemble block region (b rows x c columns): 24 x 24 tiles
CF_ENS_REGION_24X40 = 1  # 16 x 32 tiles

CF_VAR_FIXED_CURVATURE = 1
CF_VAR_SQRT_POTENTIAL = 2
CF_VAR_COMOVING_CURVATURE = 3
CF_VAR_APPROX_COMOVING_CURVATURE = 4

CF_SPEC_GAUSS = 0
CF_SPEC_KM3_GAUSS = 1
CF_SPEC_SCALE_INV = 2
CF_SPEC_POWER_CUTOFF = 3
CF_SPEC_DERIV_POWER_CUTOFF = 4


class CfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"cfhip error {code}: {msg}")
        self.code = code


class Cosmology(C.Structure):
    _fields_ = [("a1", C.c_double), ("H1", C.c_double), ("t1", C.c_double), ("p", C.c_double)]


class KGParams(C.Structure):
    _fields_ = [("N", C.c_int64), ("L", C.c_double), ("m", C.c_double), ("lam", C.c_double),
                ("cosmo", Cosmology)]


vp = C.c_void_p
i64 = C.c_int64
dbl = C.c_double
st = C.c_int32

# (name, restype, argtypes)
_SIGS = [
    ("cf_last_error", C.c_char_p, []),
    ("cf_version", C.c_char_p, []),
    ("cf_device_count", st, [C.POINTER(C.c_int)]),
    ("cf_set_device", st, [C.c_int]),
    ("cf_sync", st, [vp]),
    ("cf_stream_create", st, [C.POINTER(vp)]),
    ("cf_stream_destroy", st, [vp]),
    ("cf_alloc", st, [C.c_size_t, C.POINTER(vp)]),
    ("cf_free", st, [vp]),
    ("cf_h2d", st, [vp, vp, C.c_size_t, vp]),
    ("cf_d2h", st, [vp, vp, C.c_size_t, vp]),
    ("cf_d2d", st, [vp, vp, C.c_size_t, vp]),
    ("cf_fill_zero", st, [vp, C.c_size_t, vp]),
    ("cf_cosmo_a", dbl, [C.POINTER(Cosmology), dbl]),
    ("cf_cosmo_H", dbl, [C.POINTER(Cosmology), dbl]),
    ("cf_cosmo_eta", dbl, [C.POINTER(Cosmology), dbl]),
    ("cf_inv_ah_sqr", dbl, [i64, dbl, dbl]),
    ("cf_rhs_f64", st, [i64, vp, vp, dbl, dbl, dbl, dbl, vp]),
    ("cf_scale_sum", st, [C.c_int, vp, C.POINTER(vp), C.POINTER(dbl), i64, vp]),
    ("cf_scale_sum2", st, [vp, vp, vp, dbl, dbl, i64, vp]),
    ("cf_scale_sum5", st, [vp, vp, vp, vp, vp, vp, dbl, dbl, dbl, dbl, dbl, i64, vp]),
    ("cf_norm_inf", st, [vp, i64, C.POINTER(dbl), vp]),
    ("cf_count_nonfinite", st, [vp, i64, C.POINTER(i64), vp]),
    ("cf_rk4_work_bytes", C.c_size_t, [i64, C.c_int]),
    ("cf_rk4_step_f64", st, [C.POINTER(KGParams), vp, dbl, dbl, vp, vp]),
    ("cf_rk4_run_f64", st, [C.POINTER(KGParams), vp, dbl, i64, i64, dbl, vp, vp]),
    ("cf_rk4_step_f32c", st, [C.POINTER(KGParams), vp, vp, dbl, dbl, vp, vp]),
    ("cf_rk4_run_f32c", st, [C.POINTER(KGParams), vp, vp, dbl, i64, i64, dbl, vp, vp]),
    ("cf_rk4_tb_work_bytes", C.c_size_t, [i64, C.c_int]),
    ("cf_rk4_run_tb_f64", st, [C.POINTER(KGParams), vp, dbl, i64, i64, dbl, C.c_int, vp, vp]),
    ("cf_rk4_run_tb_f64_noalloc", st, [C.POINTER(KGParams), vp, dbl, i64, i64, dbl, C.c_int, vp]),
    ("cf_rk4_step_tb_f64", st, [C.POINTER(KGParams), vp, vp, dbl, dbl, C.c_int, vp]),
    ("cf_rk4_step_tb_f32c", st, [C.POINTER(KGParams), vp, vp, vp, vp, dbl, dbl, C.c_int, vp]),
    ("cf_rk4_step_tb_f64_slab", st, [C.POINTER(KGParams), i64, vp, vp, dbl, dbl, C.c_int, vp]),
    ("cf_rk4_step_tb_f32c_slab", st, [C.POINTER(KGParams), i64, vp, vp, vp, vp, dbl, dbl, C.c_int, vp]),
    ("cf_rk4_step_tb_f32s_slab", st, [C.POINTER(KGParams), i64, vp, vp, vp, vp, dbl, dbl, C.c_int, vp]),
    ("cf_rk4_run_tb_f32c", st, [C.POINTER(KGParams), vp, vp, dbl, i64, i64, dbl, C.c_int, vp, vp]),
    ("cf_rk4_step_tb_f32s", st, [C.POINTER(KGParams), vp, vp, vp, vp, dbl, dbl, C.c_int, vp]),
    ("cf_rk4_run_tb_f32s", st, [C.POINTER(KGParams), vp, vp, dbl, i64, i64, dbl, C.c_int, vp, vp]),
    ("cf_rk4_tb_tune_pace", st, [C.POINTER(KGParams), vp, vp, dbl, dbl, C.c_int, vp, vp, C.POINTER(C.c_int)]),
    ("cf_rk4_tb_set_pace", st, [i64, C.c_int, C.c_int]),
    ("cf_rk4_tb_get_pace", C.c_int, [i64, C.c_int]),
    ("cf_rk4_tb_f32s_region", st, [i64, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    ("cf_rk4_tb_set_chunk", st, [i64, C.c_int]),
    ("cf_rk4_tb_get_chunk", C.c_int, [i64]),
    ("cf_ensemble_create", st, [C.c_int, vp, vp, dbl, dbl, C.POINTER(Cosmology), C.c_int, C.POINTER(vp)]),
    ("cf_ensemble_destroy", st, [vp]),
    ("cf_ensemble_sites", i64, [vp]),
    ("cf_ensemble_state", st, [vp, C.c_int, C.POINTER(vp), C.POINTER(vp)]),
    ("cf_ensemble_init_grf", st, [vp, vp, C.c_int, vp, dbl, vp]),
    ("cf_ensemble_set_plan", st, [vp, i64, C.c_int, C.c_int]),
    ("cf_ensemble_get_plan", st, [vp, i64, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    ("cf_ensemble_set_region", st, [vp, i64, C.c_int]),
    ("cf_ensemble_get_region", st, [vp, i64, C.POINTER(C.c_int)]),
    ("cf_ensemble_tune", st, [vp, dbl, C.c_int, C.POINTER(C.c_int)]),
    ("cf_ensemble_set_pace", st, [vp, C.c_int]),
    ("cf_ensemble_get_pace", C.c_int, [vp]),
    ("cf_ensemble_run", st, [vp, dbl, i64, i64, dbl, C.c_int, vp]),
    ("cf_ensemble_density_spectrum", st, [vp, C.c_int, dbl, vp, vp]),
    ("cf_adjoint_work_bytes", C.c_size_t, [i64, i64, i64]),
    ("cf_loss_spectrum_adjoint", st, [vp, C.POINTER(KGParams), vp, dbl, vp, vp, vp, C.POINTER(dbl),
                                      C.POINTER(dbl), vp]),
    ("cf_rk4_step_adjoint_f64", st, [C.POINTER(KGParams), vp, dbl, dbl, vp, vp, vp, vp]),
    ("cf_adjoint_spectrum_loss", st, [vp, C.POINTER(KGParams), vp, dbl, dbl, i64, i64, vp, vp, C.POINTER(dbl),
                                      C.POINTER(dbl), vp, vp]),
    ("cf_variant_create", st, [C.POINTER(KGParams), C.c_int, dbl, i64, vp, C.POINTER(vp), vp]),
    ("cf_variant_destroy", None, [vp]),
    ("cf_variant_rhs", st, [vp, vp, vp, dbl, vp]),
    ("cf_variant_rk4_step", st, [vp, vp, dbl, dbl, vp, vp]),
    ("cf_variant_rk4_run", st, [vp, vp, dbl, i64, i64, dbl, vp, vp]),
    ("cf_variant_rk4_run_tb", st, [vp, vp, dbl, i64, i64, dbl, C.c_int, vp, vp]),
    ("cf_variant_rk4_step_tb", st, [vp, vp, vp, dbl, dbl, C.c_int, vp]),
    ("cf_variant_energy_density", st, [vp, vp, vp, dbl, vp]),
    ("cf_variant_psi", st, [vp, dbl, vp, vp, vp]),
    ("cf_comoving_psi_fft", st, [i64, vp, vp, vp, dbl, dbl, vp]),
    ("cf_cutoff_fouriers", st, [i64, i64, vp, vp, vp]),
    ("cf_wkb_create", st, [C.POINTER(KGParams), vp, dbl, C.POINTER(vp), vp]),
    ("cf_wkb_evaluate", st, [vp, dbl, vp, vp]),
    ("cf_wkb_destroy", None, [vp]),
    ("cf_wkb_tables_host", st, [i64, dbl, dbl, dbl, dbl, dbl, dbl, vp, vp, vp, vp]),
    ("cf_split_f64", st, [vp, vp, vp, i64, vp]),
    ("cf_join_f64", st, [vp, vp, vp, i64, vp]),
    ("cf_energy_density_f64", st, [i64, vp, vp, dbl, dbl, dbl, vp]),
    ("cf_axis_average", st, [i64, vp, vp, vp]),
    ("cf_spectrum_plan_create", st, [i64, C.POINTER(vp)]),
    ("cf_spectrum_plan_destroy", st, [vp]),
    ("cf_spectrum_bins", i64, [i64]),
    ("cf_fft_d2z", st, [vp, vp, vp, vp]),
    ("cf_fft_z2d", st, [vp, vp, vp, vp]),
    ("cf_fft_z2d_inplace", st, [vp, vp, vp]),
    ("cf_power_spectrum", st, [vp, vp, vp, vp]),
    ("cf_power_spectrum_consume", st, [vp, vp, vp, vp]),
    ("cf_mode_power_spectrum", st, [vp, vp, dbl, dbl, dbl, vp, vp]),
    ("cf_bin_power", st, [vp, vp, vp, vp]),
    ("cf_bin_mode_power", st, [vp, vp, vp, dbl, dbl, dbl, vp, vp]),
    ("cf_radial_bins", st, [vp, vp, vp, vp]),
    ("cf_normals_host", st, [C.c_uint32, i64, vp]),
    ("cf_spectrum_table_host", st, [C.c_int, i64, dbl, vp, vp]),
    ("cf_grf", st, [vp, vp, vp, vp, vp, vp]),
    ("cf_scaled_fourier_modes", st, [vp, vp, vp, vp, vp]),
    ("cf_rescale_rms", st, [vp, i64, dbl, vp]),
    ("cf_init_grf_state", st, [vp, C.c_uint32, C.c_int, vp, dbl, dbl, vp, vp]),
]

EXPORTED = [s[0] for s in _SIGS]

_lib = None


def load():
    """Load libcfhip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"libcfhip.so not found at {LIB_PATH}; build it with `python -c \"import "
            f"__graft_entry__ as g; g.build()\"` or `make -C cosmic-fields-lite_amd`")
    lib = C.CDLL(LIB_PATH)
    for name, res, args in _SIGS:
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def check(code):
    if code != CF_OK:
        raise CfError(code, load().cf_last_error().decode(errors="replace"))
    return code
