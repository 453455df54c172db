"""ctypes binding of libmgw.so (the C-ABI declared in include/mgw.h).

The library is loaded from the package directory; there is no fallback: a
missing or unloadable library raises immediately.
"""
from __future__ import annotations

import ctypes as C
import os

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'libmgw.so')

MGW_OK = 0
MGW_MEM_HOST, MGW_MEM_DEVICE = 0, 1
MGW_FP64, MGW_FP32, MGW_BF16X3, MGW_BF16X6, MGW_FP16X3 = 0, 1, 2, 3, 4
MGW_TECH_ONE_SHOT, MGW_TECH_IMAGINARY_PADE, MGW_TECH_CONTOUR = 0, 1, 2


class Ranges(C.Structure):
    """mgw_ranges: 1-based state ranges (module variables of the reference)."""
    _fields_ = [('ncore_G', C.c_int32), ('nvirtual_G', C.c_int32), ('nsemin', C.c_int32),
                ('nsemax', C.c_int32), ('nhomo_G', C.c_int32), ('ncore_W', C.c_int32),
                ('nvirtual_W', C.c_int32), ('nlumo_G', C.c_int32)]


class G0W0Params(C.Structure):
    _fields_ = [('technique', C.c_int32), ('nomega_sigma', C.c_int32), ('step_sigma', C.c_double),
                ('eta', C.c_double), ('nomega_chi_imag', C.c_int32), ('nomega_sigma_calc', C.c_int32),
                ('step_sigma_calc', C.c_double), ('spin_fact', C.c_double), ('precision', C.c_int32),
                ('nomega_chi_real', C.c_int32)]


class Shells(C.Structure):
    """mgw_shells: a set of contracted Gaussian shells (host arrays)."""
    _fields_ = [('nshell', C.c_int32), ('am', C.POINTER(C.c_int32)), ('x0', C.POINTER(C.c_double)),
                ('prim_offset', C.POINTER(C.c_int32)), ('alpha', C.POINTER(C.c_double)),
                ('coeff', C.POINTER(C.c_double)), ('transform_offset', C.POINTER(C.c_int32)),
                ('transform', C.POINTER(C.c_double)), ('function_offset', C.POINTER(C.c_int32))]


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_vp = C.c_void_p

# name -> (restype, argtypes); exactly the functions declared in include/mgw.h
SIGNATURES = {
    'mgw_abi_version': (C.c_int32, []),
    'mgw_last_error': (C.c_char_p, []),
    'mgw_ctx_create': (C.c_int, [C.c_int, C.POINTER(_vp)]),
    'mgw_ctx_destroy': (C.c_int, [_vp]),
    'mgw_ctx_synchronize': (C.c_int, [_vp]),
    'mgw_ctx_trim': (C.c_int, [_vp, C.c_int]),
    'mgw_autotune': (C.c_int, [C.c_int, C.POINTER(C.c_int)]),
    'mgw_ctx_graphs': (C.c_int, [_vp, C.c_int]),
    'mgw_set_energy_g': (C.c_int, [_vp, C.POINTER(C.c_double)]),
    'mgw_memory_report': (C.c_int, [_vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    'mgw_device_pool_report': (C.c_int, [C.c_int, C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    'mgw_device_pool_limit': (C.c_int, [C.c_int, C.c_int64]),
    'mgw_device_alloc': (C.c_int, [C.c_int, C.c_int64, C.POINTER(C.c_void_p)]),
    'mgw_device_free': (C.c_int, [C.c_void_p]),
    'mgw_device_copy': (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int]),
    'mgw_device_memory': (C.c_int, [C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    'mgw_precision_escalations': (C.c_int, [_vp, C.POINTER(C.c_int64)]),
    'mgw_footprint_bytes': (C.c_int64, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    'mgw_footprint_bytes_nse': (C.c_int64, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                            C.c_int]),
    'mgw_eri_2center_ao': (C.c_int, [_vp, C.POINTER(Shells), C.c_int, _vp]),
    'mgw_eri_3center_ao': (C.c_int, [_vp, C.POINTER(Shells), C.POINTER(Shells), C.c_int64, _ip, C.c_int, _vp]),
    'mgw_hartree_exchange_ri': (C.c_int, [_vp, _dp, C.c_int, C.c_int, C.c_double, _dp, _dp]),
    'mgw_exchange_m_vxc': (C.c_int, [_vp, _dp, _dp, _dp]),
    'mgw_set_eri_3center': (C.c_int, [_vp, _vp, C.c_int64, C.c_int, _ip, C.c_int, C.c_int]),
    'mgw_destroy_eri_3center': (C.c_int, [_vp]),
    'mgw_inverse_sqrt_eri_2center': (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.POINTER(C.c_int), _dp, _vp]),
    'mgw_build_eri_3center': (C.c_int, [_vp, _vp, C.c_int64, C.c_int64, C.c_int64, _ip, C.c_int, C.c_int,
                                        _vp]),
    'mgw_set_mo': (C.c_int, [_vp, _vp, C.c_int, C.c_int, _vp, _vp, C.c_int]),
    'mgw_set_mo_spin': (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int, _vp, _vp, C.c_int]),
    'mgw_set_spin': (C.c_int, [_vp, C.c_int]),
    'mgw_calculate_eri_3center_mo': (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, _dp]),
    'mgw_destroy_eri_3center_mo': (C.c_int, [_vp]),
    'mgw_calculate_eri_3center_mo_noalloc': (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p]),
    'mgw_npole': (C.c_int, [_vp, C.POINTER(Ranges), C.POINTER(C.c_int)]),
    'mgw_set_spectral_function': (C.c_int, [_vp, C.c_int, C.c_int, _vp, _vp, C.c_int]),
    'mgw_gw_selfenergy_qs': (C.c_int, [_vp, C.POINTER(Ranges), C.c_double, C.c_double, _dp, _dp, _dp]),
    'mgw_gw_ingredients': (C.c_int, [_vp, C.POINTER(Ranges), _dp, _dp]),
    'mgw_polarizability': (C.c_int, [_vp, C.POINTER(Ranges), C.c_double, C.c_int, _dp, _dp, _dp, _dp]),
    'mgw_polarizability_grid': (C.c_int, [_vp, C.POINTER(Ranges), C.c_double, C.c_int, C.c_int, _dp, _dp,
                                          _dp, _dp, _dp]),
    'mgw_gw_selfenergy': (C.c_int, [_vp, C.POINTER(Ranges), C.c_double, C.c_double, C.c_int, _dp, _dp,
                                    _dp, _dp]),
    'mgw_gw_selfenergy_imag': (C.c_int, [_vp, C.POINTER(Ranges), C.c_int, _dp, _dp]),
    'mgw_gw_selfenergy_contour': (C.c_int, [_vp, C.POINTER(Ranges), C.c_double, C.c_int, C.c_int, C.c_double,
                                            C.c_int, _dp, _dp, _dp, _dp]),
    'mgw_pade_fit': (C.c_int, [_vp, C.c_int, C.c_int, _dp, _dp, C.c_int, _dp, C.c_double, _dp]),
    'mgw_find_qp_energy': (C.c_int, [_vp, C.POINTER(Ranges), C.c_int, _dp, _dp, _dp, _dp, _dp, C.c_int,
                                     _dp, _dp, _dp]),
    'mgw_g0w0': (C.c_int, [_vp, C.POINTER(G0W0Params), C.POINTER(Ranges), _dp, _dp, _dp, _dp, _dp, _dp]),
    'mgw_profile_enable': (C.c_int, [_vp, C.c_int]),
    'mgw_profile_read': (C.c_int, [_vp, C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_double),
                                   C.POINTER(C.c_double), C.POINTER(C.c_double)]),
}

KCLASS = ['gemm_ao2mo', 'gemm_casida', 'gemm_sigma', 'sigma_poles', 'eigensolver', 'gemm_chi0',
          'cholesky', 'gemm_wb', 'gemm_other', 'other']

_lib = None


def load() -> C.CDLL:
    """Load libmgw.so (raises OSError when it is missing: no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f'{LIB_PATH} not found: build it with `python -m molgw_amd.build`')
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


class MgwError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f'libmgw error {code}: {msg}')
        self.code = code


def check(code: int) -> None:
    if code != MGW_OK:
        raise MgwError(code, load().mgw_last_error().decode(errors='replace'))
