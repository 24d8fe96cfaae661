"""ctypes binding of numrs_amd/libnrfft.so (the C ABI declared in include/nrfft.h).

The library is built in-tree by ``__graft_entry__.build()`` (make -C
numrs_amd/csrc).  There is no fallback: if the shared library is missing the
import of any compute entry point raises.
"""
from __future__ import annotations

import ctypes
import os
import re

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnrfft.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "nrfft.h")

# Status codes (include/nrfft.h).
OK = 0
E_ISIGN = -1
E_DIMS = -2
E_NOT_POW2 = -3
E_LEN = -4
E_EMPTY = -5
E_RESP_TOO_LONG = -6
E_HIP = -7
E_NOMEM = -8
E_UNSUPPORTED = -9
E_INVALID = -10

F32 = 0
F64 = 1
SPLIT_F16 = 1
SPLIT_BF16 = 2

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_vp = ctypes.c_void_p
_sz = ctypes.c_size_t
_szp = ctypes.POINTER(ctypes.c_size_t)
_i = ctypes.c_int

_SIGS = {
    "nrfft_version": (ctypes.c_char_p, []),
    "nrfft_status_string": (ctypes.c_char_p, [_i]),
    "nrfft_last_error": (ctypes.c_char_p, []),
    "nrfft_init": (_i, [_i]),
    "nrfft_four1_f64": (_i, [_dp, _sz, _i]),
    "nrfft_four1_f32": (_i, [_fp, _sz, _i]),
    "nrfft_four1_batch_f64": (_i, [_dp, _sz, _sz, _i]),
    "nrfft_four1_batch_f32": (_i, [_fp, _sz, _sz, _i]),
    "nrfft_realft_f64": (_i, [_dp, _sz, _i]),
    "nrfft_realft_f32": (_i, [_fp, _sz, _i]),
    "nrfft_realft_batch_f64": (_i, [ctypes.POINTER(_dp), _sz, _sz, _i]),
    "nrfft_realft_batch_f32": (_i, [ctypes.POINTER(_fp), _sz, _sz, _i]),
    "nrfft_fourn_f64": (_i, [_dp, _szp, _sz, _i]),
    "nrfft_fourn_f32": (_i, [_fp, _szp, _sz, _i]),
    "nrfft_convlv_f64": (_i, [_dp, _sz, _dp, _sz, _i, _dp]),
    "nrfft_convlv_f32": (_i, [_fp, _sz, _fp, _sz, _i, _fp]),
    "nrfft_convlv_batch_f64": (_i, [_dp, _sz, _dp, _sz, _sz, _i, _dp]),
    "nrfft_convlv_batch_f32": (_i, [_fp, _sz, _fp, _sz, _sz, _i, _fp]),
    "nrfft_correl_f64": (_i, [_dp, _dp, _sz, _dp]),
    "nrfft_correl_f32": (_i, [_fp, _fp, _sz, _fp]),
    "nrfft_plan_c2c": (_i, [ctypes.POINTER(_vp), _i, _szp, _sz, _sz]),
    "nrfft_execute_c2c": (_i, [_vp, _vp, _vp, _i, _vp]),
    "nrfft_plan_c2c_split": (_i, [ctypes.POINTER(_vp), _szp, _sz, _sz, _i]),
    "nrfft_plan_realft": (_i, [ctypes.POINTER(_vp), _i, _sz, _sz]),
    "nrfft_execute_realft": (_i, [_vp, _vp, _i, _vp]),
    "nrfft_plan_convlv": (_i, [ctypes.POINTER(_vp), _i, _sz, _sz, _sz]),
    "nrfft_execute_convlv": (_i, [_vp, _vp, _vp, _vp, _i, _vp]),
    "nrfft_twofft_f64": (_i, [_dp, _dp, _sz, _dp, _dp]),
    "nrfft_twofft_f32": (_i, [_fp, _fp, _sz, _fp, _fp]),
    "nrfft_cosft1_f64": (_i, [_dp, _sz]),
    "nrfft_cosft1_f32": (_i, [_fp, _sz]),
    "nrfft_cosft2_f64": (_i, [_dp, _sz, _i]),
    "nrfft_cosft2_f32": (_i, [_fp, _sz, _i]),
    "nrfft_rlft3_f64": (_i, [_dp, _dp, _sz, _sz, _sz, _i]),
    "nrfft_rlft3_f32": (_i, [_fp, _fp, _sz, _sz, _sz, _i]),
    "nrfft_fourfs_f64": (_i, [_dp, _dp, _szp, _sz, _i, _sz]),
    "nrfft_fourfs_f32": (_i, [_fp, _fp, _szp, _sz, _i, _sz]),
    "nrfft_plan_twofft": (_i, [ctypes.POINTER(_vp), _i, _sz, _sz]),
    "nrfft_execute_twofft": (_i, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "nrfft_plan_cosft1": (_i, [ctypes.POINTER(_vp), _i, _sz, _sz]),
    "nrfft_execute_cosft1": (_i, [_vp, _vp, _vp, _vp]),
    "nrfft_plan_cosft2": (_i, [ctypes.POINTER(_vp), _i, _sz, _sz]),
    "nrfft_execute_cosft2": (_i, [_vp, _vp, _vp, _i, _vp]),
    "nrfft_plan_rlft3": (_i, [ctypes.POINTER(_vp), _i, _sz, _sz, _sz, _sz]),
    "nrfft_execute_rlft3": (_i, [_vp, _vp, _vp, _i, _vp]),
    "nrfft_plan_correl": (_i, [ctypes.POINTER(_vp), _i, _sz, _sz]),
    "nrfft_execute_correl": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "nrfft_execute_c2c_launches": (_i, [_vp, _vp, _vp, _i, _vp, _i, _i]),
    "nrfft_plan_c2c_tuned": (_i, [ctypes.POINTER(_vp), _i, _szp, _sz, _sz, ctypes.POINTER(_i), _sz]),
    "nrfft_autotune_c2c": (_i, [_i, _szp, _sz, _sz, _i, ctypes.POINTER(_i), ctypes.POINTER(ctypes.c_double)]),
    "nrfft_plan_tune": (_i, [_vp, ctypes.POINTER(_i), _sz]),
    "nrfft_tune_cache_save": (_i, [ctypes.c_char_p]),
    "nrfft_tune_cache_load": (_i, [ctypes.c_char_p]),
    "nrfft_plan_uneven": (_i, [ctypes.POINTER(_vp), _i, _sz, ctypes.POINTER(_i), _szp, _szp]),
    "nrfft_execute_uneven": (_i, [_vp, _vp, _vp, _i, _vp]),
    "nrfft_execute_uneven_nograph": (_i, [_vp, _vp, _vp, _i, _vp]),
    "nrfft_uneven_check": (_i, [_vp, _vp]),
    "nrfft_uneven_info": (_i, [_vp, _szp, _szp, _szp, ctypes.POINTER(ctypes.c_uint)]),
    "nrfft_uneven_destroy": (_i, [_vp]),
    "nrfft_autotune_uneven": (_i, [_i, _i, _i, _i, ctypes.POINTER(_i), ctypes.POINTER(ctypes.c_float)]),
    "nrfft_fft_batch_f64": (_i, [ctypes.POINTER(_dp), _szp, _sz, _i]),
    "nrfft_fft_batch_f32": (_i, [ctypes.POINTER(_fp), _szp, _sz, _i]),
    "nrfft_gen_uniform_f32": (_i, [_vp, _sz, ctypes.c_uint64, ctypes.c_uint64, _vp]),
    "nrfft_gen_uniform_f64": (_i, [_vp, _sz, ctypes.c_uint64, ctypes.c_uint64, _vp]),
    "nrfft_execute_four1_batch_noalloc": (_i, [_i, _vp, _vp, _sz, _sz, _i, _vp]),
    "nrfft_plan_destroy": (_i, [_vp]),
    "nrfft_plan_workspace_bytes": (_sz, [_vp]),
    "nrfft_plan_num_launches": (_i, [_vp]),
}

_lib = None


def header_symbols() -> list[str]:
    """Names of every function include/nrfft.h declares."""
    with open(HEADER_PATH) as f:
        text = f.read()
    decl = re.compile(r"^\s*(?:const\s+)?(?:int|char\s*\*|size_t|void)\s*\*?\s*(nrfft_[a-z0-9_]+)\s*\(", re.M)
    return sorted(set(decl.findall(text)))


def lib() -> ctypes.CDLL:
    """Load libnrfft.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build the HIP extension first "
                "(python -c 'import __graft_entry__ as g; g.build()')")
        # torch ships its own libamdhip64 / libhsa-runtime64.  Loaded first,
        # they satisfy this library's libamdhip64.so.7 dependency and the
        # process has one HIP runtime; loaded after this library has started
        # the system runtime, torch's second copy finds no GPU.  So torch (when
        # installed) is imported before the library is opened.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def status_string(code: int) -> str:
    return lib().nrfft_status_string(code).decode()


def last_error() -> str:
    return lib().nrfft_last_error().decode()
