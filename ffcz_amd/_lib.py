"""ctypes binding of ``libffcz.so`` (the C ABI declared in ``include/ffcz.h``).

The library is the product path: if it is missing this module raises, there is
no fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

_PKG = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("FFCZ_LIB", _PKG / "libffcz.so"))

FFCZ_OK = 0
FFCZ_ERR_INVALID = 1
FFCZ_ERR_STATE = 2
FFCZ_ERR_HIP = 3
FFCZ_ERR_RANGE = 4
FFCZ_ERR_ALLOC = 5
FFCZ_ERR_ZSTD = 6

FFCZ_F32 = 0
FFCZ_F64 = 1
MODE_DIRECT, MODE_ABS, MODE_PTW, MODE_PARTIAL_PTW = 0, 1, 2, 3
QUANT_REF16, QUANT_STEP = 1, 2
FFT_VALU, FFT_BF16X3, FFT_FP16X3 = 0, 1, 2
FFT_VARIANTS = {"valu-fp32": FFT_VALU, "mfma-bf16x3": FFT_BF16X3, "mfma-fp16x3": FFT_FP16X3}


class FFCzError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[ffcz status {code}] {msg}")
        self.code = code


class Cert(C.Structure):
    _fields_ = [
        ("spatial_violations", C.c_ulonglong),
        ("freq_violations", C.c_ulonglong),
        ("spatial_max_excess", C.c_double),
        ("freq_max_excess", C.c_double),
        ("ok", C.c_int),
    ]


_vp = C.c_void_p
_sz = C.c_size_t
_szp = C.POINTER(C.c_size_t)
_i = C.c_int
_d = C.c_double
_f = C.c_float
_h = C.c_void_p  # opaque handles

# name -> argtypes (all return int unless listed in _RESTYPES)
SIGNATURES = {
    "ffcz_last_error": [],
    "ffcz_abi_version": [],
    "ffcz_create": [_i, _sz, _sz, _sz, C.POINTER(_h)],
    "ffcz_destroy": [_h],
    "ffcz_init_all_abs_f32": [_h, _vp, _f, _f],
    "ffcz_init_all_abs_f64": [_h, _vp, _d, _d],
    "ffcz_init_direct_f32": [_h, _vp, _vp, _vp, _f, _sz],
    "ffcz_init_direct_f64": [_h, _vp, _vp, _vp, _d, _sz],
    "ffcz_init_all_ptw_f32": [_h, _vp, _f, _f],
    "ffcz_init_all_ptw_f64": [_h, _vp, _d, _d],
    "ffcz_init_partial_ptw_f32": [_h, _vp, _vp, _f, _f, _sz],
    "ffcz_init_partial_ptw_f64": [_h, _vp, _vp, _d, _d, _sz],
    "ffcz_solve_f32": [_h, _vp, _i, _f],
    "ffcz_solve_f64": [_h, _vp, _i, _d],
    "ffcz_solve_all_abs_f32": [_h, _vp, _i, _f],
    "ffcz_solve_all_abs_f64": [_h, _vp, _i, _d],
    "ffcz_solve_all_ptw_f32": [_h, _vp, _i, _f],
    "ffcz_solve_all_ptw_f64": [_h, _vp, _i, _d],
    "ffcz_solve_partial_ptw_f32": [_h, _vp, _i, _f],
    "ffcz_solve_partial_ptw_f64": [_h, _vp, _i, _d],
    "ffcz_set_ptw_semantics": [_h, _i],
    "ffcz_solve_loop": [_h, _i, _vp, _i, _d],
    "ffcz_get_status": [_h, C.POINTER(_i), C.POINTER(_d), C.POINTER(_i)],
    "ffcz_extract": [_h],
    "ffcz_reset_edit_accumulators": [_h],
    "ffcz_get_results": [_h, _vp, _sz, _szp],
    "ffcz_get_freq_flags": [_h, _vp, _sz, _szp],
    "ffcz_get_spatial_flags": [_h, _vp, _sz, _szp],
    "ffcz_get_freq_flags_pack": [_h, _vp, _sz, _szp],
    "ffcz_get_spatial_flags_pack": [_h, _vp, _sz, _szp],
    "ffcz_get_active_freq_quant": [_h, _vp, _sz, _szp],
    "ffcz_get_active_spatial_quant": [_h, _vp, _sz, _szp],
    "ffcz_get_active_freq_edits": [_h, _vp, _sz, _szp],
    "ffcz_get_active_spatial_edits": [_h, _vp, _sz, _szp],
    "ffcz_get_freq_edits_full": [_h, _vp, _sz, _szp],
    "ffcz_get_spatial_edits_full": [_h, _vp, _sz, _szp],
    "ffcz_get_edit_range": [_h, C.POINTER(_d), C.POINTER(_d), C.POINTER(_d), C.POINTER(_d)],
    "ffcz_set_margins": [_h, _d, _d],
    "ffcz_set_quantizer": [_h, _i, _d, _d],
    "ffcz_batch_create": [_i, _sz, C.POINTER(_sz), C.POINTER(_h)],
    "ffcz_batch_destroy": [_h],
    "ffcz_batch_size": [_h, _szp],
    "ffcz_batch_field": [_h, _sz, C.POINTER(_h)],
    "ffcz_batch_solve": [_h, _i, C.POINTER(_vp), _i, _d, _i],
    "ffcz_batch_status": [_h, C.POINTER(_i), C.POINTER(_d), C.POINTER(_i), _sz, C.POINTER(_i)],
    "ffcz_batch_time_iteration": [_h, _i, C.POINTER(_f)],
    "ffcz_batch_extract": [_h],
    "ffcz_batch_certify": [_h, _i, C.POINTER(Cert), _sz],
    "ffcz_set_fft_precision": [_h, _i],
    "ffcz_autotune": [_h, _i, _i, C.POINTER(_i), C.POINTER(_f)],
    "ffcz_get_tile_vars": [_h, C.POINTER(_i)],
    "ffcz_get_tune_source": [_h, C.POINTER(_i)],
    "ffcz_set_plan_cache": [C.c_char_p],
    "ffcz_get_plan_cache": [C.c_char_p, _sz, _szp],
    "ffcz_set_tile_vars": [_h, C.POINTER(_i)],
    "ffcz_set_cert_pair": [_h, _i],
    "ffcz_get_fft_precision": [_h, C.POINTER(_i)],
    "ffcz_certify": [_h, _i, C.POINTER(Cert)],
    "ffcz_certify_both": [_h, C.POINTER(Cert)],
    "ffcz_certify_both_enqueue": [_h],
    "ffcz_certify_both_results": [_h, C.POINTER(Cert)],
    "ffcz_set_tuning": [_i, C.c_long],
    "ffcz_get_tuning": [C.POINTER(_i), C.POINTER(C.c_long)],
    "ffcz_get_stream": [_h, C.POINTER(_vp)],
    "ffcz_time_iteration": [_h, _i, C.POINTER(_f)],
    "ffcz_set_pass_timing": [_h, _i],
    "ffcz_set_graph": [_h, _i],
    "ffcz_set_borrow_e0": [_h, _i],
    "ffcz_get_pass_timing": [_h, C.POINTER(_f), C.POINTER(C.c_long)],
    "ffcz_get_pass_counts": [_h, C.POINTER(C.c_long)],
    "ffcz_huffman_zstd_compress_u16": [_vp, _sz, _i, C.POINTER(_h)],
    "ffcz_huffman_zstd_compress_u8": [_vp, _sz, _i, C.POINTER(_h)],
    "ffcz_huffman_encode_u16": [_vp, _sz, C.POINTER(_h)],
    "ffcz_huffman_encode_u8": [_vp, _sz, C.POINTER(_h)],
    "ffcz_coded_info": [_h, _szp, _szp, _szp, _szp, C.POINTER(C.c_ulonglong)],
    "ffcz_coded_data": [_h, _vp, _sz],
    "ffcz_coded_device_ms": [_h, C.POINTER(_f)],
    "ffcz_coded_bitstream": [_h, _vp, _sz],
    "ffcz_coded_table": [_h, _vp, _vp, _vp, _sz],
    "ffcz_coded_free": [_h],
    "ffcz_huffman_zstd_decompress_u16": [_vp, _sz, _vp, _vp, _vp, _sz, _sz, _sz, _vp, _szp],
    "ffcz_huffman_zstd_decompress_u8": [_vp, _sz, _vp, _vp, _vp, _sz, _sz, _sz, _vp, _szp],
    "ffcz_container_pack": [_h, _i, C.POINTER(_h)],
    "ffcz_blob_size": [_h, _szp],
    "ffcz_blob_copy": [_h, _vp, _sz],
    "ffcz_blob_free": [_h],
    "ffcz_container_info": [_vp, _sz, C.POINTER(_i), _szp, C.POINTER(_d), C.POINTER(_d), C.POINTER(_i), _szp],
    "ffcz_container_decode": [_vp, _sz, _vp, _vp, _vp],
    "ffcz_base_compress": [_i, _vp, _sz, _sz, _sz, _d, _i, _vp, C.POINTER(_h), _vp],
    "ffcz_base_info": [_vp, _sz, C.POINTER(_i), _szp, C.POINTER(_d), _szp],
    "ffcz_base_decompress": [_vp, _sz, _vp, _vp],
    "ffcz_fft_r2c": [_i, _sz, _sz, _sz, _vp, _vp, _vp],
    "ffcz_fft_c2r": [_i, _sz, _sz, _sz, _vp, _vp, _vp],
    "ffcz_fft_r2c_prec": [_i, _i, _sz, _sz, _sz, _vp, _vp, _vp],
    "ffcz_fft_c2r_prec": [_i, _i, _sz, _sz, _sz, _vp, _vp, _vp],
    "ffcz_k_project_spatial": [_i, _vp, _d, _vp, _sz, _vp],
    "ffcz_k_project_freq_abs": [_i, _vp, _d, _vp, _sz, _vp],
    "ffcz_k_convergence": [_i, _vp, _d, _sz, C.POINTER(_d)],
    "ffcz_k_quantize_dequantize": [_i, _vp, _vp, _d, _d, _sz, _d, _vp],
    "ffcz_k_extract": [_i, _vp, _sz, _i, _d, _vp, _vp, _vp, _szp, C.POINTER(_d), C.POINTER(_d), _vp],
    "ffcz_k_extract_noalloc": [_i, _vp, _sz, _i, _d, _vp, _vp, _vp, _szp, C.POINTER(_d), C.POINTER(_d), _vp],
}
_RESTYPES = {"ffcz_last_error": C.c_char_p, "ffcz_coded_free": None, "ffcz_blob_free": None}

_lib = None


def load() -> C.CDLL:
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} not found: build it with `python -m ffcz_amd.build`")
    lib = C.CDLL(str(LIB_PATH))
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RESTYPES.get(name, C.c_int)
    _lib = lib
    return lib


def check(status: int) -> None:
    if status != FFCZ_OK:
        msg = load().ffcz_last_error()
        raise FFCzError(status, msg.decode() if msg else "")


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args))


def set_tuning(force_generic: int = -1, slab_mb: int | None = None):
    """Process-wide kernel tuning (see ffcz_set_tuning); returns the previous
    (force_generic, slab_mb)."""
    g, m = C.c_int(), C.c_long()
    call("ffcz_get_tuning", C.byref(g), C.byref(m))
    call("ffcz_set_tuning", int(force_generic), int(m.value if slab_mb is None else slab_mb))
    return g.value, m.value
