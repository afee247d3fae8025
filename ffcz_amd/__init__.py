"""MI355X-native FFCz alternating-projection post-processor.

The compute path is ``libffcz.so`` (hand-written HIP kernels for gfx950 behind
the C ABI of ``include/ffcz.h``); this package is its Python host mirror.
"""
from ._lib import FFCzError, load  # noqa: F401
from .solver import ProjectionSolver, CertResult, set_plan_cache, plan_cache, MODE_ABS, MODE_DIRECT, MODE_PARTIAL_PTW, MODE_PTW, QUANT_REF16, QUANT_STEP  # noqa: F401
from ._lib import FFT_VALU, FFT_BF16X3, FFT_FP16X3, FFT_VARIANTS  # noqa: F401
from .entropy import (  # noqa: F401
    CompressionResult,
    cuda_huffman_zstd_compress,
    cuda_huffman_zstd_decompress,
    huffman_encode,
)
from .fft import fft_r2c, fft_c2r  # noqa: F401
from .batch import ProjectionBatch  # noqa: F401
from . import container  # noqa: F401

__all__ = [
    "ProjectionSolver", "ProjectionBatch", "CertResult", "FFCzError", "load", "fft_r2c", "fft_c2r",
    "cuda_huffman_zstd_compress", "cuda_huffman_zstd_decompress", "huffman_encode", "CompressionResult",
    "MODE_ABS", "MODE_DIRECT", "MODE_PARTIAL_PTW", "MODE_PTW", "QUANT_REF16", "QUANT_STEP",
    "FFT_VALU", "FFT_BF16X3", "FFT_FP16X3", "FFT_VARIANTS", "container", "set_plan_cache", "plan_cache",
]
