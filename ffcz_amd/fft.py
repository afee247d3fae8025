"""3D real FFT entry points of the library (cuFFT R2C/C2R layout:
compact half spectrum [Nx][Ny][Nz//2+1], interleaved complex, unnormalised
inverse).  Inputs and outputs are GPU torch tensors."""
from __future__ import annotations

import torch

from ._lib import FFCZ_F32, FFCZ_F64, call


def _dt(t: torch.Tensor) -> int:
    if t.dtype in (torch.float32, torch.complex64):
        return FFCZ_F32
    if t.dtype in (torch.float64, torch.complex128):
        return FFCZ_F64
    raise TypeError(f"unsupported dtype {t.dtype}")


def fft_r2c(x: torch.Tensor, variant: int = 0) -> torch.Tensor:
    """3D R2C of a device field (cuFFT layout, unnormalised); ``variant``
    selects the FFT kernels (0 vector ALU, 1 MFMA BF16x3, 2 MFMA FP16x3)."""
    assert x.is_cuda and x.dim() == 3
    x = x.contiguous()
    nx, ny, nz = x.shape
    cdt = torch.complex64 if x.dtype == torch.float32 else torch.complex128
    out = torch.empty((nx, ny, nz // 2 + 1), dtype=cdt, device=x.device)
    torch.cuda.synchronize(x.device)
    call("ffcz_fft_r2c_prec", _dt(x), int(variant), nx, ny, nz, x.data_ptr(), out.data_ptr(), None)
    return out


def fft_c2r(X: torch.Tensor, nz: int, variant: int = 0) -> torch.Tensor:
    """3D C2R (unnormalised) of a device half spectrum; ``variant`` as for fft_r2c."""
    assert X.is_cuda and X.dim() == 3
    X = X.contiguous()
    nx, ny, nh = X.shape
    assert nh == nz // 2 + 1
    rdt = torch.float32 if X.dtype == torch.complex64 else torch.float64
    out = torch.empty((nx, ny, nz), dtype=rdt, device=X.device)
    torch.cuda.synchronize(X.device)
    call("ffcz_fft_c2r_prec", _dt(X), int(variant), nx, ny, nz, X.data_ptr(), out.data_ptr(), None)
    return out
