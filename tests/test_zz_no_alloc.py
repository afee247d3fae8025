"""ffcz_k_extract_noalloc, the allocation-free extraction entry, against the
allocating ffcz_k_extract it shadows, through the C ABI: flags, compact
values, symbols, count and range identical bit for bit, on two calls in a row
(the second with a smaller array, so scratch left by the first must not
leak into it), for both quantisers and both types; and the size limit is an
error before any device work."""
import ctypes as C

import numpy as np
import pytest
import torch

from ffcz_amd._lib import FFCZ_F32, FFCZ_F64, FFCzError, call, load
from oracle import ffcz_oracle as O

pytestmark = pytest.mark.gpu

NOALLOC_MAX = 1 << 30
ERR_CAPACITY = 8


def _run(name, dt, dv, n, kind, step, stream):
    dev = dv.device
    pack = torch.full((((n + 63) // 64) * 8,), 0xA5, dtype=torch.uint8, device=dev)
    comp = torch.full((n,), -7.0, dtype=dv.dtype, device=dev)
    quant = torch.full((n,), 0x5A5A, dtype=torch.int16, device=dev)
    cnt, lo, hi = C.c_size_t(), C.c_double(), C.c_double()
    torch.cuda.synchronize()
    call(name, dt, dv.data_ptr(), n, kind, step, pack.data_ptr(), comp.data_ptr(), quant.data_ptr(), C.byref(cnt),
         C.byref(lo), C.byref(hi), stream)
    k = cnt.value
    return (pack.cpu().numpy()[: (n + 7) // 8], comp.cpu().numpy()[:k], quant.cpu().numpy().view(np.uint16)[:k], k,
            lo.value, hi.value)


@pytest.mark.parametrize("dtype, dt", [(np.float32, FFCZ_F32), (np.float64, FFCZ_F64)])
@pytest.mark.parametrize("kind, step", [(1, 0.0), (2, 2.5e-5)])
def test_noalloc_matches_allocating_entry_on_two_calls(gpu, dtype, dt, kind, step):
    s = torch.cuda.Stream(device=gpu)
    for n, seed in ((1_000_003, 1), (70_001, 2)):
        rng = np.random.default_rng(seed)
        v = (rng.standard_normal(n) * 1e-3).astype(dtype)
        v[rng.random(n) < 0.6] = 0.0
        dv = torch.from_numpy(v).to(gpu)
        a = _run("ffcz_k_extract", dt, dv, n, kind, step, None)
        b = _run("ffcz_k_extract_noalloc", dt, dv, n, kind, step, C.c_void_p(s.cuda_stream))
        for x, y in zip(a[:3], b[:3]):
            assert np.array_equal(x, y)
        assert a[3:] == b[3:]
        # and the reference arithmetic (projection_algorithm.cu:363-550)
        if kind == 1:
            ref = O.extract_one(v)
            assert b[3] == ref.compact.size and np.array_equal(b[0], ref.pack)
            assert np.array_equal(b[2], ref.quant) and (b[4], b[5]) == (ref.vmin, ref.vmax)
        else:
            q = O.quantize_step(v, step)
            assert b[3] == int(np.count_nonzero(q))
            assert np.array_equal(b[2], O.zigzag(q[q != 0]).astype(np.uint16))


def test_noalloc_rejects_arrays_beyond_its_scratch(gpu):
    dv = torch.zeros(64, dtype=torch.float32, device=gpu)
    cnt, lo, hi = C.c_size_t(), C.c_double(), C.c_double()
    rc = load().ffcz_k_extract_noalloc(FFCZ_F32, dv.data_ptr(), NOALLOC_MAX + 1, 1, 0.0, dv.data_ptr(),
                                       dv.data_ptr(), dv.data_ptr(), C.byref(cnt), C.byref(lo), C.byref(hi), None)
    assert rc == ERR_CAPACITY
    assert b"2^30" in load().ffcz_last_error()
    with pytest.raises(FFCzError):
        call("ffcz_k_extract_noalloc", FFCZ_F32, dv.data_ptr(), NOALLOC_MAX + 1, 1, 0.0, dv.data_ptr(),
             dv.data_ptr(), dv.data_ptr(), C.byref(cnt), C.byref(lo), C.byref(hi), None)
