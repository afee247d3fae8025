"""The command line end to end on the GPU: raw files in, the container and
the decoded field out; the decoded field meets the certified bound, and the
decompress-only path (-x) writes the same bytes."""
import numpy as np
import pytest

from ffcz_amd import cli, container
from oracle import ffcz_oracle as O

pytestmark = pytest.mark.gpu


# (69, 69, 115): a field shape with prime factor 23 (run-time planned kernels)
@pytest.mark.parametrize("shape, flag, dtype, tflag", [((16, 20, 30), "-3", np.float32, "-f"),
                                                       ((1, 64, 64), "-2", np.float64, "-d"),
                                                       ((69, 69, 115), "-3", np.float32, "-f")])
def test_compress_then_decompress(gpu, tmp_path, shape, flag, dtype, tflag):
    rng = np.random.default_rng(5)
    x = rng.random(shape).astype(dtype)
    eps = 1e-3
    base = (np.round(x / (2 * eps)) * (2 * eps)).astype(dtype)  # uniform quantisation, |base - x| <= eps
    fo, fb, fz, fd, fd2 = (str(tmp_path / n) for n in ("o.raw", "b.raw", "c.ffcz", "d.raw", "d2.raw"))
    x.tofile(fo)
    base.tofile(fb)
    dims = [str(d) for d in (shape[1:] if flag == "-2" else shape)]
    lines = []
    argv = ["-i", fo, "-e", fb, "-z", fz, "-o", fd, flag, *dims, tflag, "-M", "ABS", str(eps), "-F", "REL", "1e-3",
            "--margin", "0.125", "--step-quant", "--zstd", "3"]
    r = cli.compress(cli.parse(argv), out=lines.append)
    assert any(l.startswith("# Active spatial edits:") for l in lines)
    assert r["certified"]
    # REL frequency bound: 1e-3 * max |F(x)|
    assert np.isclose(r["delta"], 1e-3 * np.abs(np.fft.rfftn(x.astype(np.float64))).max(), rtol=1e-5)
    xh = np.fromfile(fd, dtype=dtype).reshape(shape)
    assert np.abs(xh.astype(np.float64) - x).max() <= eps * (1 + 1e-6)
    # the frequency bound too, checked by the oracle in FP64
    assert O.certify(xh.astype(np.float64) - x, eps * (1 + 1e-6), r["delta"] * (1 + 1e-6))["ok"]
    info = container.info(open(fz, "rb").read())
    assert info["shape"] == shape and info["n_active_spatial"] == r["n_active_spatial"]
    assert cli.main(["-x", "-z", fz, "-e", fb, "-o", fd2]) == 0
    assert open(fd, "rb").read() == open(fd2, "rb").read()


def test_built_in_base_compressor(gpu, tmp_path):
    shape = (16, 20, 30)
    x = np.random.default_rng(6).random(shape).astype(np.float32)
    eps = 1e-3
    fo, fz, fd, fd2 = (str(tmp_path / n) for n in ("o.raw", "c.ffcz", "d.raw", "d2.raw"))
    x.tofile(fo)
    lines = []
    argv = ["-i", fo, "--base", "-z", fz, "-o", fd, "-3", *map(str, shape), "-f", "-M", "ABS", str(eps), "-F", "REL",
            "1e-3", "--margin", "0.125", "--step-quant", "--zstd", "3"]
    r = cli.compress(cli.parse(argv), out=lines.append)
    assert r["certified"] and any(l.startswith("Base stream:") for l in lines)
    xh = np.fromfile(fd, dtype=np.float32).reshape(shape)
    assert np.abs(xh.astype(np.float64) - x).max() <= eps * (1 + 1e-6)
    assert cli.main(["-x", "--base", "-z", fz, "-o", fd2]) == 0
    assert open(fd, "rb").read() == open(fd2, "rb").read()
