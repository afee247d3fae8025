"""``fftz``-compatible command line (projection_example.cu:30-265) over the
MI355X library, with raw file I/O (fileIO.cpp:35-76).

Compress (the reference's flags)::

    python -m ffcz_amd.cli -i orig.f32 -e base.f32 -3 512 512 512 -f \\
        -M ABS 1e-4 -F REL 1e-4 -z edits.ffcz [-o decompressed.f32]

Decompress (no counterpart in the reference, which cannot decode its output)::

    python -m ffcz_amd.cli -x -z edits.ffcz -e base.f32 -o decompressed.f32

With ``--base`` the library's own GPU base compressor (dual-quantisation
Lorenzo, bound -M) produces x' instead of the -e file, and its stream is
written to ``<-z>.base``; ``-x --base -z edits.ffcz -o out`` decodes both.

Differences from the reference driver:
  * ``-M REL xi`` and ``-F REL xi`` work (the reference parses and ignores
    them): the spatial bound becomes xi * (max x - min x), the frequency bound
    xi * max |F(x)|, the form the reference's configurations state.
  * ``-z`` receives the self-describing container of the coded edits and
    ``-o`` the decoded field x_hat (both unused in the reference).
  * Timings are measured wall times, not the reference's synthetic formula.
  * Errors exit with status 2 and a message instead of ``exit(EXIT_FAILURE)``
    after the usage text; the usage text is the reference's, plus the new flags.
"""
from __future__ import annotations

import os
import sys
import time
from dataclasses import dataclass, field

import numpy as np

USAGE = """Usage:
  -i <file_path>  : Specify the file of original data
  -e <file_path>  : Specify the reconstructed file of base compressor
  -z <file_path>  : Specify the compressed file
  -o <file_path>  : Specify the decompressed file (optional)
  -1 <nx>                : 1D data with <nx> values
  -2 <nx> <ny>           : 2D data with <nx> * <ny> values
  -3 <nx> <ny> <nz>      : 3D data with <nx> * <ny> * <nz> values
  -f              : Use float data type
  -d              : Use double data type
  -M ABS <xi>     : Specify the absolute error bound in spatial domain
  -M REL <xi>     : Specify the relative error bound in spatial domain
  -F ABS <xi>     : Specify the absolute error bound in frequency domain
  -F REL <xi>     : Specify the relative error bound in frequency domain
  -x              : Decompress -z with the base reconstruction -e into -o
  --max-iter <n>  : Projection iterations (default 100)
  --tol <t>       : Convergence tolerance (default 1e-7)
  --zstd <level>  : ZSTD level of the coded streams (default 20)
  --step-quant    : Quantise edits with step eps/8 and delta/8 (default: the reference's 16-bit quantiser)
  --margin <m>    : Iterate against bounds shrunk by m, m in [0, 1) (default 0)
  --base          : Produce x' with the built-in GPU base compressor (no -e); its stream goes to <-z>.base
"""


class UsageError(Exception):
    pass


@dataclass
class Args:
    original: str | None = None
    base: str | None = None
    compressed: str | None = None
    decompressed: str | None = None
    shape: tuple = ()
    dtype: type = np.float32
    spatial_abs: bool = True
    spatial: float | None = None
    freq_abs: bool = True
    freq: float | None = None
    decompress: bool = False
    max_iter: int = 100
    tol: float = 1e-7
    zstd: int = 20
    step_quant: bool = False
    margin: float = 0.0
    base_gpu: bool = False
    extra: dict = field(default_factory=dict)


def parse(argv) -> Args:
    """The reference's argument loop (projection_example.cu:54-154), with its
    dimension conventions: -1 n is (1, 1, n) and -2 nx ny is (1, nx, ny)."""
    a = Args()
    i = 0

    def need(k, what):
        if i + k >= len(argv):
            raise UsageError(f"Missing {what}")

    while i < len(argv):
        arg = argv[i]
        if arg in ("-i", "-e", "-z", "-o"):
            need(1, {"-i": "input file path", "-e": "base reconstructed file path", "-z": "compressed file path",
                     "-o": "decompressed file path"}[arg])
            i += 1
            setattr(a, {"-i": "original", "-e": "base", "-z": "compressed", "-o": "decompressed"}[arg], argv[i])
        elif arg in ("-1", "-2", "-3"):
            k = int(arg[1])
            need(k, "size of the data")
            dims = [int(v) for v in argv[i + 1: i + 1 + k]]
            if any(d < 1 for d in dims):
                raise UsageError("Sizes must be positive")
            a.shape = tuple([1] * (3 - k) + dims)
            i += k
        elif arg == "-f":
            a.dtype = np.float32
        elif arg == "-d":
            a.dtype = np.float64
        elif arg in ("-M", "-F"):
            need(2, "spatial error bound" if arg == "-M" else "frequency error bound")
            kind, val = argv[i + 1], float(argv[i + 2])
            if kind not in ("ABS", "REL"):
                raise UsageError(f"{arg} takes ABS or REL")
            if val <= 0:
                raise UsageError("Error bounds must be positive")
            if arg == "-M":
                a.spatial_abs, a.spatial = kind == "ABS", val
            else:
                a.freq_abs, a.freq = kind == "ABS", val
            i += 2
        elif arg == "-x":
            a.decompress = True
        elif arg in ("--max-iter", "--tol", "--zstd", "--margin"):
            need(1, arg)
            i += 1
            conv = int if arg in ("--max-iter", "--zstd") else float
            setattr(a, {"--max-iter": "max_iter", "--tol": "tol", "--zstd": "zstd", "--margin": "margin"}[arg],
                    conv(argv[i]))
        elif arg == "--step-quant":
            a.step_quant = True
        elif arg == "--base":
            a.base_gpu = True
        else:
            raise UsageError("Unknown argument")
        i += 1
    if a.decompress:
        if not (a.compressed and (a.base or a.base_gpu) and a.decompressed):
            raise UsageError("Decompression needs the compressed file (-z), the base reconstruction (-e or --base) "
                             "and -o")
        return a
    if a.base_gpu and a.base:
        raise UsageError("--base replaces -e")
    if a.base_gpu and not a.compressed:
        raise UsageError("--base needs -z (its stream goes to <-z>.base)")
    if not a.original or not (a.base or a.base_gpu):
        raise UsageError("Input files of original data (-i) and base decompressed data (-e) are mandatory")
    if not a.shape:
        raise UsageError("Missing size of the data (-1, -2 or -3)")
    if a.spatial is None or a.freq is None:
        raise UsageError("Both error bounds (-M and -F) are required")
    if not 0.0 <= a.margin < 1.0:
        raise UsageError("--margin must lie in [0, 1)")
    return a


def read_raw(path: str, n: int, dtype) -> np.ndarray:
    """readRawArrayBinary (fileIO.cpp:35-76): the file must hold exactly n
    elements of dtype."""
    size = os.path.getsize(path)
    want = n * np.dtype(dtype).itemsize
    if size != want:
        raise UsageError(f"File size does not match array size: {path} has {size} bytes, expected {want}")
    return np.fromfile(path, dtype=dtype, count=n)


def bounds(x, a: Args, torch, dev):
    """Absolute bounds from -M / -F (REL: relative to the value range and to
    max |F(x)|)."""
    import ffcz_amd

    eps = a.spatial if a.spatial_abs else a.spatial * float(x.max() - x.min())
    if a.freq_abs:
        delta = a.freq
    else:
        X = ffcz_amd.fft_r2c(x)
        delta = a.freq * float(X.abs().max())
        del X
    if not eps > 0 or not delta > 0:
        raise UsageError("The error bounds are zero (a constant field with a relative bound)")
    return eps, delta


def compress(a: Args, out=print) -> dict:
    import torch

    import ffcz_amd
    from ffcz_amd import ProjectionSolver, QUANT_STEP, container

    ffcz_amd.load()
    dev = torch.device("cuda:0")
    n = int(np.prod(a.shape))
    x = torch.from_numpy(read_raw(a.original, n, a.dtype).reshape(a.shape)).to(dev)
    eps, delta = bounds(x, a, torch, dev)
    if a.base_gpu:
        tb = time.perf_counter()
        base, xb = container.base_compress(x, eps, zstd_level=a.zstd)
        torch.cuda.synchronize()
        out(f"Base stream: {len(base)} bytes ({1e3 * (time.perf_counter() - tb):.1f} ms)")
        with open(a.compressed + ".base", "wb") as f:
            f.write(base)
    else:
        xb = torch.from_numpy(read_raw(a.base, n, a.dtype).reshape(a.shape)).to(dev)
    e0 = xb - x  # projection_example.cu:173-176
    t0 = time.perf_counter()
    with ProjectionSolver(*a.shape, dtype=a.dtype) as s:
        s.initialize_all_abs(x, eps, delta)
        if a.margin:
            s.set_margins(a.margin, a.margin)
        if a.step_quant:
            s.set_quantizer(QUANT_STEP, delta / 8, eps / 8)
        s.solve_all_abs(e0, a.max_iter, a.tol)
        torch.cuda.synchronize()
        t_solve = time.perf_counter() - t0
        cert = s.certify(True)
        t1 = time.perf_counter()
        buf = container.pack(s, zstd_level=a.zstd)
        t_pack = time.perf_counter() - t1
        info = container.info(buf)
        iters, conv = s.iteration_count, s.status()[2]
    if a.compressed:
        with open(a.compressed, "wb") as f:
            f.write(buf)
    out(f"# Active frequency edits: {info['n_active_freq']}")
    out(f"# Active spatial edits: {info['n_active_spatial']}")
    out(f"Iterations: {iters} ({'converged' if conv else 'not converged'}); eps = {eps:.6g}, delta = {delta:.6g}")
    out(f"Certified (FP64, quantised edits): {'yes' if cert.ok else 'no'} "
        f"({cert.spatial_violations} spatial, {cert.freq_violations} frequency violations)")
    out(f"Compressed edits: {len(buf)} bytes ({n * np.dtype(a.dtype).itemsize / len(buf):.2f}x beside the base)")
    out(f"Compression time: {1e3 * (t_solve + t_pack):.1f} ms (solve {1e3 * t_solve:.1f} ms, pack {1e3 * t_pack:.1f} ms)")
    result = {"bytes": len(buf), "certified": cert.ok, "iterations": iters, "eps": eps, "delta": delta,
              "n_active_freq": info["n_active_freq"], "n_active_spatial": info["n_active_spatial"]}
    if a.decompressed:
        t2 = time.perf_counter()
        xh = container.decode(buf, xb)
        torch.cuda.synchronize()
        out(f"Decompression time: {1e3 * (time.perf_counter() - t2):.1f} ms")
        xh.cpu().numpy().tofile(a.decompressed)
    return result


def decompress(a: Args, out=print) -> None:
    import torch

    import ffcz_amd
    from ffcz_amd import container

    ffcz_amd.load()
    with open(a.compressed, "rb") as f:
        buf = f.read()
    info = container.info(buf)
    n = int(np.prod(info["shape"]))
    if a.base_gpu:
        with open(a.compressed + ".base", "rb") as f:
            xb = container.base_decompress(f.read(), torch.empty(info["shape"], dtype=getattr(
                torch, np.dtype(info["dtype"]).name), device="cuda:0"))
    else:
        xb = torch.from_numpy(read_raw(a.base, n, info["dtype"])).to(torch.device("cuda:0"))
    t0 = time.perf_counter()
    xh = container.decode(buf, xb)
    torch.cuda.synchronize()
    out(f"Decompression time: {1e3 * (time.perf_counter() - t0):.1f} ms")
    xh.cpu().numpy().tofile(a.decompressed)


def main(argv=None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    try:
        a = parse(argv)
        if a.decompress:
            decompress(a)
        else:
            compress(a)
    except UsageError as e:
        print(e, file=sys.stderr)
        print(USAGE, file=sys.stderr)
        return 2
    return 0


if __name__ == "__main__":
    sys.exit(main())
