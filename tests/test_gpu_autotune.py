"""FFT plan autotuner: tile variants of every pass kernel, and the radix order
of the iteration's vector-ALU y and x passes.

The tile variants change only how pencils are grouped into workgroups (64- or
128-byte column tiles for y/x, 128 or 256 threads for z); each pencil's
stages, twiddles and operation order are the same, so every tile variant gives
bit-identical iterates, edits and certification counts.  That is the parity
check here, over all 2^7 combinations on small fields and on the tuner's own
choice.  The radix-order variant (2) runs a length's radices reversed: a
different rounding, checked against the default within the FP32 / FP64
tolerances."""
import itertools
import json
import os
import subprocess
import sys
import textwrap
from pathlib import Path

import numpy as np
import pytest

from ffcz_amd import FFT_BF16X3, FFT_FP16X3, MODE_ABS, ProjectionSolver
from ffcz_amd._lib import FFCzError
from tests._cases import binding_case

pytestmark = pytest.mark.gpu

NAMES = ProjectionSolver.TILE_VAR_NAMES


def _run(shape, dtype, case, vars_=None, tune=False, fft=None):
    x, e0, eps, delta = case
    with ProjectionSolver(*shape, dtype=dtype) as s:
        s.initialize_all_abs(x, eps, delta)
        if fft is not None:
            s.set_fft_precision(fft)
        chosen = None
        if tune:
            chosen = s.autotune(reps=2, use_cache=False, timings=True)
        if vars_ is not None:
            s.set_tile_vars(vars_)
            assert s.tile_vars() == vars_
        s.set_margins(0.125, 0.125)
        s.solve_loop(MODE_ABS, e0, 40, 1e-7)
        c = s.certify(False)
        return s.status(), s.get_results(), s.get_spatial_edits_full(), s.get_freq_edits_full(), c, chosen


@pytest.mark.parametrize("shape,dtype", [((32, 64, 128), np.float32), ((30, 36, 50), np.float64),
                                         ((16, 60, 120), np.float32)])
def test_every_tile_variant_gives_identical_results(gpu, shape, dtype):
    case = binding_case(shape, seed=3, dtype=dtype, spread=1.4, dfrac=0.9)
    ref = _run(shape, dtype, case)
    combos = list(itertools.product((0, 1), repeat=7))
    # all combinations on one shape, a spread of them on the others
    if shape != (32, 64, 128):
        combos = combos[::9] + [(1,) * 7]
    for combo in combos:
        got = _run(shape, dtype, case, dict(zip(NAMES, combo)))
        assert got[0] == ref[0], combo
        for a, b in zip(got[1:4], ref[1:4]):
            assert np.array_equal(a, b), combo
        assert got[4] == ref[4], combo


@pytest.mark.parametrize("fft", [FFT_BF16X3, FFT_FP16X3])
@pytest.mark.parametrize("shape", [(32, 64, 512), (16, 128, 256), (64, 32, 96)])
def test_mfma_iteration_under_every_tile_variant(gpu, shape, fft):
    """The MFMA passes run their full-width tiles whatever the tile variants
    say; the variants reach the vector-ALU kernels of the MFMA iteration (the
    fused z pass) and the FP64 certification, which regroup pencils without
    changing any operation: the results stay bit-identical.  The tuner times
    only those variants under an MFMA precision.  The shapes cover the 16 x 16
    row transform of the z passes (nz = 512), radix-16 stages with vector-ALU
    remainders and the strided passes."""
    dtype = np.float32
    case = binding_case(shape, seed=6, dtype=dtype, spread=1.4, dfrac=0.9)
    ref = _run(shape, dtype, case, fft=fft)
    for combo in [(1,) * 7, (1, 0, 1, 0, 0, 0, 0), (0, 1, 0, 1, 0, 0, 0)]:
        got = _run(shape, dtype, case, dict(zip(NAMES, combo)), fft=fft)
        assert got[0] == ref[0], combo
        for a, b in zip(got[1:4], ref[1:4]):
            assert np.array_equal(a, b), combo
    tuned = _run(shape, dtype, case, tune=True, fft=fft)
    ms = tuned[5]["ms"]
    # z forward, y and x are not timed: their slots repeat the default time
    assert ms[1] == ms[2] == ms[3] == ms[0]
    assert tuned[5]["z_fwd"] == tuned[5]["y"] == tuned[5]["x"] == 0


def test_autotune_choice_cache_and_errors(gpu):
    shape, dtype = (32, 64, 128), np.float32
    case = binding_case(shape, seed=4, dtype=dtype)
    ref = _run(shape, dtype, case)
    tuned = _run(shape, dtype, case, tune=True)
    chosen = tuned[5]
    assert set(chosen) == set(NAMES) | {"ms"} and len(chosen["ms"]) == 9
    assert all(t > 0 for t in chosen["ms"][:8]) and chosen["ms"][5] <= chosen["ms"][0]
    for a, b in zip(tuned[1:4], ref[1:4]):
        assert np.array_equal(a, b)
    with ProjectionSolver(*shape, dtype=dtype) as s:
        s.initialize_all_abs(case[0], case[2], case[3])
        again = s.autotune(reps=1, use_cache=True)  # from the cache: no timing
        assert again == {k: chosen[k] for k in NAMES}
        assert s.tune_source() == "process"
        with pytest.raises(FFCzError):
            s.set_tile_vars({"z_fwd": 2})
    with ProjectionSolver(*shape, dtype=dtype) as s:
        with pytest.raises(FFCzError):
            s.autotune()  # before initialize


def test_pass_timing_inside_solves(gpu):
    """Events between the launches of real solves: one sample per iteration
    that did work, positive per-pass means, results unchanged."""
    shape, dtype = (32, 64, 128), np.float32
    case = binding_case(shape, seed=5, dtype=dtype)
    ref = _run(shape, dtype, case)
    x, e0, eps, delta = case
    with ProjectionSolver(*shape, dtype=dtype) as s:
        s.initialize_all_abs(x, eps, delta)
        s.set_margins(0.125, 0.125)
        s.set_pass_timing(True)
        total = 0
        for _ in range(2):
            s.reset_edit_accumulators()
            s.solve_loop(MODE_ABS, e0, 40, 1e-7)
            total += s.status()[0]
        ms, n = s.pass_timing()
        # the fused form: the forward z pass of the initial error opens each
        # solve, and the first launch after each host check (that pass, or a
        # y forward pass at the start of a later chunk) is left out
        assert n == total and ms[0] == 0 and all(t > 0 for t in ms[1:])
        cnt = s.pass_counts()
        assert cnt[0] == 0 and 0 < cnt[1] <= total and cnt[2:5] == [total] * 3
        assert np.array_equal(s.get_results(), ref[1])
        s.set_pass_timing(False)
        assert s.pass_timing()[1] == 0


_PLAN_SCRIPT = textwrap.dedent("""
    import json, sys
    import numpy as np
    sys.path.insert(0, sys.argv[1])
    from ffcz_amd import ProjectionSolver
    from tests._cases import binding_case
    shape = tuple(int(v) for v in sys.argv[2].split(","))
    x, e0, eps, delta = binding_case(shape, seed=4, dtype=np.float32)
    with ProjectionSolver(*shape, dtype=np.float32) as s:
        s.initialize_all_abs(x, eps, delta)
        v = s.autotune(reps=1, use_cache=True)
        print(json.dumps({"vars": v, "source": s.tune_source()}))
""")


def _plan_process(path, shape):
    env = dict(os.environ, FFCZ_PLAN_CACHE=str(path))
    root = str(Path(__file__).resolve().parents[1])
    out = subprocess.run([sys.executable, "-c", _PLAN_SCRIPT, root, ",".join(map(str, shape))], env=env,
                         capture_output=True, text=True, timeout=240)
    assert out.returncode == 0, out.stderr[-2000:]
    return json.loads(out.stdout.strip().splitlines()[-1])


def test_plan_file_reused_by_a_second_process(gpu, tmp_path):
    """A plan tuned in one process is written to the plan file; the next
    process reads it from there and gets the same variants without timing.
    Another shape in the same file is tuned and kept beside it."""
    path = tmp_path / "cache" / "plans.txt"
    first = _plan_process(path, (32, 64, 128))
    assert first["source"] == "timed"
    second = _plan_process(path, (32, 64, 128))
    assert second == {"vars": first["vars"], "source": "disk"}
    third = _plan_process(path, (16, 60, 120))
    assert third["source"] == "timed"
    plans = [line.split() for line in path.read_text().splitlines() if line and not line.startswith("#")]
    assert len(plans) == 2
    assert sorted(tuple(p[3:6]) for p in plans) == [("16", "60", "120"), ("32", "64", "128")]
    again = _plan_process(path, (16, 60, 120))
    assert again == {"vars": third["vars"], "source": "disk"}


def _rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("shape,dtype,tol", [((60, 120, 96), np.float32, 2e-5), ((150, 60, 48), np.float64, 1e-11)])
def test_radix_order_variant_agrees_with_the_default(gpu, shape, dtype, tol):
    """Variant 2 of the y and x passes runs the length's radices reversed
    (60 = 4*5*3 as 3*5*4, 120 = 8*5*3 as 3*5*8, 150 = 2*5*5*3 as 3*5*5*2): the
    same transform with another rounding, so the loop takes as many
    iterations, its edits agree within the precision's tolerance, the
    certification passes, and the bits differ (the reversed kernels ran)."""
    case = binding_case(shape, seed=8, dtype=dtype, spread=1.4, dfrac=0.9)
    ref = _run(shape, dtype, case)
    got = _run(shape, dtype, case, dict(zip(NAMES, (0, 0, 2, 2, 0, 0, 0))))
    assert got[0][0] == ref[0][0] and got[0][2] == ref[0][2]
    for a, b in zip(got[1:4], ref[1:4]):
        assert _rel(a, b) < tol
    assert any(not np.array_equal(a, b) for a, b in zip(got[1:4], ref[1:4]))
    assert got[4].ok and ref[4].ok
    with ProjectionSolver(*shape, dtype=dtype) as s:
        with pytest.raises(FFCzError):
            s.set_tile_vars(dict(zip(NAMES, (2, 0, 0, 0, 0, 0, 0))))
        with pytest.raises(FFCzError):
            s.set_tile_vars(dict(zip(NAMES, (0, 0, 0, 0, 0, 2, 0))))
