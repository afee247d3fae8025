"""The C ABI boundary: the library loads, exports every entry point declared
in include/ffcz.h, and the C++ drop-in headers compile against it.  The GPU
test runs the drop-in program (the reference driver's workflow)."""
import re
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
LIB = ROOT / "ffcz_amd" / "libffcz.so"


def declared():
    text = (ROOT / "include" / "ffcz.h").read_text()
    return set(re.findall(r"\b(ffcz_\w+)\s*\(", text))


def test_library_exports_every_declared_symbol():
    import ffcz_amd

    lib = ffcz_amd.load()
    out = subprocess.run(["nm", "-D", "--defined-only", str(LIB)], capture_output=True, text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    missing = declared() - exported
    assert not missing, missing
    for name in declared():
        assert hasattr(lib, name)
    assert lib.ffcz_abi_version() == 1


def test_python_binding_covers_the_header():
    from ffcz_amd._lib import SIGNATURES

    assert declared() == set(SIGNATURES)


def test_null_handle_is_an_error_not_a_crash():
    import ffcz_amd
    from ffcz_amd._lib import FFCZ_ERR_INVALID

    lib = ffcz_amd.load()
    assert lib.ffcz_extract(None) == FFCZ_ERR_INVALID
    assert b"null solver handle" in lib.ffcz_last_error()


def test_plan_cache_path_is_settable_without_a_device(tmp_path):
    """The autotuner's plan file path (host-side state, no device call)."""
    import ffcz_amd

    old = ffcz_amd.plan_cache()
    try:
        ffcz_amd.set_plan_cache(tmp_path / "plans.txt")
        assert ffcz_amd.plan_cache() == str(tmp_path / "plans.txt")
        ffcz_amd.set_plan_cache(None)
        assert ffcz_amd.plan_cache() == ""
    finally:
        ffcz_amd.set_plan_cache(old)


def _compile_dropin(out: Path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    subprocess.run([gxx, "-std=c++17", "-O2", f"-I{ROOT / 'include'}", str(ROOT / "tests" / "cpp" / "dropin_check.cpp"),
                    f"-L{ROOT / 'ffcz_amd'}", "-lffcz", f"-Wl,-rpath,{ROOT / 'ffcz_amd'}", "-o", str(out)],
                   check=True, capture_output=True, text=True)


def test_cpp_dropin_headers_compile(tmp_path):
    _compile_dropin(tmp_path / "dropin_check")
    assert (tmp_path / "dropin_check").exists()


@pytest.mark.gpu
def test_cpp_dropin_runs(gpu, tmp_path):
    exe = tmp_path / "dropin_check"
    _compile_dropin(exe)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = [l for l in r.stdout.splitlines() if l.startswith("threw=")]
    assert len(lines) == 2
    for l in lines:
        kv = dict(p.split("=") for p in l.split())
        assert kv["threw"] == "1" and kv["roundtrip"] == "1"
        assert float(kv["max_abs_e"]) <= 1e-3 + 1e-7
