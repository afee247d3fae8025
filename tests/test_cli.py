"""The fftz-compatible command line: argument parsing with the reference's
flags and dimension conventions (projection_example.cu:30-154) and the raw
reader's size check (fileIO.cpp:35-76), on the CPU."""
import numpy as np
import pytest

from ffcz_amd import cli


def test_reference_flags_and_dimension_conventions():
    a = cli.parse("-i o.f32 -e b.f32 -z c.ffcz -o d.f32 -3 4 5 6 -f -M ABS 1e-3 -F REL 2e-4".split())
    assert (a.original, a.base, a.compressed, a.decompressed) == ("o.f32", "b.f32", "c.ffcz", "d.f32")
    assert a.shape == (4, 5, 6) and a.dtype == np.float32
    assert a.spatial_abs and a.spatial == 1e-3 and not a.freq_abs and a.freq == 2e-4
    # -1 n is (1, 1, n) and -2 nx ny is (1, nx, ny), as the reference sets Nx, Ny, Nz
    assert cli.parse("-i o -e b -1 7 -d -M REL 1e-2 -F ABS 3".split()).shape == (1, 1, 7)
    a2 = cli.parse("-i o -e b -2 8 9 -d -M REL 1e-2 -F ABS 3".split())
    assert a2.shape == (1, 8, 9) and a2.dtype == np.float64 and not a2.spatial_abs
    a3 = cli.parse("-x -z c -e b -o d".split())
    assert a3.decompress


@pytest.mark.parametrize("argv, msg", [
    ("-e b -3 2 2 2 -M ABS 1 -F ABS 1", "mandatory"),
    ("-i o -e b -M ABS 1 -F ABS 1", "Missing size"),
    ("-i o -e b -3 2 2 2 -F ABS 1", "Both error bounds"),
    ("-i o -e b -3 2 2", "Missing size of the data"),
    ("-i o -e b -3 2 2 2 -M XYZ 1 -F ABS 1", "ABS or REL"),
    ("-i o -e b -3 2 2 2 -M ABS 0 -F ABS 1", "positive"),
    ("-i o -e b -3 2 0 2 -M ABS 1 -F ABS 1", "positive"),
    ("-i o -e b -3 2 2 2 -M ABS 1 -F ABS 1 -q", "Unknown argument"),
    ("-i", "Missing input file path"),
    ("-x -z c -o d", "Decompression needs"),
    ("-i o -e b -3 2 2 2 -M ABS 1 -F ABS 1 --margin 1.5", "margin"),
    ("-i o -e b --base -z c -3 2 2 2 -M ABS 1 -F ABS 1", "replaces -e"),
    ("-i o --base -3 2 2 2 -M ABS 1 -F ABS 1", "needs -z"),
])
def test_bad_arguments_are_reported(argv, msg, capsys):
    with pytest.raises(cli.UsageError, match=msg):
        cli.parse(argv.split())
    assert cli.main(argv.split()) == 2
    assert "Usage:" in capsys.readouterr().err


def test_raw_reader_checks_the_file_size(tmp_path):
    p = tmp_path / "a.f32"
    np.arange(10, dtype=np.float32).tofile(p)
    assert np.array_equal(cli.read_raw(str(p), 10, np.float32), np.arange(10, dtype=np.float32))
    with pytest.raises(cli.UsageError, match="File size does not match"):
        cli.read_raw(str(p), 11, np.float32)
    with pytest.raises(cli.UsageError, match="File size does not match"):
        cli.read_raw(str(p), 10, np.float64)
