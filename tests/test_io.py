"""The reference CLI's file formats (turbotops_amd.io, src/turbotops.cpp:126-230):
readers against hand-made files, the diagram writer against the compiled
reference CLI's own output on the same input (oracle/_ref)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

CLI = os.path.join(ROOT, "oracle", "_ref", "turbotops_nosimd")


def test_read_text_points_separators_comments_bom(tmp_path):
    from turbotops_amd.io import read_text_points
    p = tmp_path / "pts.csv"
    p.write_bytes(b"\xef\xbb\xbf# header\n0,1,2\n\n3\t4;5  # trailing\n 6 7 8\n")
    x = read_text_points(str(p))
    assert x.dtype == np.float32 and np.array_equal(x, np.arange(9, dtype=np.float32).reshape(3, 3))
    # lines end at '\n' only; \v, \f and \r separate tokens, as getline and >> do
    p.write_bytes(b"0\x0b1\r\n2\x0c3\n")
    assert read_text_points(str(p)).tolist() == [[0, 1], [2, 3]]
    p.write_text("0 1\n2 3 4\n")
    with pytest.raises(ValueError, match="inconsistent row length at point 1"):
        read_text_points(str(p))
    p.write_text("0 x\n")
    with pytest.raises(ValueError, match="bad coordinate 'x'"):
        read_text_points(str(p))
    with pytest.raises(ValueError, match="cannot open"):
        read_text_points(str(tmp_path / "missing.txt"))


def test_read_bin_points(tmp_path):
    from turbotops_amd.io import read_bin_points
    x = np.random.default_rng(1).normal(size=(7, 3)).astype(np.float32)
    p = tmp_path / "pts.bin"
    x.tofile(p)
    assert np.array_equal(read_bin_points(str(p), 3), x)
    with pytest.raises(ValueError, match="multiple of 4"):
        read_bin_points(str(p), 4)
    with pytest.raises(ValueError, match="--dim is required"):
        read_bin_points(str(p), 0)


def test_read_edges(tmp_path):
    from turbotops_amd.io import read_edges
    p = tmp_path / "g.txt"
    p.write_text("# i j w\n0 1 2.5\n1,0,1.5\n2\t3\t0.5\n3 3 9\n1 3 1\n")
    g = read_edges(str(p))
    assert g.n == 4 and g.valid()
    assert g.row_ptr.tolist() == [0, 1, 3, 4, 6]
    assert g.col_idx.tolist() == [1, 0, 3, 3, 1, 2]
    assert g.weights.tolist() == [1.5, 1.5, 1.0, 0.5, 1.0, 0.5]
    # strtoull: a sign is accepted and '-1' wraps to 2^64 - 1
    p.write_text("+0 1 1\n")
    assert read_edges(str(p)).n == 2
    p.write_text("0 -1 1\n")
    with pytest.raises(ValueError, match="vertex id too large"):
        read_edges(str(p))
    p.write_text("0 18446744073709551616 1\n")
    with pytest.raises(ValueError, match="bad vertex"):
        read_edges(str(p))
    p.write_text("0 1\n")
    with pytest.raises(ValueError, match="expected 'i j w'"):
        read_edges(str(p))
    p.write_text("0 1 -1\n")
    with pytest.raises(ValueError, match="bad weight"):
        read_edges(str(p))


@pytest.mark.skipif(not os.path.exists(CLI), reason="oracle/_ref CLI not built")
def test_diagram_writer_matches_reference_cli(tmp_path, reference):
    from turbotops_amd.io import diagram_lines, write_diagram
    x = np.random.default_rng(3).normal(size=(60, 3)).astype(np.float32)
    p = tmp_path / "pts.txt"
    np.savetxt(p, x, fmt="%.9g")
    cli = subprocess.run([CLI, "run", str(p), "--maxdim", "1"], capture_output=True, text=True, timeout=60)
    assert cli.returncode == 0, cli.stderr
    want = [ln for ln in cli.stdout.splitlines() if ln and not ln.startswith("#")]
    from turbotops_amd.io import read_text_points
    xr = read_text_points(str(p))
    dims, b, e = reference.persistence(xr, maxdim=1)[0]
    dgms = [np.stack([b[dims == d], e[dims == d]], 1) for d in range(2)]
    assert diagram_lines(dgms) == want
    out = tmp_path / "d.txt"
    write_diagram(str(out), dgms)
    assert out.read_text().splitlines() == want


def test_parse_f32_rounds_once_like_strtof():
    """The text readers convert each token as the C library's strtof does
    (src/turbotops.cpp:118-123), including the decimals whose nearest
    float64 lies exactly halfway between two float32 values."""
    import ctypes
    from fractions import Fraction

    from turbotops_amd.io import _parse_f32
    libc = ctypes.CDLL(None)
    libc.strtof.restype = ctypes.c_float
    libc.strtof.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p)]

    def strtof(s):
        return np.float32(libc.strtof(s.encode(), None))

    # 1 + 2^-24 is halfway between 1 and 1 + 2^-23; a decimal just above or
    # below it rounds to that float64 midpoint, but strtof resolves it
    mid = Fraction(1) + Fraction(1, 2 ** 24)
    def dec(q):   # exact decimal text of a dyadic rational
        k = 0
        while (q * 10 ** k).denominator != 1:
            k += 1
        i = q * 10 ** k
        s = str(i.numerator).rjust(k + 1, "0")
        return s[:-k] + "." + s[-k:]
    toks = [dec(mid + Fraction(1, 2 ** 80)), dec(mid - Fraction(1, 2 ** 80)), dec(mid),
            "-" + dec(mid + Fraction(1, 2 ** 80)), "0.1", "1e-45", "7e-46", "3.4028235e38", "3.4028236e38",
            "1e39", "-0", "inf", "-Infinity", "0x1.8p3", "1.17549435e-38", "2.5e-42"]
    rng = np.random.default_rng(5)
    toks += [f"{v:.{p}g}" for v, p in zip(rng.normal(size=200) * 10.0 ** rng.integers(-40, 38, 200),
                                             rng.integers(6, 20, 200))]
    for t in toks:
        got, want = _parse_f32(t, "coordinate"), strtof(t)
        assert got.view(np.uint32) == want.view(np.uint32), t
    assert _parse_f32(dec(mid + Fraction(1, 2 ** 80)), "x") == np.float32(1 + 2 ** -23)
    with pytest.raises(ValueError, match="bad coordinate '1_0'"):
        _parse_f32("1_0", "coordinate")
