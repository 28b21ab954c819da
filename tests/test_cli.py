"""The command-line `run` (turbotops_amd/turbotops, host/cli.cpp) against the
reference's own tool (src/turbotops.cpp:784-933), compiled from its sources
into oracle/_ref/turbotops_nosimd: the same exit status and message for
every argument and input error, and the same diagram bytes on stdout for
text, CSV, .bin and edge-list inputs under each metric, graph and mode.

The error cases stop before any GPU work and run on the CPU; the diagram
cases run the front end on the GPU."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

OURS = os.path.join(ROOT, "turbotops_amd", "turbotops")
REF = os.path.join(ROOT, "oracle", "_ref", "turbotops_nosimd")


def _cli(exe, args, cwd, timeout=120):
    return subprocess.run([exe] + args, capture_output=True, text=True, timeout=timeout, cwd=cwd)


@pytest.fixture(scope="module")
def ours():
    if not os.path.exists(OURS):
        pytest.skip("turbotops_amd/turbotops not built")
    return OURS


@pytest.fixture(scope="module")
def inputs(tmp_path_factory):
    d = tmp_path_factory.mktemp("cli")
    (d / "two.txt").write_text("0 0\n1 0\n")
    (d / "empty.txt").write_text("")
    (d / "ragged.txt").write_text("0 1\n2 3 4\n")
    (d / "bad.txt").write_text("0 x\n")
    (d / "short_edge.txt").write_text("0 1 2\n1;2 3\n")
    (d / "neg_edge.txt").write_text("0 1 -1\n")
    (d / "big_vertex.txt").write_text("0 4294967295 1\n")
    (d / "pts.bin").write_bytes(np.zeros(5, np.float32).tobytes())
    return d


# (arguments, first line of stderr); every case exits with status 2
ERRORS = [
    ([], "usage:"),
    (["run"], "turbotops: run: missing input file"),
    (["bogus"], "turbotops: unknown command 'bogus' (try --help)"),
    (["run", "empty.txt"], "turbotops: empty.txt: no points (empty file, directory, or wrong format?)"),
    (["run", "ragged.txt"], "turbotops: ragged.txt: inconsistent row length at point 1"),
    (["run", "bad.txt"], "turbotops: bad coordinate 'x'"),
    (["run", "missing.txt"], "turbotops: cannot open missing.txt"),
    (["run", "pts.bin"], "turbotops: --dim is required for .bin input"),
    (["run", "pts.bin", "--dim", "2"], "turbotops: pts.bin: size is not a multiple of 4 * dim"),
    (["run", "a", "b"], "turbotops: unexpected argument 'b'"),
    (["run", "a", "--frob"], "turbotops: unknown option '--frob'"),
    (["run", "a", "--maxdim"], "turbotops: missing value for --maxdim"),
    (["run", "a", "--maxdim", "-1"], "turbotops: bad maxdim '-1' (out of range)"),
    (["run", "a", "-k", "x"], "turbotops: bad k 'x'"),
    (["run", "a", "--mode", "foo"], "turbotops: mode must be auto, boundary or coboundary"),
    (["run", "a", "--mode", "boundary", "--coboundary"],
     "turbotops: --coboundary conflicts with an earlier mode/engine flag"),
    (["run", "a", "--engine", "zz"], "turbotops: engine must be auto, reducer or redzed"),
    (["run", "a", "--graph", "ring"], "turbotops: --graph must be full or knn"),
    (["run", "a", "--approximate"], "turbotops: --approximate requires --graph knn"),
    (["run", "a", "--verify-recall", "3"], "turbotops: --verify-recall requires --approximate"),
    (["run", "a", "--memory-budget", "12x"], "turbotops: bad byte count '12x'"),
    (["run", "a", "--memory-budget", "0"], "turbotops: bad byte count '0'"),
    (["run", "a", "--threshold", "1.5q"], "turbotops: bad threshold '1.5q'"),
    (["run", "a", "--output", "/nonexistent/x"], "turbotops: cannot write /nonexistent/x"),
    (["run", "short_edge.txt", "--edges"], "turbotops: short_edge.txt:2: expected 'i j w'"),
    (["run", "neg_edge.txt", "--edges"], "turbotops: neg_edge.txt:1: bad weight"),
    (["run", "big_vertex.txt", "--edges"], "turbotops: vertex id too large"),
    (["run", "empty.txt", "--edges"], "turbotops: empty.txt: no vertices (empty edge list?)"),
    # option checks inside persistence(), before any device work
    (["run", "two.txt", "--threshold", "-1"], "turbotops: persistence: threshold must be non-negative or +inf"),
    (["run", "two.txt", "--metric", "foo"],
     "turbotops: metric: unknown 'foo' (expected l1 | l2 | linf | lp:<p>, p >= 1)"),
    (["run", "two.txt", "--metric", "lp:0.5"], "turbotops: metric: p must be in [1, inf]"),
    (["run", "two.txt", "--graph", "knn", "-k", "0"], "turbotops: persistence: k must be >= 1"),
    (["run", "two.txt", "--maxdim", "40"], "turbotops: rips: maxdim must be <= 30"),
    (["run", "two.txt", "--maxdim", "40", "--boundary"], "turbotops: redzed: maxdim must be <= 30"),
]


@pytest.mark.parametrize("args,msg", ERRORS, ids=[" ".join(a) or "(none)" for a, _ in ERRORS])
def test_cli_errors(ours, inputs, args, msg):
    got = _cli(ours, args, inputs)
    assert got.returncode == 2 and got.stdout == ""
    assert got.stderr.splitlines()[0] == msg
    if os.path.exists(REF):   # the expectations are the reference tool's own output
        want = _cli(REF, args, inputs)
        assert (want.returncode, want.stderr.splitlines()[0]) == (2, msg)


def test_cli_help_and_version(ours):
    h = _cli(ours, ["--help"], ROOT)
    assert h.returncode == 0 and "turbotops run <input> [options]" in h.stdout
    v = _cli(ours, ["--version"], ROOT)
    assert v.returncode == 0 and v.stdout.startswith("turbotops ")
    g = _cli(ours, ["gen", "uniform", "10", "2", "1", "x.txt"], ROOT)
    assert g.returncode == 2 and "not part of this tool" in g.stderr


# ---- diagrams on the GPU ------------------------------------------------------------------

def _torus(n, seed):
    rng = np.random.default_rng(seed)
    t, p = rng.uniform(0, 2 * np.pi, (2, n))
    x = np.stack([(1 + 0.3 * np.cos(t)) * np.cos(p), (1 + 0.3 * np.cos(t)) * np.sin(p), 0.3 * np.sin(t)], 1)
    return (x + rng.normal(0, 0.01, x.shape)).astype(np.float32)


@pytest.fixture(scope="module")
def clouds(tmp_path_factory):
    if not os.path.exists(REF):
        pytest.skip("oracle/_ref/turbotops_nosimd not built")
    d = tmp_path_factory.mktemp("clouds")
    rng = np.random.default_rng(11)
    tor = _torus(300, 1)
    np.savetxt(d / "torus.txt", tor, fmt="%.9g")
    tor.tofile(d / "torus.bin")
    u8 = rng.uniform(0, 1, (400, 8)).astype(np.float32)
    np.savetxt(d / "u8.txt", u8, fmt="%.9g")
    g2 = rng.normal(size=(1000, 2)).astype(np.float32)
    np.savetxt(d / "g2.txt", g2, fmt="%.9g")
    # CSV with a BOM, comments, blank lines and mixed separators
    rows = [",".join(f"{v:.9g}" for v in r) if i % 3 else ";".join(f"{v:.9g}" for v in r) + "  # c"
            for i, r in enumerate(_torus(200, 4))]
    (d / "torus.csv").write_bytes(b"\xef\xbb\xbf# x,y,z\n" + ("\n\n".join(rows) + "\n").encode())
    # an edge list: a random geometric graph with repeated edges and self loops
    p = rng.uniform(0, 1, (250, 2))
    dd = np.sqrt(((p[:, None] - p[None]) ** 2).sum(-1))
    i, j = np.nonzero(np.triu(dd < 0.16, 1))
    lines = [f"{a} {b} {dd[a, b]:.9g}" for a, b in zip(i, j)]
    lines += [f"{b}\t{a}\t{dd[a, b] * 1.5:.9g}" for a, b in zip(i[::7], j[::7])]   # heavier repeats
    lines += [f"{a},{a},0.5" for a in range(0, 250, 50)]                           # self loops
    (d / "edges.txt").write_text("# i j w\n" + "\n".join(lines) + "\n")
    return d


DIAGRAMS = [
    ["torus.txt"],
    ["torus.txt", "--coboundary"],
    ["torus.txt", "--boundary"],
    ["torus.txt", "--metric", "l1"],
    ["torus.txt", "--metric", "linf", "--maxdim", "2", "--threshold", "0.45"],
    ["torus.bin", "--dim", "3", "--maxdim", "2", "--threshold", "0.5"],
    ["torus.csv"],
    ["u8.txt", "--graph", "knn", "-k", "8", "--maxdim", "2"],
    ["u8.txt", "--metric", "chebyshev", "--threshold", "0.55"],
    ["g2.txt", "--maxdim", "0"],
    ["g2.txt", "--maxdim", "0", "--threshold", "0.05"],
    ["edges.txt", "--edges", "--maxdim", "2"],
    ["edges.txt", "--edges", "--threshold", "0.1"],
]


@pytest.mark.gpu
@pytest.mark.parametrize("args", DIAGRAMS, ids=[" ".join(a) for a in DIAGRAMS])
def test_cli_diagram_bytes(gpu, ours, clouds, args):
    got = _cli(ours, ["run"] + args + ["--quiet"], clouds)
    want = _cli(REF, ["run"] + args + ["--quiet"], clouds)
    assert want.returncode == 0, want.stderr
    assert got.returncode == 0, got.stderr
    assert got.stdout.count("\n") > 0 and got.stdout == want.stdout


@pytest.mark.gpu
def test_cli_lp_within_tolerance(gpu, ours, clouds):
    """Lp distances are within 2 ulp of the reference's, so the diagram is
    compared bar for bar with a relative tolerance."""
    args = ["run", "torus.txt", "--metric", "lp:1.5", "--quiet"]
    got, want = _cli(ours, args, clouds), _cli(REF, args, clouds)
    assert got.returncode == 0 and want.returncode == 0

    def bars(text):
        rows = [ln.split() for ln in text.splitlines()]
        return {d: np.array(sorted((float(b), float(e)) for dd, b, e in rows if int(dd) == d)) for d in (0, 1)}
    g, w = bars(got.stdout), bars(want.stdout)
    for d in (0, 1):
        assert g[d].shape == w[d].shape
        assert np.allclose(g[d], w[d], rtol=1e-5, atol=0)


@pytest.mark.gpu
def test_cli_output_file_and_summary(gpu, ours, clouds, tmp_path):
    out = tmp_path / "d.txt"
    got = _cli(ours, ["run", "torus.txt", "--output", str(out)], clouds)
    assert got.returncode == 0 and got.stdout == ""
    want = _cli(REF, ["run", "torus.txt", "--quiet"], clouds)
    assert out.read_text() == want.stdout
    err = got.stderr.splitlines()
    assert err[0].startswith("turbotops: mode: auto: ")
    assert err[1].startswith("turbotops: method=")
    h1 = sum(1 for ln in want.stdout.splitlines() if ln.startswith("1 "))
    assert f"turbotops: H1 {h1} finite, 0 essential" in err
