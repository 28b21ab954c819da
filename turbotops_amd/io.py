"""Point, graph and diagram file formats of the reference's command-line tool
(src/turbotops.cpp:126-230), so that files written for one read the same in
the other. Errors raise ValueError with the reference's message text.

* text points: one point per line; ',', tab and ';' separate coordinates as
  spaces do; '#' starts a comment; blank lines are skipped; a UTF-8 BOM is
  ignored; every row must have the first row's length.
* .bin points: raw little-endian float32, row-major, dimension given.
* edge lists: 'i j w' per line (',' or tab also separate); self loops are
  dropped; a repeated edge keeps its smallest weight.
* diagrams: 'dim birth death' per bar with %.9g numbers and 'inf', in the
  canonical order (dimension, birth, death).
"""
import math

import numpy as np

from .frontend import CsrGraph


def _parse_f32(tok, what):
    try:
        v = np.float32(float(tok))
    except ValueError:
        raise ValueError(f"bad {what} '{tok}'") from None
    return v


def _parse_u64(tok, what):
    if not tok.isdigit():
        raise ValueError(f"bad {what} '{tok}'")
    return int(tok)


def read_text_points(path):
    """read_text_points (src/turbotops.cpp:126-155) -> (n, dim) float32."""
    try:
        raw = open(path, "rb").read()
    except OSError:
        raise ValueError(f"cannot open {path}") from None
    if raw.startswith(b"\xef\xbb\xbf"):
        raw = raw[3:]
    rows, dim = [], 0
    for line in raw.decode("utf-8", "replace").splitlines():
        line = line.split("#", 1)[0]
        for ch in ",\t;":
            line = line.replace(ch, " ")
        toks = line.split()
        if not toks:
            continue
        vals = [_parse_f32(t, "coordinate") for t in toks]
        if not dim:
            dim = len(vals)
        elif len(vals) != dim:
            raise ValueError(f"{path}: inconsistent row length at point {len(rows)}")
        rows.append(vals)
    return np.array(rows, np.float32).reshape(len(rows), dim)


def read_bin_points(path, dim):
    """read_bin_points (src/turbotops.cpp:157-169) -> (n, dim) float32."""
    if not dim:
        raise ValueError("--dim is required for .bin input")
    try:
        raw = open(path, "rb").read()
    except OSError:
        raise ValueError(f"cannot open {path}") from None
    if len(raw) % (dim * 4):
        raise ValueError(f"{path}: size is not a multiple of 4 * dim")
    return np.frombuffer(raw, dtype="<f4").astype(np.float32).reshape(-1, dim)


def read_edges(path):
    """read_edges (src/turbotops.cpp:171-203) -> CsrGraph."""
    try:
        lines = open(path).read().splitlines()
    except OSError:
        raise ValueError(f"cannot open {path}") from None
    rows = {}
    for lineno, line in enumerate(lines, 1):
        line = line.split("#", 1)[0].replace(",", " ").replace("\t", " ")
        toks = line.split()
        if not toks:
            continue
        if len(toks) < 3:
            raise ValueError(f"{path}:{lineno}: expected 'i j w'")
        i, j = _parse_u64(toks[0], "vertex"), _parse_u64(toks[1], "vertex")
        d = _parse_f32(toks[2], "weight")
        if i == j:
            continue
        if not (d >= 0) or not math.isfinite(d):
            raise ValueError(f"{path}:{lineno}: bad weight")
        if max(i, j) >= 0xFFFFFFFF:
            raise ValueError("vertex id too large")
        for a, b in ((i, j), (j, i)):
            r = rows.setdefault(a, {})
            if b not in r or d < r[b]:
                r[b] = d
    n = (max(rows) + 1) if rows else 0
    rp = np.zeros(n + 1, np.uint64)
    col, w = [], []
    for v in range(n):
        r = rows.get(v, {})
        for u in sorted(r):
            col.append(u)
            w.append(r[u])
        rp[v + 1] = len(col)
    return CsrGraph(n, rp, np.array(col, np.uint32), np.array(w, np.float32))


def diagram_lines(dgms):
    """write_diagram (src/turbotops.cpp:205-230): the canonical text lines."""
    bars = []
    for d, a in enumerate(dgms):
        a = np.asarray(a, np.float32).reshape(-1, 2)
        bars.extend((d, float(b), float(e)) for b, e in a)
    bars.sort(key=lambda t: (t[0], t[1], t[2]))
    return [f"{d} {b:.9g} inf" if math.isinf(e) else f"{d} {b:.9g} {e:.9g}" for d, b, e in bars]


def write_diagram(f, dgms):
    """Writes the canonical diagram text to a path or an open text file."""
    text = "".join(line + "\n" for line in diagram_lines(dgms))
    if hasattr(f, "write"):
        f.write(text)
    else:
        with open(f, "w") as fh:
            fh.write(text)
