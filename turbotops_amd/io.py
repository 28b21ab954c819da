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
import re
from fractions import Fraction

import numpy as np

from .frontend import CsrGraph

_U64 = re.compile(r"([+-]?)([0-9]+)")
_U64_MAX = (1 << 64) - 1
_SPACE = re.compile(r"[ \t\n\v\f\r]+")


def _parse_f32(tok, what):
    """The C library's strtof on a whole token: the decimal (or hexadecimal)
    value rounded once to float32. Going through float64 first rounds twice,
    which differs where the float64 lands exactly halfway between two
    float32 values; that case is settled on the exact decimal value."""
    try:
        if "_" in tok or not tok.isascii():
            raise ValueError
        low = tok.lower().lstrip("+-")
        d = float.fromhex(tok) if low.startswith("0x") else float(tok)
    except ValueError:
        raise ValueError(f"bad {what} '{tok}'") from None
    with np.errstate(over="ignore"):
        f = np.float32(d)
        g = np.nextafter(f, np.float32(math.copysign(math.inf, d - float(f)))) if np.isfinite(f) else f
    if np.isfinite(f) and float(f) != d:
        mid = (Fraction(float(f)) + Fraction(float(g))) / 2 if np.isfinite(g) else None
        if mid is not None and Fraction(d) == mid and not low.startswith("0x"):
            exact = Fraction(tok)
            if exact != mid:
                f = min(f, g) if exact < mid else max(f, g)
    return f


def _parse_u64(tok, what):
    """strtoull on a whole token: an optional sign and decimal digits; a
    minus sign wraps modulo 2^64 and values past 2^64 - 1 are errors."""
    m = _U64.fullmatch(tok)
    if m is None or int(m.group(2)) > _U64_MAX:
        raise ValueError(f"bad {what} '{tok}'")
    v = int(m.group(2))
    return (-v) % (_U64_MAX + 1) if m.group(1) == "-" else v


def _lines(path):
    """The file's lines, split at '\n' only, as getline does; bytes that are
    not UTF-8 are kept (they fail the numeric conversion, as in C)."""
    try:
        with open(path, "rb") as fh:
            raw = fh.read()
    except OSError:
        raise ValueError(f"cannot open {path}") from None
    return raw, raw.decode("utf-8", "surrogateescape").split("\n")


def _tokens(line, seps):
    """Cut at '#', map each character of seps to a space and split at the C
    library's white space (space, \\t, \\n, \\v, \\f, \\r)."""
    line = line.split("#", 1)[0]
    for ch in seps:
        line = line.replace(ch, " ")
    return [t for t in _SPACE.split(line) if t]


def read_text_points(path):
    """read_text_points (src/turbotops.cpp:126-155) -> (n, dim) float32."""
    raw, lines = _lines(path)
    if raw.startswith(b"\xef\xbb\xbf"):
        lines[0] = lines[0][1:]   # the BOM decodes to one character
    rows, dim = [], 0
    for line in lines:
        toks = _tokens(line, ",\t;")
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
    raw, _ = _lines(path)
    if len(raw) % (dim * 4):
        raise ValueError(f"{path}: size is not a multiple of 4 * dim")
    return np.frombuffer(raw, dtype="<f4").astype(np.float32).reshape(-1, dim)


def read_edges(path):
    """read_edges (src/turbotops.cpp:171-203) -> CsrGraph."""
    _, lines = _lines(path)
    rows = {}
    for lineno, line in enumerate(lines, 1):
        toks = _tokens(line, ",\t")
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
