"""CPU reference gradient of a persistence diagram — TEST INFRASTRUCTURE ONLY.

The reference has no gradient (include/turbotops/rips.hpp:52-67: a Barcode
keeps values only), so this is a restatement from the mathematics: every
finite bar value of a Rips diagram is the length d(a, b) of its critical
edge, hence

    dL/dx_a = sum over edges e = (a, b) of  g_e * d d(x_a, x_b) / d x_a,

with the derivative of the Minkowski distance of include/turbotops/
metric.hpp:92-207 (L2 (x_a - x_b)/d; L1 sign; L-inf the sign on the first
coordinate of largest |diff|; Lp sign(diff) |diff|^(p-1) / d^(p-1)), and
d/dx_b = -d/dx_a. Evaluated in float64 on the float32 coordinates.

Pins (tests/test_oracle.py): central finite differences in float64 of
L(x) = sum_e g_e d(x_a, x_b) for every metric, where a wrong sign, a dropped
endpoint, a swapped operand or a wrong Lp exponent each fail; and torch's
own autograd of the same L for L2, L1 and L-inf.
"""
import numpy as np

KINDS = {"l1": 0, "l2": 1, "linf": 2, "lp": 3}


def _kind(metric):
    m = str(metric).lower()
    if m.startswith("lp:"):
        p = float(m[3:])
        if p == 1:
            return 0, 1.0
        if p == 2:
            return 1, 2.0
        if np.isinf(p):
            return 2, np.inf
        return 3, p
    return {"l1": (0, 1.0), "manhattan": (0, 1.0), "l2": (1, 2.0), "euclidean": (1, 2.0),
            "linf": (2, np.inf), "chebyshev": (2, np.inf)}[m]


def dist64(x, a, b, metric):
    """Minkowski distance of rows a and b of x in float64."""
    kind, p = _kind(metric)
    t = np.abs(np.asarray(x, np.float64)[a] - np.asarray(x, np.float64)[b])
    if kind == 0:
        return t.sum(-1)
    if kind == 1:
        return np.sqrt((t * t).sum(-1))
    if kind == 2:
        return t.max(-1)
    return (t ** p).sum(-1) ** (1.0 / p)


def edge_grad64(x, edges, grad, metric):
    """dL/dx (float64, x's shape) for L = sum_e grad[e] * d(x[a_e], x[b_e])."""
    kind, p = _kind(metric)
    x = np.asarray(x, np.float32).astype(np.float64)
    e = np.asarray(edges, np.int64).reshape(-1, 2)
    g = np.asarray(grad, np.float64).reshape(-1)
    out = np.zeros_like(x)
    if e.shape[0] == 0:
        return out
    a, b = e[:, 0], e[:, 1]
    diff = x[a] - x[b]
    if kind == 1:
        s = np.sqrt((diff * diff).sum(1))
        with np.errstate(invalid="ignore", divide="ignore"):
            dd = np.where(s[:, None] > 0, diff / s[:, None], 0.0)
    elif kind == 0:
        dd = np.sign(diff)
    elif kind == 2:
        arg = np.argmax(np.abs(diff), axis=1)          # first largest, as the kernel
        dd = np.zeros_like(diff)
        rows = np.arange(diff.shape[0])
        dd[rows, arg] = np.sign(diff[rows, arg])
    else:
        s = (np.abs(diff) ** p).sum(1)
        d = s ** (1.0 / p)
        with np.errstate(invalid="ignore", divide="ignore"):
            dd = np.where(s[:, None] > 0, np.sign(diff) * np.abs(diff) ** (p - 1) / d[:, None] ** (p - 1), 0.0)
    dd[a == b] = 0.0
    contrib = g[:, None] * dd
    np.add.at(out, a, contrib)
    np.add.at(out, b, -contrib)
    return out
