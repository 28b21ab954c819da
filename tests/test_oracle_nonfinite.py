"""Dense matrices of clouds with non-finite coordinates: the restatement
(oracle/ttoracle.c) against the reference compiled from its own sources
(NO_SIMD build, oracle/_ref), bit for bit, NaN payloads included; and the L2
NaN cells against a from-scratch statement of the x86 rules they follow
(a - b returns the NaN operand quieted, a first, or the default NaN
0xffc00000 for inf - inf; fma(d, d, s) returns d's NaN; so a cell holds the
NaN of the last NaN term of d(x_i, x_j), i > j). The GPU kernels are held to
the restatement on the same clouds by test_dist_gpu.py."""
import numpy as np
import pytest

from conftest import bits

EXACT = ["l1", "l2", "linf"]


def _cloud(n, d, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n, d)).astype(np.float32)
    x[rng.integers(0, n, 12), rng.integers(0, d, 12)] = np.nan
    x[rng.integers(0, n, 6), rng.integers(0, d, 6)] = np.inf
    x[rng.integers(0, n, 6), rng.integers(0, d, 6)] = -np.inf
    x[rng.integers(0, n, 3), rng.integers(0, d, 3)] = -np.float32(np.nan)
    return x


def _term_nan(ak, bk):
    """The x86 result of a - b when it is NaN, else None."""
    if (ak & 0x7fffffff) > 0x7f800000:
        return ak | 0x00400000
    if (bk & 0x7fffffff) > 0x7f800000:
        return bk | 0x00400000
    if (ak & 0x7fffffff) == 0x7f800000 and ak == bk:
        return 0xffc00000
    return None


def _x86_l2_nan(a, b, last=True):
    terms = [_term_nan(int(ak), int(bk)) for ak, bk in zip(a.view(np.uint32), b.view(np.uint32))]
    terms = [t for t in terms if t is not None]
    return terms[-1] if last else terms[0]


@pytest.mark.parametrize("n,d", [(120, 3), (90, 4), (60, 21)])
def test_restatement_matches_reference_on_nonfinite(restatement, reference, n, d):
    x = _cloud(n, d, seed=n + d)
    for metric in EXACT:
        want_full, want = reference.make_dense(x, metric)
        assert want_full
        got = restatement.make_dense(x, metric, True)
        assert np.array_equal(bits(got), bits(want)), metric
        assert not np.any(bits(np.diagonal(got))), metric


@pytest.mark.parametrize("n,d", [(120, 3), (90, 4), (60, 21)])
def test_l2_nan_cells_follow_the_x86_rule(restatement, n, d):
    x = _cloud(n, d, seed=n + d)
    m = bits(restatement.make_dense(x, "l2", True))
    nan_cells = 0
    for i in range(n):
        for j in range(i):
            if (int(m[i, j]) & 0x7fffffff) > 0x7f800000:
                nan_cells += 1
                want = _x86_l2_nan(x[i], x[j])
                assert int(m[i, j]) == want and int(m[j, i]) == want, (i, j)
    assert nan_cells > 100
    # a planted mistake, the first NaN term instead of the last, disagrees
    first = sum(int(m[i, j]) != _x86_l2_nan(x[i], x[j], last=False)
                for i in range(n) for j in range(i) if (int(m[i, j]) & 0x7fffffff) > 0x7f800000)
    assert first > 0
