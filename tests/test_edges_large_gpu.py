"""Dimension-0 edges of a dense complex with more than 2^32 vertex pairs
(n = 93000: C(n, 2) = 4,324,453,500), which needs 64-bit row offsets, edge
indices of 33 bits and a radix sort over more than 2^32 keys (64-bit digit
offsets). The reference bounds this case only by its memory budget
(rips.hpp:391-416). Too large for the oracle, so the checks are the ones the
mathematics fixes, on the device:
  - every pair is kept below a threshold above the largest distance, so the
    edge count is C(n, 2) and the indices are a permutation of [0, C(n, 2)):
    their sum is C(n, 2)(C(n, 2) - 1)/2 and their per-row counts are i;
  - the order is `precedes` (rips.hpp:116-118): diameter ascending, index
    descending among equal diameters;
  - each diameter is the matrix entry its index decodes to (sampled), and the
    sum of the diameters' bits equals the matrix's lower triangle's;
  - the apparent flag of sampled edges equals a brute-force search for the
    zero-pivot cofacet and facet (rips.hpp:585-665) over the matrix rows."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _decode(idx):
    import torch
    v = ((1.0 + torch.sqrt(1.0 + 8.0 * idx.double())) * 0.5).long()
    v = torch.where(v * (v - 1) // 2 > idx, v - 1, v)
    v = torch.where((v + 1) * v // 2 <= idx, v + 1, v)
    return v, idx - v * (v - 1) // 2


def test_dense_edges_beyond_2_32_pairs(gpu):
    import torch
    from turbotops_amd import dense_edges_dev, dist_matrix_dev
    n = 93000
    M = n * (n - 1) // 2
    assert M > 2 ** 32
    rng = np.random.default_rng(70)
    a = rng.uniform(0, 2 * np.pi, n)
    x = np.stack([np.cos(a), np.sin(a)], 1).astype(np.float32)
    print(f"dense edges n={n} pairs={M}", flush=True)
    xt = torch.from_numpy(x).cuda()
    mat = dist_matrix_dev(xt, "l2", "full").view(n, n)
    k, idx, d, ap = dense_edges_dev(mat, 1e30, apparent=True)
    assert k == M
    print("count ok; checking the order", flush=True)
    # indices: a permutation of [0, M): the sums of k and k^2 over [0, M),
    # in int64 arithmetic that wraps modulo 2^64 as torch's does
    def wrapped(v):
        v %= 1 << 64
        return v - (1 << 64) if v >= 1 << 63 else v
    s1 = s2 = 0
    for s0 in range(0, k, 1 << 28):
        part = idx[s0:s0 + (1 << 28)]
        s1 += int(part.sum(dtype=torch.int64))
        s2 += int((part * part).sum(dtype=torch.int64))
    assert wrapped(s1) == wrapped(M * (M - 1) // 2)
    assert wrapped(s2) == wrapped((M - 1) * M * (2 * M - 1) // 6)
    assert int(idx.min()) == 0 and int(idx.max()) == M - 1
    step = 1 << 28
    for s0 in range(0, k - 1, step):
        s1 = min(k, s0 + step + 1)
        dd, ii = d[s0:s1], idx[s0:s1]
        up = dd[1:] > dd[:-1]
        tie = dd[1:] == dd[:-1]
        assert bool(torch.all(up | tie))
        assert bool(torch.all(~tie | (ii[1:] < ii[:-1])))
    # per-row counts (a permutation has exactly i indices in row i)
    rows, cols = _decode(idx[:: 4099])
    assert bool(torch.all((cols >= 0) & (cols < rows) & (rows < n)))
    # diameters: sampled against the matrix, and the bit sums of all
    g = torch.Generator(device="cuda").manual_seed(71)
    pos = torch.randint(0, k, (1 << 20,), device="cuda", generator=g)
    r, c = _decode(idx[pos])
    assert torch.equal(d[pos].view(torch.int32), mat[r, c].view(torch.int32))
    total = 0
    for i0 in range(0, n, 4096):
        blk = mat[i0:i0 + 4096]
        low = torch.tril(torch.ones_like(blk, dtype=torch.bool), diagonal=i0 - 1)
        total += int((blk.view(torch.int32).long() * low).sum())
    dsum = 0
    for s0 in range(0, k, step):
        dsum += int(d[s0:s0 + step].view(torch.int32).long().sum())
    assert dsum == total
    print("order and diameters ok; checking sampled apparent flags", flush=True)
    # apparent flags by brute force on 3000 sampled edges
    sel = torch.randint(0, k, (3000,), device="cuda", generator=g)
    si, sj = _decode(idx[sel])
    ws = torch.arange(n, device="cuda")
    for e in range(3000):
        i, j, dv = int(si[e]), int(sj[e]), float(d[sel[e]])
        ri, rj = mat[i], mat[j]
        ok = (torch.maximum(ri, rj) <= dv) & (ws != i) & (ws != j)
        hit = torch.nonzero(ok)
        if hit.numel() == 0:
            want = False
        else:
            w = int(hit.max())
            diw, djw = float(ri[w]), float(rj[w])
            if w > i:
                want = True
            elif w > j:
                want = djw != dv
            else:
                want = djw != dv and diw != dv
        assert bool(ap[sel[e]]) == want, (i, j)
    n_ap = int(ap.sum(dtype=torch.int64))
    print(f"apparent flags ok ({n_ap} apparent of {k})", flush=True)
