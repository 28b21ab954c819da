"""The backward of a diagram (tt_diagram_backward[_dev], turbotops_amd.grad).

* The kernel against the CPU reference gradient (oracle/grad.py) on random
  edge sets, every metric, with repeated, reversed and zero-length edges.
* Critical edges: each finite bar value equals its critical edge's length in
  the filtration's own float32 distances, bit for bit.
* The differentiable form: autograd through diagram() equals central
  differences in float64 of the loss with the pairing fixed, and central
  differences of the real diagram (float32 forward) wherever the pairing
  does not change (noisy circle, n = 256, seed 8; loss = -sum(death - birth)
  over the top 8 H1 bars).
* Torch's own autograd of the same loss over the critical edges.
* The batched form: every cloud's gradient equals the single-cloud one.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GRAD_METRICS = ["l2", "l1", "linf", "lp:1.5", "lp:3"]


def _edges(seed, n, m):
    rng = np.random.default_rng(seed)
    e = rng.integers(0, n, size=(m, 2))
    e = np.concatenate([e, e[:5], e[:5, ::-1], [[0, 0], [n - 1, n - 1]]])
    g = rng.normal(size=e.shape[0]).astype(np.float32)
    return e, g


@pytest.mark.parametrize("metric", GRAD_METRICS)
@pytest.mark.parametrize("n,dim,m", [(300, 3, 900), (257, 8, 40), (64, 1, 500), (2, 16, 3)])
def test_backward_kernel_vs_oracle(gpu, metric, n, dim, m):
    import torch
    from oracle.grad import edge_grad64
    from turbotops_amd import diagram_backward
    rng = np.random.default_rng(n * 7 + dim)
    x = rng.normal(size=(n, dim)).astype(np.float32)
    x[1] = x[0]                                          # a zero-length edge (0, 1)
    e, g = _edges(n + m, n, m)
    e = np.concatenate([e, [[0, 1]]])
    g = np.concatenate([g, [1.5]]).astype(np.float32)
    print(f"case backward metric={metric} n={n} dim={dim} m={len(e)}", flush=True)
    want = edge_grad64(x, e, g, metric).astype(np.float32)
    host = diagram_backward(x, e, g, metric)
    dev = diagram_backward(torch.tensor(x, device="cuda"), torch.tensor(e, device="cuda"),
                           torch.tensor(g, device="cuda"), metric).cpu().numpy()
    np.testing.assert_allclose(host, want, rtol=2e-6, atol=2e-6)
    assert np.array_equal(host.view(np.uint32), dev.view(np.uint32))   # one kernel, deterministic


def test_backward_empty_and_errors(gpu):
    from turbotops_amd import diagram_backward
    x = np.ones((5, 2), np.float32)
    assert not diagram_backward(x, np.zeros((0, 2), np.int64), np.zeros(0, np.float32)).any()
    with pytest.raises(ValueError, match="out of range"):
        diagram_backward(x, np.array([[0, 5]]), np.ones(1, np.float32))
    with pytest.raises(ValueError, match="one gradient per edge"):
        diagram_backward(x, np.array([[0, 1]]), np.ones(2, np.float32))


def _circle(n, seed):
    from turbotops_amd.synthetic import noisy_circle
    return noisy_circle(n=n, seed=seed)


@pytest.mark.parametrize("metric", ["l2", "linf", "lp:3"])
def test_critical_edges_carry_bar_values(gpu, metric):
    from turbotops_amd import make_dense, persistence, persistence_pairs
    x = _circle(200, 3)
    dgms, be, de, _ = persistence_pairs(x, maxdim=1, metric=metric)
    want, _ = persistence(x, maxdim=1, metric=metric)
    _, mat, _ = make_dense(x, metric)
    for d in range(2):
        assert np.array_equal(np.asarray(dgms[d]), np.asarray(want[d]))
        bars = np.asarray(dgms[d])
        for col, edges in ((0, be[d]), (1, de[d])):
            v = bars[:, col]
            has = edges[:, 0] >= 0
            if d == 0 and col == 0:
                assert not has.any() and (v == 0).all()
                continue
            assert np.array_equal(has, np.isfinite(v))
            assert (edges[has, 0] != edges[has, 1]).all()
            assert np.array_equal(mat[edges[has, 0], edges[has, 1]].view(np.uint32), v[has].view(np.uint32))


def _loss_fixed64(x64, be, de, top):
    """-sum(death - birth) over the chosen H1 bars with their critical edges
    fixed, in float64."""
    from oracle.grad import dist64
    b = dist64(x64, be[top, 0], be[top, 1], "l2")
    d = dist64(x64, de[top, 0], de[top, 1], "l2")
    return -float((d - b).sum())


def _top8(bars):
    life = np.where(np.isfinite(bars[:, 1]), bars[:, 1] - bars[:, 0], -np.inf)
    return np.argsort(-life, kind="stable")[:8]


def test_autograd_fd_circle_n256(gpu):
    import torch
    from turbotops_amd import diagram, persistence_pairs, top_k_persistence_loss
    x = _circle(256, 8)
    print("case autograd fd circle n=256 seed=8", flush=True)
    xt = torch.tensor(x, device="cuda", requires_grad=True)
    dgms = diagram(xt, maxdim=1)
    loss = top_k_persistence_loss(dgms[1], 8)
    loss.backward()
    grad = xt.grad.cpu().numpy().astype(np.float64)
    dg, be, de, _ = persistence_pairs(x, maxdim=1)
    top = _top8(np.asarray(dg[1]))
    assert abs(float(loss.detach()) - (-(dg[1][top, 1] - dg[1][top, 0]).sum())) < 1e-5

    # (1) float64 central differences of the fixed-pairing loss, every coordinate
    x64 = x.astype(np.float64)
    h = 1e-6
    fd = np.zeros_like(x64)
    for i in range(x.shape[0]):
        for k in range(x.shape[1]):
            zp, zm = x64.copy(), x64.copy()
            zp[i, k] += h
            zm[i, k] -= h
            fd[i, k] = (_loss_fixed64(zp, be[1], de[1], top) - _loss_fixed64(zm, be[1], de[1], top)) / (2 * h)
    np.testing.assert_allclose(grad, fd, rtol=1e-5, atol=1e-5)
    assert np.abs(grad).sum() > 1.0                  # the loss does move the points

    # (2) central differences of the real diagram (float32 forward) on the
    # coordinates the loss touches, where the pairing stays put
    ends = np.unique(np.concatenate([be[1][top].ravel(), de[1][top].ravel()]))
    checked = 0
    hh = 1e-3
    for i in ends[:16]:
        for k in range(3):
            vals = []
            for s in (1, -1):
                z = x64.copy()
                z[i, k] += s * hh
                zf = z.astype(np.float32)
                d2, be2, de2, _ = persistence_pairs(zf, maxdim=1)
                t2 = _top8(np.asarray(d2[1]))
                same = (np.array_equal(be2[1][t2], be[1][top]) and np.array_equal(de2[1][t2], de[1][top]))
                vals.append((same, -float((d2[1][t2, 1].astype(np.float64) - d2[1][t2, 0]).sum())))
            if vals[0][0] and vals[1][0]:
                # the perturbation actually applied is the float32 rounding of z
                step = float(np.float32(x64[i, k] + hh)) - float(np.float32(x64[i, k] - hh))
                fdr = (vals[0][1] - vals[1][1]) / step
                assert abs(fdr - grad[i, k]) < 2e-3, (i, k, fdr, grad[i, k])
                checked += 1
    assert checked >= 24                             # of 48; the rest switch pairing


@pytest.mark.parametrize("metric,order", [("l2", 2), ("l1", 1), ("linf", math.inf)])
def test_autograd_vs_torch(gpu, metric, order):
    """Torch's autograd of the same loss over the critical edges (float64 on
    the CPU) against the HIP backward."""
    import torch
    from turbotops_amd import diagram, persistence_pairs
    x = _circle(160, 5)
    xt = torch.tensor(x, device="cuda", requires_grad=True)
    dgms = diagram(xt, maxdim=1, metric=metric)
    w0 = torch.linspace(0.5, 1.5, dgms[0].shape[0], device="cuda")
    fin0 = torch.isfinite(dgms[0][:, 1])
    fin1 = torch.isfinite(dgms[1][:, 1])
    loss = (w0[fin0] * dgms[0][fin0, 1]).sum() + (dgms[1][fin1, 1] - 2 * dgms[1][fin1, 0]).sum()
    loss.backward()
    _, be, de, _ = persistence_pairs(x, maxdim=1, metric=metric)
    xr = torch.tensor(x, dtype=torch.float64, requires_grad=True)

    def length(ed):
        ed = torch.as_tensor(ed)
        return torch.linalg.vector_norm(xr[ed[:, 0]] - xr[ed[:, 1]], ord=order, dim=1)

    f0 = fin0.cpu().numpy()
    f1 = fin1.cpu().numpy()
    ref = (w0.cpu().double()[f0] * length(de[0][f0])).sum() + (length(de[1][f1]) - 2 * length(be[1][f1])).sum()
    ref.backward()
    np.testing.assert_allclose(xt.grad.cpu().numpy(), xr.grad.numpy(), rtol=1e-5, atol=1e-5)


def test_autograd_host_tensor_and_float64(gpu):
    """A CPU tensor and a float64 tensor take the same kernel; the gradient
    comes back on the input's device and dtype."""
    import torch
    from turbotops_amd import diagram, top_k_persistence_loss
    x = _circle(120, 4)
    g = []
    for xt in (torch.tensor(x, device="cuda", requires_grad=True), torch.tensor(x, requires_grad=True),
               torch.tensor(x.astype(np.float64), requires_grad=True)):
        top_k_persistence_loss(diagram(xt, maxdim=1)[1], 8).backward()
        assert xt.grad.device == xt.device and xt.grad.dtype == xt.dtype
        g.append(xt.grad.cpu().float().numpy())
    assert np.array_equal(g[0], g[1]) and np.array_equal(g[0], g[2])


def test_autograd_tiny_clouds(gpu):
    import torch
    from turbotops_amd import diagram
    for n in (1, 2, 3):
        xt = torch.zeros((n, 2), device="cuda", requires_grad=True)
        with torch.no_grad():
            xt += torch.arange(n * 2, device="cuda", dtype=torch.float32).reshape(n, 2)
        dg = diagram(xt, maxdim=1)
        fin = torch.isfinite(dg[0][:, 1])
        (dg[0][fin, 1].sum() + dg[1].sum() * 0).backward()
        assert xt.grad.shape == (n, 2)
        if n == 1:
            assert not xt.grad.any()


def test_batch_grad_equals_single(gpu):
    import torch
    from turbotops_amd import DiagramBatch, diagram, top_k_persistence_loss
    from turbotops_amd.synthetic import uneven_clouds
    clouds = uneven_clouds(count=10, nmin=30, nmax=220, seed=9)
    print("case batch grad clouds=10", flush=True)
    db = DiagramBatch([c.shape for c in clouds])
    xs = [torch.tensor(c, device="cuda", requires_grad=True) for c in clouds]
    outs = db(xs)
    loss = sum(top_k_persistence_loss(o[1], 8) + o[0][torch.isfinite(o[0][:, 1]), 1].sum() for o in outs)
    loss.backward()
    for c, xt, o in zip(clouds, xs, outs):
        x1 = torch.tensor(c, device="cuda", requires_grad=True)
        d1 = diagram(x1, maxdim=1)
        assert torch.equal(d1[0], o[0]) and torch.equal(d1[1], o[1])
        (top_k_persistence_loss(d1[1], 8) + d1[0][torch.isfinite(d1[0][:, 1]), 1].sum()).backward()
        assert torch.equal(x1.grad, xt.grad)
