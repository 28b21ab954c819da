"""Differentiable persistence diagrams.

The reference computes diagrams only (include/turbotops/rips.hpp:52-67: a
Barcode keeps bar values, not simplices). Here each bar also carries its
critical edges: the longest edge of its birth simplex and of its death
simplex. Every finite bar value is the length of its critical edge, so the
gradient of a loss on the diagram is a scatter of dL/d(birth, death) into
the two endpoints of each critical edge, times the derivative of the
Minkowski distance (include/turbotops/metric.hpp:92-207). That scatter is
one HIP kernel behind the C ABI (tt_diagram_backward[_dev]); the forward
is the persistence() pipeline itself.

The pairing is piecewise constant in the points, so the gradient is exact
wherever no two filtration values tie; at ties it is the one-sided gradient
of the pairing the forward chose.
"""
import ctypes as C
import math

import numpy as np
import torch

from . import capi
from .api import persistence_pairs
from .frontend import parse_metric


def _stream_ptr(t):
    return C.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


def diagram_backward(x, edges, grad, metric="l2"):
    """dL/dx for a loss whose gradient with respect to the lengths of the
    edges (m, 2) is grad (m,). x, edges and grad are torch tensors (all on
    one HIP device, or all on the CPU) or numpy arrays; the result has x's
    shape, float32, where x lives. The work runs in the library's kernel
    either way."""
    kind, p = parse_metric(metric) if not isinstance(metric, tuple) else metric
    lib = capi.lib()
    if isinstance(x, torch.Tensor) and x.is_cuda:
        xs = x.detach().to(torch.float32).contiguous()
        e = torch.as_tensor(edges, device=x.device).to(torch.int32).contiguous()
        g = torch.as_tensor(grad, device=x.device).to(torch.float32).contiguous()
        n, dim = xs.shape
        m = e.shape[0]
        if e.numel() and (int(e.min()) < 0 or int(e.max()) >= n):
            raise ValueError("diagram_backward: edge endpoint out of range")
        if g.shape[0] != m:
            raise ValueError("diagram_backward: one gradient per edge")
        dx = torch.empty_like(xs)
        lib.call("tt_diagram_backward_dev", C.c_void_p(xs.data_ptr()), n, dim, kind, p,
                 C.c_void_p(e.data_ptr() if m else 0), C.c_void_p(g.data_ptr() if m else 0), m,
                 C.c_void_p(dx.data_ptr() if dx.numel() else 0), _stream_ptr(xs))
        return dx
    xs = np.ascontiguousarray(x.detach().cpu().numpy() if isinstance(x, torch.Tensor) else x, np.float32)
    e = np.ascontiguousarray(np.asarray(edges).reshape(-1, 2), np.uint32)
    g = np.ascontiguousarray(np.asarray(grad).reshape(-1), np.float32)
    n, dim = xs.shape
    if g.shape[0] != e.shape[0]:
        raise ValueError("diagram_backward: one gradient per edge")
    dx = np.empty_like(xs)
    ad = lambda a: C.c_void_p(a.ctypes.data) if a.size else None
    lib.call("tt_diagram_backward", ad(xs), n, dim, kind, p, ad(e), ad(g), e.shape[0], ad(dx))
    return torch.from_numpy(dx) if isinstance(x, torch.Tensor) else dx


class _Pairing:
    """The critical edges of one diagram and where each bar's gradient goes:
    chunks (dim, column 0 = birth / 1 = death, bar indices) in edge order."""

    def __init__(self, birth_edges, death_edges):
        chunks, edges = [], []
        for d, (be, de) in enumerate(zip(birth_edges, death_edges)):
            for col, e in ((0, np.asarray(be)), (1, np.asarray(de))):
                sel = np.nonzero(e[:, 0] >= 0)[0] if e.size else np.zeros(0, np.int64)
                if sel.size:
                    chunks.append((d, col, sel))
                    edges.append(e[sel])
        self.chunks = chunks
        self.edges = np.concatenate(edges, 0) if edges else np.zeros((0, 2), np.int64)

    def edge_grad(self, grads, device):
        """Gradients of the critical-edge lengths from dL/d(diagram)."""
        parts = []
        for d, col, sel in self.chunks:
            g = grads[d] if d < len(grads) else None
            if g is None:
                parts.append(torch.zeros(sel.size, dtype=torch.float32, device=device))
            else:
                idx = torch.as_tensor(sel, device=g.device)
                parts.append(g[idx, col].to(device=device, dtype=torch.float32))
        if not parts:
            return torch.zeros(0, dtype=torch.float32, device=device)
        return torch.cat(parts)


def _as_bars(dgms, like):
    return tuple(torch.as_tensor(np.asarray(a, np.float32)).to(like.device) for a in dgms)


class Diagram(torch.autograd.Function):
    """x (n, d) -> one (m_k, 2) tensor of bars per dimension k <= maxdim,
    equal to persistence(x, ...); differentiable in x."""

    @staticmethod
    def forward(ctx, x, maxdim, threshold, metric, threads):
        host = x.detach().to("cpu", torch.float32).contiguous().numpy()
        dgms, be, de, _ = persistence_pairs(host, maxdim=max(1, maxdim), threshold=threshold, metric=metric,
                                            threads=threads)
        dgms, be, de = dgms[: maxdim + 1], be[: maxdim + 1], de[: maxdim + 1]
        ctx.pairing = _Pairing(be, de)
        ctx.metric = metric
        ctx.save_for_backward(x)
        ctx.set_materialize_grads(False)
        return _as_bars(dgms, x)

    @staticmethod
    def backward(ctx, *grads):
        (x,) = ctx.saved_tensors
        pr = ctx.pairing
        g = pr.edge_grad(grads, x.device)
        e = torch.as_tensor(pr.edges, device=x.device)
        dx = diagram_backward(x, e, g, ctx.metric)
        return dx.to(x.dtype), None, None, None, None


def diagram(x, maxdim=1, threshold=math.inf, metric="l2", threads=0):
    """persistence(x)'s diagrams as tensors that carry gradients to x."""
    if not isinstance(x, torch.Tensor):
        raise TypeError("diagram: x must be a torch tensor")
    if x.dim() != 2:
        raise ValueError("x must be a 2-d array (n, d)")
    return Diagram.apply(x, int(maxdim), float(threshold), str(metric), int(threads))


class _BatchDiagram(torch.autograd.Function):
    @staticmethod
    def forward(ctx, runner, metric, threads, *xs):
        host = [x.detach().to("cpu", torch.float32).contiguous().numpy() for x in xs]
        res = runner(host, use_graph=True, threads=threads, pairs=True)
        ctx.pairings = [_Pairing(be, de) for _, be, de, _ in res]
        ctx.metric = metric
        ctx.save_for_backward(*xs)
        ctx.set_materialize_grads(False)
        out = []
        for x, (dgms, _, _, _) in zip(xs, res):
            out.extend(_as_bars(dgms, x))
        ctx.ndims = [len(r[0]) for r in res]
        return tuple(out)

    @staticmethod
    def backward(ctx, *grads):
        xs = ctx.saved_tensors
        # one kernel launch per point dimension: clouds of equal dimension are
        # packed end to end with their edges offset
        per_cloud, at = [], 0
        for nd in ctx.ndims:
            per_cloud.append(grads[at: at + nd])
            at += nd
        groups = {}
        for c, x in enumerate(xs):
            groups.setdefault((x.shape[1], x.device), []).append(c)
        out = [None] * len(xs)
        for (_, device), members in groups.items():
            pts, edges, gs, offs, off = [], [], [], [], 0
            for c in members:
                x = xs[c]
                pr = ctx.pairings[c]
                pts.append(x.detach().to(torch.float32))
                edges.append(torch.as_tensor(pr.edges, device=device) + off)
                gs.append(pr.edge_grad(per_cloud[c], device))
                offs.append(off)
                off += x.shape[0]
            dx = diagram_backward(torch.cat(pts), torch.cat(edges), torch.cat(gs), ctx.metric)
            for c, o in zip(members, offs):
                out[c] = dx[o: o + xs[c].shape[0]].to(xs[c].dtype)
        return (None, None, None, *out)


class DiagramBatch:
    """Differentiable diagrams of many clouds through the batched front end
    (one recorded launch sequence for all clouds, host reductions in
    parallel). Create once per set of shapes; call with a list of (n_i, d_i)
    tensors for a list of per-cloud tuples of bar tensors (H0, H1). The
    backward is one kernel launch per distinct point dimension."""

    def __init__(self, shapes, metric="l2"):
        from .api import BatchPersistence
        self.metric = str(metric)
        self.bp = BatchPersistence(shapes, metric=metric, maxdim=1)

    def __call__(self, xs, threads=0):
        xs = list(xs)
        flat = _BatchDiagram.apply(self._run, self.metric, int(threads), *xs)
        out, at = [], 0
        for _ in xs:
            out.append(tuple(flat[at: at + 2]))
            at += 2
        return out

    def _run(self, host, use_graph, threads, pairs):
        return self.bp(host, use_graph=use_graph, threads=threads, pairs=pairs)

    @property
    def device_ms(self):
        return self.bp.device_ms


def top_k_persistence_loss(bars, k=8):
    """-sum(death - birth) over the k longest finite bars of one dimension."""
    if bars.shape[0] == 0:
        return bars.sum() * 0
    fin = torch.isfinite(bars[:, 1])
    b = bars[fin]
    life = b[:, 1] - b[:, 0]
    top = torch.topk(life, min(k, life.shape[0])).values
    return -top.sum()


__all__ = ["diagram", "Diagram", "DiagramBatch", "diagram_backward", "top_k_persistence_loss"]
