"""turbotops_amd — the MI355X (gfx950) filtration front end of turbotops.

A drop-in for the data-parallel front end of dnbaker/turbotops v0.1.0:
Minkowski distance matrices, the enclosing radius, exact radius graphs,
exact kNN and kNN graphs, H0 via Boruvka, the apparent-pair scan, a batched
form for many clouds and the backward of a diagram, as hand-written HIP
kernels behind the C ABI in include/ttgpu.h.
"""
from . import capi  # noqa: F401
from .api import BatchPersistence, persistence, persistence_batch, persistence_pairs  # noqa: F401
from .frontend import (  # noqa: F401
    CsrGraph,
    DeviceComplex,
    build_edges,
    build_edges_csr,
    dense_edges,
    dense_edges_dev,
    dense_layout,
    dist_matrix_dev,
    knn_graph,
    knn_graph_exact,
    knn_lists_exact,
    make_dense,
    mst_deaths,
    mst_h0,
    parse_metric,
    radius_graph,
    sparse_edges,
)
from .grad import DiagramBatch, diagram, diagram_backward, top_k_persistence_loss  # noqa: F401

__version__ = "0.1.0"
