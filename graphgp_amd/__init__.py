"""graphgp_amd -- the GraphGP hot path on MI355X (gfx950).

Drop-in for the reference's public API (graphgp/__init__.py:1-42): the same names,
arguments and errors, backed by hand-written HIP kernels reached through the C ABI of
``include/graphgp_amd.h``.  Arrays are torch tensors on the GPU (numpy inputs are moved).
"""
from .tree import build_tree, query_preceding_neighbors, query_neighbors
from .graph import Graph, check_graph, build_graph, compute_depths, order_by_depth
from .refine import (
    generate,
    generate_inv,
    generate_logdet,
    generate_vjp,
    generate_inv_vjp,
    generate_logdet_vjp,
    generate_dense,
    generate_dense_inv,
    generate_dense_logdet,
    refine,
    refine_inv,
    refine_logdet,
    compute_cov_matrix,
)
from .store import save_graph, load_graph
from . import extras

__all__ = [
    "build_tree",
    "query_preceding_neighbors",
    "query_neighbors",
    "Graph",
    "check_graph",
    "build_graph",
    "compute_depths",
    "order_by_depth",
    "generate",
    "generate_inv",
    "generate_logdet",
    "generate_vjp",
    "generate_inv_vjp",
    "generate_logdet_vjp",
    "generate_dense",
    "generate_dense_inv",
    "generate_dense_logdet",
    "refine",
    "refine_inv",
    "refine_logdet",
    "compute_cov_matrix",
    "save_graph",
    "load_graph",
    "extras",
]
