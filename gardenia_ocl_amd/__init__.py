"""gardenia_ocl_amd -- MI355X-native pull PageRank and CSR SpMV.

A drop-in for the PageRank (src/pr), SpMV (src/spmv) and SymGS (src/symgs)
solver plug-ins of
gardenia_ocl: hand-written HIP kernels for gfx950 behind a C ABI
(include/gardenia_hip.h, include/gardenia_graph.h), with Python mirrors of the
reference's ``PRSolver`` / ``SpmvSolver`` and device-resident plans over torch
tensors.
"""
from . import graph
from ._native import GardeniaError, LIB_PATH
from .pr import (EPSILON, K_DAMP, MAX_ITER, PageRankPlan, PRSolver, PRVerifier, pagerank_push,
                 verify_residual)
from .spmv import SpmvSolver, SpmvPlan
from .symgs import SymGSPlan, SymGSSolver
from .ppr import PersonalizedPageRankPlan
from .multigraph import MultiGraphPageRank

__all__ = [
    "graph", "GardeniaError", "PersonalizedPageRankPlan", "MultiGraphPageRank", "LIB_PATH", "PRSolver", "PRVerifier", "pagerank_push", "verify_residual", "PageRankPlan", "SpmvSolver", "SpmvPlan",
    "SymGSPlan", "SymGSSolver",
    "EPSILON", "K_DAMP", "MAX_ITER",
]
