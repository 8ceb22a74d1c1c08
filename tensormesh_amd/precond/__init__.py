"""Learned preconditioner: a U-shaped edge message-passing network over
nested node aggregates (BF16 MFMA kernels with FP32 accumulation, hand-written
forward and backward), trained on plain Jacobi-PCG data and applied inside
flexible CG with Jacobi-PCG as the fallback."""
from .hierarchy import Hierarchy, build_host_levels
from .network import Adam, GNNPreconditioner
from .fcg import LearnedFCG

__all__ = ["Hierarchy", "build_host_levels", "GNNPreconditioner", "Adam", "LearnedFCG"]
