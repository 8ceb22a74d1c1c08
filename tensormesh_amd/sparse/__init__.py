from .matrix import SparseMatrix
from .ops import CSR, pcg_raw
from .nonlinear import nonlinear_solve, NonLinearSolveFunction
