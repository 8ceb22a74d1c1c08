from .matrix import SparseMatrix
from .ops import CSR, pcg_raw
