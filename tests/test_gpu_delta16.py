"""The compact int16 column format of the scalar Jacobi-PCG (``tm_csr_delta16``,
``tm_pcg_jacobi_d16``): the same iteration as the int32 path bit for bit, and
the int32 path when an entry lies too far from the diagonal."""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

import tensormesh_amd as tm
from tensormesh_amd.mesh import Mesh, kuhn_cube_arrays
from tensormesh_amd.sparse import ops

pytestmark = pytest.mark.gpu


def _solve(csr, vals, b, use16):
    prev = ops.COL_DELTA16
    ops.COL_DELTA16 = use16
    try:
        return ops.pcg_raw(csr, vals, b, rtol=1e-10, maxiter=5000)
    finally:
        ops.COL_DELTA16 = prev


@pytest.mark.parametrize("vdt", [torch.float32, torch.float64])
def test_delta16_matches_int32_bit_for_bit(cuda, vdt):
    print(f"case delta16 poisson {vdt}", flush=True)
    pts, conn, bnd = kuhn_cube_arrays(14)
    m = Mesh(torch.from_numpy(pts), {"tetra": torch.from_numpy(conn)}, {"is_boundary": torch.from_numpy(bnd)})
    asm = tm.KappaLaplaceElementAssembler.from_mesh(m).to(cuda)
    asm.value_dtype = vdt
    kappa = torch.tensor(np.exp(np.random.default_rng(3).standard_normal(len(pts))), dtype=torch.float32,
                         device=cuda)
    K = asm(point_data={"kappa": kappa})
    Ki, fi = tm.Condenser(m.boundary_mask.to(cuda))(K, tm.const_node_assembler(1.0).from_mesh(m).to(cuda)())
    csr = Ki.csr()
    cd = ops.col_delta16(csr)
    assert cd is not None and cd.dtype == torch.int16
    rows = torch.repeat_interleave(torch.arange(csr.n_rows, device=cuda), torch.diff(csr.row_ptr.long()))
    assert torch.equal(rows + cd.long(), csr.col_idx.long())
    x16, it16, r16, ok16 = _solve(csr, Ki.values, fi, True)
    x32, it32, r32, ok32 = _solve(csr, Ki.values, fi, False)
    assert ok16 and ok32 and it16 == it32 and r16 == r32 and torch.equal(x16, x32)


def test_delta16_falls_back_when_an_entry_is_far(cuda):
    """A banded SPD matrix with one symmetric pair 40,000 apart: no int16 copy,
    the solve runs on int32 columns and matches SciPy."""
    print("case delta16 fallback", flush=True)
    n = 50000
    main = np.full(n, 4.0)
    off = np.full(n - 1, -1.0)
    A = sp.diags([off, main, off], [-1, 0, 1], format="lil")
    A[3, 40003] = -0.5
    A[40003, 3] = -0.5
    A = A.tocsr()
    A.sort_indices()
    csr = ops.CSR(torch.tensor(A.indptr, dtype=torch.int32, device=cuda),
                  torch.tensor(A.indices, dtype=torch.int32, device=cuda), n, n)
    vals = torch.tensor(A.data, dtype=torch.float32, device=cuda)
    assert ops.col_delta16(csr) is None and csr.col_delta is False
    b = torch.ones(n, dtype=torch.float64, device=cuda)
    x, it, rel, ok = ops.pcg_raw(csr, vals, b, rtol=1e-10)
    A32 = sp.csr_matrix((A.data.astype(np.float32).astype(np.float64), A.indices, A.indptr), shape=(n, n))
    xr = sp.linalg.spsolve(A32.tocsc(), np.ones(n))
    assert ok and np.abs(x.cpu().numpy() - xr).max() <= 1e-8 * np.abs(xr).max()
