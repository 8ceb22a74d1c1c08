"""The allocation-free Jacobi-PCG entry point (``tm_pcg_jacobi_noalloc``,
workspace in the library's static device array) against ``tm_pcg_jacobi``
with a caller workspace: same kernels, so the iterate, the iteration count
and the residual must agree bit for bit.  Two calls in a row, the second on a
smaller system (the array keeps the first call's contents), and a system above
the static limit, which is refused before any device work."""
import ctypes

import numpy as np
import pytest
import torch

import tensormesh_amd as tm
from tensormesh_amd import _lib
from tensormesh_amd.mesh import Mesh, kuhn_cube_arrays
from tensormesh_amd.sparse.ops import pcg_raw

pytestmark = pytest.mark.gpu


def _system(n, cuda, seed, vdt):
    pts, conn, bnd = kuhn_cube_arrays(n)
    m = Mesh(torch.from_numpy(pts), {"tetra": torch.from_numpy(conn)}, {"is_boundary": torch.from_numpy(bnd)})
    kappa = np.exp(np.random.default_rng(seed).standard_normal(len(pts)))
    K = tm.KappaLaplaceElementAssembler.from_mesh(m).to(cuda)(
        point_data={"kappa": torch.tensor(kappa, dtype=torch.float32, device=cuda)})
    b = tm.const_node_assembler(1.0).from_mesh(m).to(cuda)()
    Ki, fi = tm.Condenser(m.boundary_mask.to(cuda))(K.to(vdt), b)
    return Ki.csr(), Ki.values.contiguous(), fi.to(torch.float64).contiguous()


def _noalloc(csr, values, b, rtol, maxiter=20000):
    L = _lib.lib()
    x = torch.zeros(csr.n_rows, dtype=torch.float64, device=values.device)
    it, rr = ctypes.c_int64(0), ctypes.c_double(0.0)
    st = L.tm_pcg_jacobi_noalloc(csr.n_rows, _lib.ptr(csr.row_ptr), _lib.ptr(csr.col_idx), _lib.ptr(values),
                                 _lib.dtype_code(values), _lib.ptr(b), _lib.ptr(x), float(rtol), 0.0, int(maxiter),
                                 0, ctypes.byref(it), ctypes.byref(rr), _lib.stream_of(values.device))
    return st, x, int(it.value), float(rr.value)


@pytest.mark.parametrize("vdt", [torch.float32, torch.float64])
def test_noalloc_pcg_matches_workspace_pcg(cuda, vdt):
    for n, seed in [(40, 1), (20, 2)]:          # second call smaller than the first
        print(f"case noalloc pcg n={n} {vdt}", flush=True)
        csr, values, b = _system(n, cuda, seed, vdt)
        st, x, it, rel = _noalloc(csr, values, b, 1e-8)
        assert st == _lib.TM_OK, _lib.lib().tm_last_error().decode()
        xr, itr, relr, ok = pcg_raw(csr, values, b, rtol=1e-8, maxiter=20000)
        assert ok and it == itr and rel == relr and it > 10
        assert torch.equal(x, xr)


def test_noalloc_pcg_refuses_oversize(cuda):
    L = _lib.lib()
    nmax = int(L.tm_pcg_noalloc_max_rows())
    assert nmax == 1 << 20
    dummy = torch.zeros(4, dtype=torch.int32, device=cuda)
    it, rr = ctypes.c_int64(0), ctypes.c_double(0.0)
    st = L.tm_pcg_jacobi_noalloc(nmax + 1, _lib.ptr(dummy), _lib.ptr(dummy), _lib.ptr(dummy), _lib.F32,
                                 _lib.ptr(dummy), _lib.ptr(dummy), 1e-8, 0.0, 10, 0, ctypes.byref(it),
                                 ctypes.byref(rr), _lib.stream_of(cuda))
    assert st == _lib.TM_ERR_LIMIT
    assert "exceed the static workspace" in L.tm_last_error().decode()
