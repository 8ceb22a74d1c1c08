"""Device condensation (tm_condense_*) against the differentiable torch path of
the same Condenser on the CPU copy of the matrix (the reference's
``_compute_layout`` algorithm, operator/condense.py:182-220) and the float64
oracle: DOF maps, edge lists and masks bit-exact; condensed right-hand side
to FP64 rounding; recover / prolong exact."""
import numpy as np
import pytest
import torch

import tensormesh_amd as tm
from oracle import condense as oc
from tensormesh_amd.mesh import Mesh, kuhn_cube_arrays

pytestmark = pytest.mark.gpu


def _system(n, cuda, seed=0):
    pts, conn, bnd = kuhn_cube_arrays(n)
    m = Mesh(torch.from_numpy(pts), {"tetra": torch.from_numpy(conn)}, {"is_boundary": torch.from_numpy(bnd)})
    kappa = np.exp(np.random.default_rng(seed).standard_normal(len(pts)))
    K = tm.KappaLaplaceElementAssembler.from_mesh(m).to(cuda)(
        point_data={"kappa": torch.tensor(kappa, dtype=torch.float32, device=cuda)})
    b = tm.const_node_assembler(1.0).from_mesh(m).to(cuda)()
    return pts, bnd, K, b


def _masks(bnd, seed):
    rng = np.random.default_rng(seed)
    n = len(bnd)
    keep = np.zeros(n, dtype=bool)
    keep[rng.choice(n, 3, replace=False)] = True
    return {"boundary": bnd, "random": rng.random(n) < 0.3, "none": np.zeros(n, dtype=bool), "almost_all": ~keep}


@pytest.mark.parametrize("kind", ["boundary", "random", "none", "almost_all"])
def test_device_layout_matches_torch_path(cuda, kind):
    print(f"case condense-layout {kind}", flush=True)
    pts, bnd, K, b = _system(4, cuda)
    mask = _masks(bnd, 7)[kind]
    dv_full = np.where(mask, pts[:, 0] - 3 * pts[:, 1] + 0.5, 0.0)
    dv = torch.tensor(dv_full[mask], dtype=torch.float64)
    cg = tm.Condenser(torch.from_numpy(mask).to(cuda), dv.to(cuda))
    cc = tm.Condenser(torch.from_numpy(mask), dv)
    Kg, fg = cg(K, b)
    Kc, fc = cc(K.cpu(), b.cpu())
    assert cg._dev is not None and "inner_dof" in cg._dev, "device layout not used"
    for name in ("inner_row", "inner_col", "ou2in_row", "ou2in_col", "is_inner_edge", "is_ou2in_edge",
                 "is_inner_dof", "is_outer_dof"):
        g, c = getattr(cg, name).cpu(), getattr(cc, name)
        assert g.dtype == c.dtype and torch.equal(g, c), name
    assert (cg.n_inner_dof, cg.n_outer_dof, cg.inner_shape, cg.ou2in_shape) == \
        (cc.n_inner_dof, cc.n_outer_dof, cc.inner_shape, cc.ou2in_shape)
    assert torch.equal(Kg.values.cpu(), Kc.values) and torch.equal(Kg.row.cpu(), Kc.row)
    # condensed CSR view equals the CSR of the condensed COO
    csr = cg._dev["csr_ii"]
    rows = np.repeat(np.arange(cg.n_inner_dof), np.diff(csr.row_ptr.cpu().numpy()))
    assert np.array_equal(rows, Kc.row.numpy()) and np.array_equal(csr.col_idx.cpu().numpy(), Kc.col.numpy())
    # f_i - K_io u_o against the float64 oracle
    kv, kr, kc, ni, fref, _ = oc.condense(K.values.double().cpu().numpy(), K.row.cpu().numpy(),
                                          K.col.cpu().numpy(), mask, b.cpu().numpy(), dv_full[mask])
    assert ni == cg.n_inner_dof
    scale = max(1.0, float(np.abs(fref).max())) if ni else 1.0
    assert np.abs(fg.cpu().numpy() - fref).max(initial=0.0) <= 1e-13 * scale
    # condense_rhs reuses the cached K_io block (the CPU torch path forms K_io u_o in the
    # values' FP32, so the FP64 oracle is the reference here)
    assert np.abs(cg.condense_rhs(2.0 * b).cpu().numpy() - fref - b.cpu().numpy()[~mask]).max(
        initial=0.0) <= 1e-13 * scale
    if ni:
        u = torch.randn(ni, dtype=torch.float64, device=cuda)
        assert torch.equal(cg.recover(u).cpu(), cc.recover(u.cpu()))
        assert torch.equal(cg.prolong(u).cpu(), cc.prolong(u.cpu()))
        assert torch.equal(cg.restrict(b).cpu(), cc.restrict(b.cpu()))


def test_device_condense_keeps_gradients(cuda):
    """With gradients required the differentiable path runs and matches the kernels' values."""
    print("case condense-grad", flush=True)
    pts, bnd, K, b = _system(3, cuda)
    c = tm.Condenser(torch.from_numpy(bnd).to(cuda))
    Ki, fi = c(K, b)
    u = torch.randn(c.n_inner_dof, dtype=torch.float64, device=cuda, requires_grad=True)
    full = c.recover(u)
    full.sum().backward()
    assert torch.equal(u.grad, torch.ones_like(u))
    with torch.no_grad():
        assert torch.equal(c.recover(u.detach()), full.detach())
