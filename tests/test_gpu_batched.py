"""GPU parity of the batched multi-mesh assembly + persistent PCG launch."""
import numpy as np
import pytest
import torch

from oracle import assemble as oa
from oracle import condense as oc
from oracle import pattern as op
from oracle import solve as osv
from tensormesh_amd.batched import BatchedPoisson
from tensormesh_amd.dataset import batched_square_meshes

pytestmark = pytest.mark.gpu


def _oracle(m):
    pts, tris, bnd, kappa = m
    u, v, pe = op.pattern_and_slots(tris, len(pts))
    k32 = kappa.astype(np.float32).astype(np.float64)
    vals = oa.reduce(pe, oa.element_matrices(pts, tris, "triangle", "kappa_laplace", kappa=k32), len(u))
    b = oa.load_vector(pts, tris, "triangle", len(pts), c=1.0)
    kv, kr, kc, ni, fi, _ = oc.condense(vals, u, v, bnd, b)
    A = osv.csr(kv, kr, kc, ni)
    x, it, _ = osv.jacobi_pcg(A, fi, rtol=1e-8)
    return osv.direct(A, fi), it


@pytest.fixture(scope="module")
def batch():
    return batched_square_meshes(12, e_min=200, e_max=20000, seed=3, grid=64)


def test_batched_solve_vs_oracle(cuda, batch):
    print("case batched eager", flush=True)
    bp = BatchedPoisson([m[:3] for m in batch], cuda, rows_per_tile=256, lanes=4)
    kappa = torch.tensor(np.concatenate([m[3] for m in batch]), dtype=torch.float32, device=cuda)
    bp.run(kappa, rtol=1e-8)
    torch.cuda.synchronize()
    xs, status, iters, rel = bp.results()
    assert (status == 1).all(), status
    assert (rel <= 1e-8).all()
    for k, m in enumerate(batch):
        xr, it_ref = _oracle(m)
        x = xs[k].cpu().numpy()
        assert np.linalg.norm(x - xr) / np.linalg.norm(xr) < 1e-5
        assert abs(int(iters[k]) - it_ref) <= max(3, 0.05 * it_ref)


@pytest.mark.parametrize("lanes,rows", [(2, 128), (8, 1024)])
def test_batched_configs_agree(cuda, batch, lanes, rows):
    bp = BatchedPoisson([m[:3] for m in batch], cuda, rows_per_tile=rows, lanes=lanes)
    kappa = torch.tensor(np.concatenate([m[3] for m in batch]), dtype=torch.float32, device=cuda)
    bp.run(kappa, rtol=1e-8)
    torch.cuda.synchronize()
    xs, status, iters, rel = bp.results()
    assert (status == 1).all()
    bp2 = BatchedPoisson([m[:3] for m in batch], cuda, rows_per_tile=256, lanes=4)
    bp2.run(kappa, rtol=1e-8)
    torch.cuda.synchronize()
    xs2, _, it2, _ = bp2.results()
    for a, b in zip(xs, xs2):
        assert (a - b).norm() / b.norm() < 1e-6


def test_batched_graph_replay(cuda, batch):
    """The captured pipeline replays with new coefficients and matches eager runs."""
    print("case batched graph", flush=True)
    bp = BatchedPoisson([m[:3] for m in batch], cuda, rows_per_tile=256, lanes=4)
    k1 = torch.tensor(np.concatenate([m[3] for m in batch]), dtype=torch.float32, device=cuda)
    bp.kappa.copy_(k1)
    g = bp.capture(rtol=1e-8)
    for rep in range(3):
        kk = k1 * (1.0 + 0.5 * rep)
        bp.kappa.copy_(kk)
        g.replay()
        torch.cuda.synchronize()
        xs, status, iters, rel = bp.results()
        xs = [x.clone() for x in xs]
        eager = BatchedPoisson([m[:3] for m in batch], cuda, rows_per_tile=256, lanes=4)
        eager.run(kk, rtol=1e-8)
        torch.cuda.synchronize()
        xe, se, ie, _ = eager.results()
        assert (status == 1).all() and (se == 1).all()
        for a, b in zip(xs, xe):
            assert (a - b).norm() / b.norm() < 1e-6
        # kappa scaled by s scales the solution by 1/s
        if rep:
            assert abs(xs[0].norm().item() * (1.0 + 0.5 * rep) / first - 1.0) < 1e-6
        else:
            first = xs[0].norm().item()
