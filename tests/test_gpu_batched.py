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


def test_autotuner_picks_fastest_and_caches(cuda, batch, tmp_path, monkeypatch):
    """``autotune_batched`` times every candidate (including two workgroups per
    CU), applies the fastest, solves correctly with it, and a second call is a
    cache hit with the same configuration."""
    print("case batched autotune", flush=True)
    from tensormesh_amd import tune
    monkeypatch.setenv("TM_TUNE_CACHE", str(tmp_path / "tune.json"))
    bp = BatchedPoisson([m[:3] for m in batch], cuda, rows_per_tile=256, lanes=4)
    bp.kappa.copy_(torch.tensor(np.concatenate([m[3] for m in batch]), dtype=torch.float32, device=cuda))
    space = {"rows_per_tile": (256, 1024), "lanes": (2, 4), "blocks_per_cu": (1, 2)}
    cfg = tune.autotune_batched(bp, space=space, use_cache=True, maxiter=10000)
    assert cfg["cached"] is False and len(cfg["trials"]) == 8
    finite = [t for t in cfg["trials"] if np.isfinite(t["seconds"])]
    assert finite and cfg["seconds"] == min(t["seconds"] for t in finite)
    ncu = torch.cuda.get_device_properties(cuda).multi_processor_count
    assert {t["blocks"] for t in cfg["trials"]} == {ncu, 2 * ncu}
    assert (bp.rows_per_tile, bp.lanes, bp.blocks) == (cfg["rows_per_tile"], cfg["lanes"], cfg["blocks"])
    bp.run(None, rtol=1e-8)
    torch.cuda.synchronize()
    xs, status, iters, rel = bp.results()
    assert (status == 1).all() and (rel <= 1e-8).all()
    for k in (0, len(batch) - 1):
        xr, _ = _oracle(batch[k])
        assert np.linalg.norm(xs[k].cpu().numpy() - xr) / np.linalg.norm(xr) < 1e-5
    bp2 = BatchedPoisson([m[:3] for m in batch], cuda, rows_per_tile=256, lanes=4)
    cfg2 = tune.autotune_batched(bp2, space=space, use_cache=True)
    assert cfg2["cached"] is True
    assert (bp2.rows_per_tile, bp2.lanes, bp2.blocks) == (cfg["rows_per_tile"], cfg["lanes"], cfg["blocks"])


def test_batched_many_larger_meshes(cuda):
    """64 uneven meshes of up to 1e5 elements in one launch: every system
    converges to 1e-8, and sampled systems match the oracle's direct solve."""
    print("case batched 64 meshes", flush=True)
    ms = batched_square_meshes(64, e_min=2000, e_max=100000, seed=7)
    bp = BatchedPoisson([m[:3] for m in ms], cuda, rows_per_tile=2048, lanes=4)
    bp.run(torch.tensor(np.concatenate([m[3] for m in ms]), dtype=torch.float32, device=cuda), rtol=1e-8)
    torch.cuda.synchronize()
    xs, status, iters, rel = bp.results()
    assert (status == 1).all() and (rel <= 1e-8).all()
    sizes = [len(m[1]) for m in ms]
    for k in (int(np.argmin(sizes)), int(np.argmax(sizes)), 31):
        xr, it_ref = _oracle(ms[k])
        # relres 1e-8 bounds the error by cond(A) * 1e-8, and cond(A) grows like
        # the element count: 1e-4 at 1e5 elements
        assert np.linalg.norm(xs[k].cpu().numpy() - xr) / np.linalg.norm(xr) < 1e-9 * sizes[k]
        assert abs(int(iters[k]) - it_ref) <= max(3, 0.05 * it_ref)
