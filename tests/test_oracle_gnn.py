"""CPU checks of the learned-preconditioner oracle and of the host hierarchy."""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

from oracle import assemble as oa
from oracle import gnn as og
from oracle import pattern as op
from oracle import solve as osv
from tensormesh_amd.mesh import kuhn_cube_arrays
from tensormesh_amd.precond.hierarchy import build_host_levels
from tensormesh_amd.precond.network import GNNPreconditioner


def _poisson(n, seed=None):
    pts, conn, bnd = kuhn_cube_arrays(n)
    eu, ev, pe = op.pattern_and_slots(conn, len(pts))
    if seed is None:
        ke = oa.element_matrices(pts, conn, "tetra", "laplace")
    else:
        kappa = np.exp(np.random.default_rng(seed).standard_normal(len(pts)))
        ke = oa.element_matrices(pts, conn, "tetra", "kappa_laplace", kappa=kappa)
    A = sp.csr_matrix((oa.reduce(pe, ke, len(eu)), (eu, ev)), shape=(len(pts),) * 2)
    inner = ~bnd
    Ai = A[inner][:, inner].tocsr()
    Ai.sort_indices()
    return pts[inner], Ai


def test_bf16_round_matches_torch():
    x = np.random.default_rng(0).standard_normal(10000) * np.exp(np.random.default_rng(1).uniform(-20, 20, 10000))
    x = np.concatenate([x, [0.0, 1.0, -1.0, 1.00390625, 1.005859375, 3.0e38]])
    ref = torch.from_numpy(x.astype(np.float32)).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(og.bf16_round(x), ref)


def test_hierarchy_nested_and_galerkin():
    coords, Ai = _poisson(6, seed=3)
    H = build_host_levels(coords, Ai.indptr, Ai.indices, 4)
    assert [h.n for h in H] == [125, 27, 8, 1]  # 5^3 interior nodes, 2:1 per axis
    levels = og.galerkin_levels(Ai, [h.agg for h in H[:-1]])
    for k in range(1, len(H)):
        h, prev = H[k], H[k - 1]
        # children lists partition the previous level and agree with agg
        assert sorted(h.child.tolist()) == list(range(prev.n))
        for I in range(h.n):
            kids = h.child[h.child_ptr[I]:h.child_ptr[I + 1]]
            assert np.all(prev.agg[kids] == I)
        # coarse pattern = structural pattern of P^T A P (all-ones values, so no
        # entry cancels), and the perm/seg sums of the fine values reproduce P^T A P
        P = levels[k - 1]["P"]
        S = sp.csr_matrix((np.ones(prev.col.size), prev.col, prev.row_ptr), shape=(prev.n, prev.n))
        Sc = (P.T @ S @ P).tocsr()
        Sc.sort_indices()
        assert np.array_equal(h.row_ptr, Sc.indptr) and np.array_equal(h.col, Sc.indices)
        fine = sp.csr_matrix(levels[k - 1]["A"])
        fine_vals = np.asarray(fine[np.repeat(np.arange(prev.n), np.diff(prev.row_ptr)), prev.col]).ravel()
        sums = np.add.reduceat(fine_vals[h.perm], h.seg[:-1])
        Ad = levels[k]["A"].toarray()
        rows = np.repeat(np.arange(h.n), np.diff(h.row_ptr))
        assert np.allclose(sums, Ad[rows, h.col], rtol=1e-12, atol=1e-14)
    # dense P^T A P on level 1
    P = levels[0]["P"].toarray()
    assert np.allclose(P.T @ Ai.toarray() @ P, levels[1]["A"].toarray())
    for L in levels:
        assert np.allclose(L["An"].diagonal(), 1.0)


def test_oracle_backward_matches_finite_differences():
    coords, Ai = _poisson(6, seed=4)
    H = build_host_levels(coords, Ai.indptr, Ai.indices, 4)
    levels = og.galerkin_levels(Ai, [h.agg for h in H[:-1]])
    net = GNNPreconditioner((0, 1, 2, 3, 1, 0), n_linear=32, device=torch.device("cpu"), homogeneous=False,
                            symmetric=False)
    rng = np.random.default_rng(1)
    p = net.params.numpy().astype(np.float64) + rng.standard_normal(net.n_params) * 0.05
    r = rng.standard_normal(Ai.shape[0])
    lk = rng.standard_normal(Ai.shape[0]) * 0.3
    z, y, s, tape = og.forward(levels, p, net.schedule, 32, r, lk)
    g = og.backward(levels, p, net.schedule, 32, r, lk, tape, s)
    idx = np.concatenate([rng.choice(net.n_params, 16, replace=False),
                          [0, 130, 197, 192 + 8192 + 3, net.n_params - 1, net.n_params - 30]])
    h = 1e-6
    gmax = np.abs(g).max()
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (og.loss(levels, og.forward(levels, pp, net.schedule, 32, r, lk)[1], r, s)
              - og.loss(levels, og.forward(levels, pm, net.schedule, 32, r, lk)[1], r, s)) / (2 * h)
        assert abs(g[i] - fd) <= 1e-6 * abs(fd) + 1e-7 * gmax, (i, g[i], fd)


def test_untrained_network_is_additive_multilevel():
    coords, Ai = _poisson(6)
    H = build_host_levels(coords, Ai.indptr, Ai.indices, 4)
    levels = og.galerkin_levels(Ai, [h.agg for h in H[:-1]])
    net = GNNPreconditioner(device=torch.device("cpu"))
    net.init_weights(0, noise=0.0)
    r = np.random.default_rng(0).standard_normal(Ai.shape[0])
    z, y, s, _ = og.forward(levels, net.params.numpy(), net.schedule, net.n_linear, r, np.zeros_like(r))
    x0 = levels[0]["d"] * r / s
    Ph = [sp.diags(levels[k]["w"]) @ levels[k]["P"] for k in range(3)]
    I = lambda m: sp.eye(m.shape[1])
    M = sp.eye(len(r)) + Ph[0] @ (I(Ph[0]) + Ph[1] @ (I(Ph[1]) + Ph[2] @ Ph[2].T) @ Ph[1].T) @ Ph[0].T
    # every channel carries the same multilevel sum; the decoder weighs them by we . wd
    scale = float(net.enc()[0][:64].double() @ net.dec()[0].double())
    assert np.allclose(y, scale * (M @ x0), rtol=1e-10, atol=1e-12)
    # symmetric positive definite operator r -> z
    Z = np.stack([og.forward(levels, net.params.numpy(), net.schedule, net.n_linear, e, np.zeros_like(r))[0]
                  / og.forward(levels, net.params.numpy(), net.schedule, net.n_linear, e, np.zeros_like(r))[2]
                  for e in np.eye(len(r))[:40]])
    assert np.allclose(Z[:, :40], Z[:, :40].T, atol=1e-12)


def _operator(levels, net, n):
    """Dense matrix of r -> z (the network is homogeneous and linear in r)."""
    cols = []
    for e in np.eye(n):
        z, _, s, _ = og.forward(levels, net.params.double().numpy(), net.schedule, net.n_linear, e, np.zeros(n))
        cols.append(z)
    return np.stack(cols, 1)


def test_symmetric_network_is_spd_for_any_weights():
    """Tied adjoint layers + tied decoder: M = sum R_l^T R_l is symmetric
    positive definite for arbitrary (random) weights, and the gradient folding
    keeps the tie under training."""
    coords, Ai = _poisson(6, seed=6)
    H = build_host_levels(coords, Ai.indptr, Ai.indices, 4)
    levels = og.galerkin_levels(Ai, [h.agg for h in H[:-1]])
    net = GNNPreconditioner(device=torch.device("cpu"))
    g = torch.Generator().manual_seed(3)
    net.params.add_(torch.randn(net.n_params, generator=g) * 0.3 * net.mask)
    net.refresh()
    M = _operator(levels, net, Ai.shape[0])
    assert np.abs(M - M.T).max() <= 1e-10 * np.abs(M).max()
    assert np.linalg.eigvalsh(0.5 * (M + M.T)).min() > 0
    # a non-symmetric network with the same schedule is not symmetric
    ns = GNNPreconditioner(device=torch.device("cpu"), symmetric=False)
    ns.params.add_(torch.randn(ns.n_params, generator=g) * 0.3 * ns.mask)
    ns.refresh()
    Mn = _operator(levels, ns, Ai.shape[0])
    assert np.abs(Mn - Mn.T).max() > 1e-3 * np.abs(Mn).max()
    # folding: the oracle gradient of the tied parameters equals the sum over their uses
    r = np.random.default_rng(2).standard_normal(Ai.shape[0])
    p = net.params.double().numpy()
    z, y, s, tape = og.forward(levels, p, net.schedule, net.n_linear, r, np.zeros_like(r))
    gfull = torch.tensor(og.backward(levels, p, net.schedule, net.n_linear, r, np.zeros_like(r), tape, s))
    net.fold(gfull)
    i = int(net.offsets[1]) + 5 * 64 + 7          # a free weight of layer 0
    h = 1e-6
    loss = []
    for sgn in (1, -1):
        q = net.params.double().clone()
        q[i] += sgn * h
        t = GNNPreconditioner(device=torch.device("cpu"))
        t.params.copy_(q.float())
        t.params.data = q                                        # keep float64 for the difference
        t.refresh()
        pp = t.params.numpy()
        loss.append(og.loss(levels, og.forward(levels, pp, t.schedule, t.n_linear, r, np.zeros_like(r))[1], r, s))
    fd = (loss[0] - loss[1]) / (2 * h)
    assert abs(float(gfull[i]) - fd) <= 1e-5 * max(abs(fd), 1e-3), (float(gfull[i]), fd)


def test_fcg_reduces_to_pcg_and_direct():
    coords, Ai = _poisson(6, seed=5)
    b = np.ones(Ai.shape[0])
    dinv = 1.0 / Ai.diagonal()
    x, it, hist = og.fcg(Ai, b, lambda r: dinv * r, rtol=1e-10)
    xr, itr, relr = osv.jacobi_pcg(Ai, b, rtol=1e-10)
    assert abs(it - itr) <= 1 and np.allclose(x, xr, rtol=1e-8, atol=1e-12)
    Ad = np.linalg.inv(Ai.toarray())
    x1, it1, _ = og.fcg(Ai, b, lambda r: Ad @ r, rtol=1e-10)
    assert it1 == 1


def test_cosine_loss_gradient_matches_finite_differences():
    coords, Ai = _poisson(6, seed=8)
    H = build_host_levels(coords, Ai.indptr, Ai.indices, 4)
    levels = og.galerkin_levels(Ai, [h.agg for h in H[:-1]])
    net = GNNPreconditioner(device=torch.device("cpu"))
    rng = np.random.default_rng(4)
    p = net.params.numpy().astype(np.float64) + rng.standard_normal(net.n_params) * 0.05 * net.mask.numpy()
    e = rng.standard_normal(Ai.shape[0])
    r = Ai @ e
    lk = np.zeros_like(r)
    z, y, s, tape = og.forward(levels, p, net.schedule, net.n_linear, r, lk)
    g = og.backward(levels, p, net.schedule, net.n_linear, r, lk, tape, s,
                    gy=og.cosine_grad_y(levels, y, r, s, e))
    L0 = og.cosine_loss(levels, y, r, s, e)
    assert 0.0 <= L0 <= 1.0
    h = 1e-6
    gmax = np.abs(g).max()
    for i in np.concatenate([rng.choice(np.nonzero(net.mask.numpy())[0], 12, replace=False), [0, 5]]):
        vals = []
        for sgn in (1, -1):
            q = p.copy()
            q[i] += sgn * h
            vals.append(og.cosine_loss(levels, og.forward(levels, q, net.schedule, net.n_linear, r, lk)[1], r, s, e))
        fd = (vals[0] - vals[1]) / (2 * h)
        assert abs(g[i] - fd) <= 1e-5 * abs(fd) + 1e-7 * gmax, (i, g[i], fd)
