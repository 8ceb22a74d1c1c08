"""GPU parity of the learned-preconditioner kernels (levels, network forward and
backward, flexible CG, training) against the float64 oracle."""
import ctypes

import numpy as np
import pytest
import scipy.sparse as sp
import torch

from oracle import assemble as oa
from oracle import gnn as og
from oracle import pattern as op
from tensormesh_amd import _lib
from tensormesh_amd.mesh import kuhn_cube_arrays
from tensormesh_amd.precond import Adam, GNNPreconditioner, Hierarchy, LearnedFCG, build_host_levels
from tensormesh_amd.sparse.ops import CSR, pcg_raw

pytestmark = pytest.mark.gpu


def _problem(n, seed, cuda, sigma=1.0):
    pts, conn, bnd = kuhn_cube_arrays(n)
    eu, ev, pe = op.pattern_and_slots(conn, len(pts))
    kappa = np.exp(sigma * np.random.default_rng(seed).standard_normal(len(pts)))
    ke = oa.element_matrices(pts, conn, "tetra", "kappa_laplace", kappa=kappa)
    A = sp.csr_matrix((oa.reduce(pe, ke, len(eu)), (eu, ev)), shape=(len(pts),) * 2)
    inner = ~bnd
    Ai = A[inner][:, inner].tocsr()
    Ai.sort_indices()
    Ai.data = Ai.data.astype(np.float32).astype(np.float64)  # the device holds FP32 values
    H = build_host_levels(pts[inner], Ai.indptr, Ai.indices, 4)
    hier = Hierarchy(H, cuda)
    values = torch.tensor(Ai.data, dtype=torch.float32, device=cuda)
    hier.set_matrix(values)
    levels = og.galerkin_levels(Ai, [h.agg for h in H[:-1]])
    csr = CSR(torch.tensor(Ai.indptr, dtype=torch.int32, device=cuda),
              torch.tensor(Ai.indices, dtype=torch.int32, device=cuda), Ai.shape[0], Ai.shape[0])
    lk = np.log(kappa[inner])
    lk = ((lk - lk.mean()) / 2.0).astype(np.float32)
    return Ai, H, hier, levels, values, csr, lk


def _perturbed(cuda, seed=7, scale=0.05, schedule=(1, 2, 3, 3, 2, 1)):
    """General (non-homogeneous, non-symmetric) network: ReLU channels, biases
    and the log-coefficient input all active."""
    net = GNNPreconditioner(schedule, device=cuda, n_linear=32, homogeneous=False, symmetric=False)
    g = torch.Generator(device="cpu").manual_seed(seed)
    net.params.add_((torch.randn(net.n_params, generator=g) * scale).to(cuda))
    net.refresh()
    return net


def _params64(net):
    """Parameters as the kernels see them: layer weights rounded to BF16."""
    return net.params.double().cpu().numpy()


def test_level_setup(cuda):
    print("case precond levels n=8", flush=True)
    Ai, H, hier, levels, values, csr, lk = _problem(8, 1, cuda)
    for k, lev in enumerate(hier.levels):
        ref = levels[k]
        assert np.allclose(lev.d.cpu().numpy(), ref["d"], rtol=1e-6)
        rows = np.repeat(np.arange(lev.n), np.diff(H[k].row_ptr))
        An = ref["An"].toarray()[rows, H[k].col]
        assert np.abs(lev.vn.cpu().numpy() - An).max() <= 1e-5 * np.abs(An).max()
        if k + 1 < len(hier.levels):
            assert np.allclose(lev.w.cpu().numpy(), ref["w"], rtol=1e-5)


def test_mp_layer_forward(cuda):
    """One fused message-passing layer against [h | A^h] W + b with BF16 rounding."""
    print("case precond mp layer n=8", flush=True)
    Ai, H, hier, levels, values, csr, lk = _problem(8, 2, cuda)
    lev = hier.levels[0]
    g = np.random.default_rng(3)
    h = og.bf16_round(g.standard_normal((lev.n, 64)))
    W = og.bf16_round(g.standard_normal((128, 64)) / 8.0)
    b = g.standard_normal(64).astype(np.float32)
    ht = torch.tensor(h, dtype=torch.bfloat16, device=cuda)
    wt = torch.tensor(W, dtype=torch.bfloat16, device=cuda).t().contiguous()
    bt = torch.tensor(b, device=cuda)
    out = torch.empty_like(ht)
    msg = torch.empty_like(ht)
    pre = torch.empty_like(ht)
    L = _lib.lib()
    _lib.check(L.tm_gnn_mp_forward(lev.n, _lib.ptr(lev.row_ptr), _lib.ptr(lev.col), _lib.ptr(lev.vn), _lib.ptr(ht),
                                   _lib.ptr(wt), _lib.ptr(bt), 32, _lib.ptr(out), _lib.ptr(msg), _lib.ptr(pre),
                                   None, _lib.stream_of(cuda)))
    m_ref = og.bf16_round(levels[0]["An"] @ h)
    assert np.abs(msg.double().cpu().numpy() - m_ref).max() <= 2 ** -7 * np.abs(m_ref).max()
    m = msg.double().cpu().numpy()  # use the device's rounded messages for the GEMM check
    u = np.concatenate([h, m], 1) @ W + b
    o_ref = h + np.concatenate([u[:, :32], np.maximum(u[:, 32:], 0)], 1)
    got = out.double().cpu().numpy()
    assert np.abs(got - o_ref).max() <= 2 ** -7 * np.abs(o_ref).max()
    assert np.linalg.norm(got - o_ref) <= 4e-3 * np.linalg.norm(o_ref)
    assert np.abs(pre.double().cpu().numpy() - u).max() <= 2 ** -7 * np.abs(u).max()


def test_small_level_layer_matches_tiled_layer(cuda):
    """Inference on levels under 65536 rows runs 16-row tiles (k_mp_small); the
    training path (activations saved) runs the 64-row tiles: same bits."""
    print("case precond mp small-vs-tiled n=8", flush=True)
    Ai, H, hier, levels, values, csr, lk = _problem(8, 2, cuda)
    L = _lib.lib()
    g = np.random.default_rng(5)
    for lev in hier.levels[:2]:
        assert lev.n < 65536
        ht = torch.tensor(og.bf16_round(g.standard_normal((lev.n, 64))), dtype=torch.bfloat16, device=cuda)
        wt = torch.tensor(og.bf16_round(g.standard_normal((128, 64)) / 8.0), dtype=torch.bfloat16,
                          device=cuda).t().contiguous()
        bt = torch.tensor(g.standard_normal(64).astype(np.float32), device=cuda)
        outs = []
        for save in (True, False):
            out = torch.empty_like(ht)
            msg, pre = (torch.empty_like(ht), torch.empty_like(ht)) if save else (None, None)
            _lib.check(L.tm_gnn_mp_forward(lev.n, _lib.ptr(lev.row_ptr), _lib.ptr(lev.col), _lib.ptr(lev.vn),
                                           _lib.ptr(ht), _lib.ptr(wt), _lib.ptr(bt), 32, _lib.ptr(out),
                                           _lib.ptr(msg), _lib.ptr(pre), None, _lib.stream_of(cuda)))
            outs.append(out)
        assert torch.equal(outs[0].view(torch.int16), outs[1].view(torch.int16)), f"level with {lev.n} rows"


def test_network_forward(cuda):
    print("case precond forward n=8", flush=True)
    Ai, H, hier, levels, values, csr, lk = _problem(8, 3, cuda)
    net = _perturbed(cuda)
    prog = net.plan(hier)
    r = np.random.default_rng(4).standard_normal(Ai.shape[0])
    z = prog.forward(torch.tensor(r, device=cuda), torch.tensor(lk, device=cuda)).cpu().numpy()
    z_ref = og.forward(levels, _params64(net), net.schedule, net.n_linear, r, lk.astype(np.float64), bf16=True)[0]
    assert np.linalg.norm(z - z_ref) <= 2e-2 * np.linalg.norm(z_ref)


@pytest.mark.parametrize("n,schedule", [(8, (1, 2, 3, 2, 1, 0)), (13, (1, 2, 3, 2, 1, 0)), (9, (1, 2, 3, 3, 2, 1))])
def test_fused_inference_matches_unfused(cuda, n, schedule):
    """Fused finest level (encode+pool, unpool+layer+decode with B = A^ P^) =
    unfused kernels, up to the BF16 rounding of the level-0 features that only
    the unfused path stores."""
    print(f"case precond fused-vs-unfused n={n}", flush=True)
    Ai, H, hier, levels, values, csr, lk = _problem(n, 10 + n, cuda)
    net = _perturbed(cuda, seed=n, schedule=schedule)
    prog = net.plan(hier)
    assert prog.fusable or prog.fusable_plain
    r = torch.tensor(np.random.default_rng(n).standard_normal(Ai.shape[0]), device=cuda)
    lkt = torch.tensor(lk, device=cuda)
    zf = prog.forward(r, lkt, fused=True).clone()
    zu = prog.forward(r, lkt, fused=False).clone()
    assert torch.linalg.vector_norm(zf - zu) <= 1e-2 * torch.linalg.vector_norm(zu)
    z_ref = og.forward(levels, _params64(net), net.schedule, net.n_linear, r.cpu().numpy(), lk.astype(np.float64), bf16=True)[0]
    assert np.linalg.norm(zf.cpu().numpy() - z_ref) <= 2e-2 * np.linalg.norm(z_ref)


@pytest.mark.parametrize("n", [9, 12])
def test_fused_symmetric_level0_matches_unfused(cuda, n):
    """Symmetric network with a level-0 layer pair: the rank-two fused finest
    level (two scalar CSR products) = the unfused 64-wide level-0 layers, and
    both match the oracle."""
    print(f"case precond fused sym0 n={n}", flush=True)
    Ai, H, hier, levels, values, csr, lk = _problem(n, 30 + n, cuda)
    net = GNNPreconditioner((0, 1, 2, 2, 1, 0), device=cuda, bottom=3)
    net.params.add_(torch.randn(net.n_params, generator=torch.Generator().manual_seed(n)).to(cuda) * 0.1 * net.mask)
    net.refresh()
    prog = net.plan(hier)
    assert prog.fusable_sym0
    r = torch.tensor(np.random.default_rng(n).standard_normal(Ai.shape[0]), device=cuda)
    lkt = torch.tensor(lk, device=cuda)
    zf = prog.forward(r, lkt, fused=True).clone()
    zu = prog.forward(r, lkt, fused=False).clone()
    assert torch.linalg.vector_norm(zf - zu) <= 1e-2 * torch.linalg.vector_norm(zu)
    z_ref = og.forward(levels, _params64(net), net.schedule, net.n_linear, r.cpu().numpy(), lk.astype(np.float64),
                       bf16=True, bottom=net.bottom)[0]
    assert np.linalg.norm(zf.cpu().numpy() - z_ref) <= 2e-2 * np.linalg.norm(z_ref)


def test_network_gradient(cuda):
    """Parameter gradient of the energy loss: hand-written backward kernels vs
    the oracle backward (itself pinned by finite differences)."""
    print("case precond gradient n=8", flush=True)
    Ai, H, hier, levels, values, csr, lk = _problem(8, 4, cuda)
    net = _perturbed(cuda, seed=8)
    prog = net.plan(hier)
    r = np.random.default_rng(5).standard_normal(Ai.shape[0])
    grad = torch.zeros_like(net.params)
    prog.loss_and_grad(torch.tensor(r, device=cuda), torch.tensor(lk, device=cuda), grad)
    p = _params64(net)
    z, y, s, tape = og.forward(levels, p, net.schedule, net.n_linear, r, lk.astype(np.float64), bf16=True)
    g_ref = og.backward(levels, p, net.schedule, net.n_linear, r, lk.astype(np.float64), tape, s, bf16=True)
    g = grad.double().cpu().numpy()
    for k in range(len(net.offsets) - 1):
        a, b = int(net.offsets[k]), int(net.offsets[k + 1])
        err = np.linalg.norm(g[a:b] - g_ref[a:b]) / max(np.linalg.norm(g_ref[a:b]), 1e-30)
        assert err <= 5e-2, (k, err)
    assert np.linalg.norm(g - g_ref) <= 2e-2 * np.linalg.norm(g_ref)


def test_cosine_loss_gradient(cuda):
    """Scale-invariant loss: device gradient vs the oracle backward."""
    print("case precond cosine gradient n=8", flush=True)
    Ai, H, hier, levels, values, csr, lk = _problem(8, 11, cuda)
    net = GNNPreconditioner(device=cuda)
    net.params.add_(torch.randn(net.n_params, generator=torch.Generator().manual_seed(5)).to(cuda) * 0.05 * net.mask)
    net.refresh()
    prog = net.plan(hier)
    e = np.random.default_rng(6).standard_normal(Ai.shape[0])
    r = Ai @ e
    zl = np.zeros_like(r)
    grad = torch.zeros_like(net.params)
    prog.loss_and_grad(torch.tensor(r, device=cuda), torch.zeros(len(r), dtype=torch.float32, device=cuda), grad,
                       torch.tensor(e, device=cuda), loss="cosine")
    p = _params64(net)
    z, y, s, tape = og.forward(levels, p, net.schedule, net.n_linear, r, zl, bf16=True)
    g_ref = og.backward(levels, p, net.schedule, net.n_linear, r, zl, tape, s, bf16=True,
                        gy=og.cosine_grad_y(levels, y, r, s, e))
    g = grad.double().cpu().numpy()
    assert np.linalg.norm(g - g_ref) <= 3e-2 * np.linalg.norm(g_ref)


@pytest.mark.parametrize("graph", [True, False])
def test_fcg_untrained_additive_multilevel(cuda, graph):
    """The untrained network is an SPD additive multilevel preconditioner: FCG
    must converge (true residual 1e-8) in fewer iterations than Jacobi-PCG
    and match the oracle FCG with the oracle network."""
    print(f"case precond fcg n=16 graph={graph}", flush=True)
    Ai, H, hier, levels, values, csr, lk = _problem(16, 5, cuda)
    net = GNNPreconditioner(device=cuda)
    solver = LearnedFCG(net, hier, csr, chunk=4, use_graph=graph)
    solver.set_system(values, torch.tensor(lk, device=cuda))
    b = torch.ones(Ai.shape[0], dtype=torch.float64, device=cuda)
    x, info = solver.solve(b, rtol=1e-8, maxiter=500)
    assert info["converged"] and info["fcg_exit"] == "converged" and info["fallback_iterations"] == 0, info
    xr = x.cpu().numpy()
    assert np.linalg.norm(np.ones(len(xr)) - Ai @ xr) <= 1.0001e-8 * np.sqrt(len(xr))
    _, itj, _, _ = pcg_raw(csr, values, b, rtol=1e-8)
    assert info["fcg_iterations"] < itj, (info["fcg_iterations"], itj)
    p = _params64(net)
    pre = lambda r: og.forward(levels, p, net.schedule, net.n_linear, r, lk.astype(np.float64), bf16=True)[0]
    _, it_ref, _ = og.fcg(Ai, np.ones(len(xr)), pre, rtol=1e-8, m=solver.m)
    assert abs(info["fcg_iterations"] - it_ref) <= max(2, 0.1 * it_ref), (info["fcg_iterations"], it_ref)


@pytest.mark.parametrize("m", [1, 2, 8])
def test_fcg_ring_sizes_vs_oracle(cuda, m):
    """FCG(m) kernels (direction ring, single reduction phase) against the
    oracle's explicitly orthogonalising FCG(m) with the same network."""
    print(f"case precond fcg m={m} n=12", flush=True)
    Ai, H, hier, levels, values, csr, lk = _problem(12, 20 + m, cuda)
    net = _perturbed(cuda, seed=m, scale=0.02)
    solver = LearnedFCG(net, hier, csr, chunk=2, use_graph=False, m=m)
    solver.set_system(values, torch.tensor(lk, device=cuda))
    b = torch.ones(Ai.shape[0], dtype=torch.float64, device=cuda)
    x, info = solver.solve(b, rtol=1e-8, maxiter=400)
    assert info["converged"], info
    p = _params64(net)
    pre = lambda r: og.forward(levels, p, net.schedule, net.n_linear, r, lk.astype(np.float64), bf16=True)[0]
    _, it_ref, _ = og.fcg(Ai, np.ones(Ai.shape[0]), pre, rtol=1e-8, m=m)
    assert abs(info["fcg_iterations"] + info["fallback_iterations"] - it_ref) <= max(3, 0.15 * it_ref), \
        (info, it_ref)


def test_fcg_graph_replay_matches_eager(cuda):
    print("case precond fcg graph-vs-eager n=12", flush=True)
    Ai, H, hier, levels, values, csr, lk = _problem(12, 6, cuda)
    net = _perturbed(cuda, seed=9, scale=0.02)
    b = torch.ones(Ai.shape[0], dtype=torch.float64, device=cuda)
    outs = []
    for graph in (False, True, True):
        solver = LearnedFCG(net, hier, csr, chunk=3, use_graph=graph)
        solver.set_system(values, torch.tensor(lk, device=cuda))
        outs.append(solver.solve(b, rtol=1e-8, maxiter=400))
    for x, info in outs[1:]:
        assert info["fcg_iterations"] == outs[0][1]["fcg_iterations"]
        assert torch.equal(x, outs[0][0])


def test_training_lowers_energy_error(cuda):
    """A few Adam steps on one system lower the relative energy error of the
    correction for residuals of the plain solver."""
    print("case precond training n=10", flush=True)
    Ai, H, hier, levels, values, csr, lk = _problem(10, 7, cuda, sigma=1.5)
    net = GNNPreconditioner(device=cuda)
    prog = net.plan(hier)
    opt = Adam(net, lr=2e-3)
    b = torch.ones(Ai.shape[0], dtype=torch.float64, device=cuda)
    xs, _, _, _ = pcg_raw(csr, values, b, rtol=1e-10)
    data = []
    for k in (0, 3, 10):
        xk = torch.zeros_like(xs) if k == 0 else pcg_raw(csr, values, b, rtol=1e-30, maxiter=k)[0]
        rk = b - torch.tensor(Ai @ xk.cpu().numpy(), device=cuda)
        data.append((rk, xs - xk))
    lkt = torch.tensor(lk, device=cuda)
    d0 = hier.levels[0].d.double()
    n = Ai.shape[0]
    grad = torch.zeros_like(net.params)

    def epoch(step):
        grad.zero_()
        errs = []
        for r, e in data:
            er = torch.dot(e, r) / torch.mean((d0 * r) ** 2)
            S = prog.loss_and_grad(r, lkt, grad, e, gy_scale=(2.0 * n) / er)
            errs.append(((S[0] - 2 * S[1] + S[2]) / S[2]).item())
        if step:
            opt.step(grad)
        return float(np.mean(errs))

    e0 = epoch(True)
    for _ in range(40):
        e1 = epoch(True)
    e_end = epoch(False)
    assert e_end < 0.8 * e0, (e0, e_end)
