"""The shipped learned preconditioner (``precond/weights/poisson96_sym.npz``,
trained on plain-solver solutions of seeds 1000-2999) on held-out 96^3
log-normal Poisson systems, against Jacobi-PCG (the reference's Jacobi Krylov,
``sparse/solve/torch_solve.py:70-128``) and against the same network
untrained; and the Jacobi-PCG fallback of flexible CG, forced by an iteration
cap, by non-finite weights and by zero weights.  Convergence is judged by the
true FP64 residual ``||b - A x|| / ||b||`` recomputed here."""
import os

import pytest
import torch

from tensormesh_amd.precond import GNNPreconditioner, LearnedFCG
from tensormesh_amd.precond.problem import PoissonCube
from tensormesh_amd.precond.train import plain_solve
from tensormesh_amd.sparse.ops import _spmv_raw

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
WEIGHTS = os.path.join(ROOT, "tensormesh_amd", "precond", "weights", "poisson96_sym.npz")


@pytest.fixture(scope="module")
def cube96(cuda):
    print("case learned set-up 96^3", flush=True)
    net = GNNPreconditioner.load(WEIGHTS, cuda)
    prob = PoissonCube(96, cuda, n_levels=net.n_levels)
    return net, prob


def _true_relres(prob, values, x, b):
    r = b - _spmv_raw(prob.csr, values, x)
    return float(torch.linalg.vector_norm(r) / torch.linalg.vector_norm(b))


def _fcg(net, prob, kappa, values, b, **kw):
    solver = LearnedFCG(net, prob.hier, prob.csr, chunk=8, m=1)
    solver.set_system(values, prob.log_kappa_feature(kappa))
    return solver.solve(b, rtol=1e-8, **kw)


@pytest.mark.parametrize("seed", [5001, 5002, 5003])
def test_trained_network_beats_untrained_and_jacobi(cube96, cuda, seed):
    net, prob = cube96
    print(f"case learned held-out seed {seed}", flush=True)
    kappa = prob.kappa(seed)
    values, b = prob.system(kappa)
    x, info = _fcg(net, prob, kappa, values, b)
    assert info["fcg_exit"] == "converged" and info["fallback_iterations"] == 0, info["fcg_exit"]
    rel = _true_relres(prob, values, x, b)
    assert rel <= 1e-8 and abs(rel - info["relres"]) <= 1e-3 * rel
    xj, itj, _, okj = plain_solve(prob, values)
    assert okj and _true_relres(prob, values, xj, b) <= 1.01e-8
    net0 = GNNPreconditioner(net.schedule, net.n_linear, cuda, homogeneous=True, symmetric=True, bottom=net.bottom)
    x0, info0 = _fcg(net0, prob, kappa, values, b)
    assert info0["converged"]
    it, it0 = info["fcg_iterations"], info0["fcg_iterations"]
    print(f"seed {seed}: trained FCG {it}, untrained FCG {it0}, Jacobi-PCG {itj}", flush=True)
    assert 10 * it <= itj, (it, itj)
    assert 2 * it <= it0, (it, it0)
    # the same solution as Jacobi-PCG (both at true relres 1e-8)
    assert float(torch.linalg.vector_norm(x - xj) / torch.linalg.vector_norm(xj)) <= 1e-6


def test_fallback_on_iteration_cap(cube96, cuda):
    net, prob = cube96
    print("case learned fallback maxiter", flush=True)
    kappa = prob.kappa(5004)
    values, b = prob.system(kappa)
    x, info = _fcg(net, prob, kappa, values, b, maxiter=3)
    assert info["fcg_exit"] == "max iterations" and info["fallback_iterations"] > 0
    assert info["converged"] and _true_relres(prob, values, x, b) <= 1.01e-8


@pytest.mark.parametrize("fill", [float("nan"), 0.0])
def test_fallback_on_broken_network(cube96, cuda, fill):
    """Non-finite weights give a non-finite correction, zero weights a zero
    correction: flexible CG breaks down (or stagnates) and Jacobi-PCG
    finishes from the last good iterate."""
    net, prob = cube96
    print(f"case learned fallback weights={fill}", flush=True)
    bad = GNNPreconditioner.load(WEIGHTS, cuda)
    bad.params.fill_(fill)
    bad.refresh()
    kappa = prob.kappa(5005)
    values, b = prob.system(kappa)
    x, info = _fcg(bad, prob, kappa, values, b)
    assert info["fcg_exit"] in ("breakdown", "stagnation"), info["fcg_exit"]
    assert info["fallback_iterations"] > 0 and info["converged"]
    assert torch.isfinite(x).all()
    assert _true_relres(prob, values, x, b) <= 1.01e-8
