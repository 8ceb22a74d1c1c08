"""Float64 CPU statement of the learned preconditioner and flexible CG (test
infrastructure only).

The network has no counterpart in the reference; this module is the
definition the HIP kernels are checked against:

* levels: normalised operators ``A^_l = D_l^-1/2 A_l D_l^-1/2`` with
  ``A_{l+1} = P_l^T A_l P_l`` (``P_l`` the 0/1 aggregation matrix), formed here
  with SciPy sparse products;
* pooling weights ``w_i = sqrt(A_ii / Ac_II)``; pool ``h_c = P^T (w h)``,
  unpool ``h = skip + w (P h_c)``;
* message passing ``h' = h + act([h | A^ h] W + b)``, act = identity on the
  first ``n_lin`` channels and ReLU on the rest;
* linear encoder ``[d r / s, lk] We + be`` with ``s = rms(d r)``, decoder
  ``y = h wd + bd``, output ``z = s d y``;
* loss ``L = (y^T A^ y / 2 - y^T r~) / n`` with ``r~ = d r / s``.

``bf16=True`` rounds every stored activation (features, messages) and the
layer weights to bfloat16 (round to nearest even) exactly where the kernels
store BF16, so forward outputs can be compared tightly; the backward pass is
plain float64 on those values.

Pinning: the backward pass is checked against central finite differences of
the float64 forward (tests/test_oracle.py), the Galerkin levels against
explicit dense ``P^T A P``, and ``fcg`` against the closed form that with the
exact inverse as preconditioner it converges in one iteration and with the
Jacobi preconditioner it reproduces PCG's iterates (FCG with a fixed SPD
preconditioner is PCG: Notay, SIAM J. Sci. Comput. 22, 2000).
"""
from __future__ import annotations

from typing import Callable, Dict, List, Sequence

import numpy as np
import scipy.sparse as sp

F = 64


def bf16_round(a: np.ndarray) -> np.ndarray:
    """Round float values to bfloat16 (nearest even), returned as float64."""
    x = np.asarray(a, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def galerkin_levels(A0: sp.csr_matrix, aggs: Sequence[np.ndarray]) -> List[Dict]:
    """Per level: A (float64 CSR), d = diag^-1/2, An (normalised), P to the next level, w."""
    levels = []
    A = sp.csr_matrix(A0, dtype=np.float64)
    for k in range(len(aggs) + 1):
        dg = A.diagonal()
        d = np.where(dg > 0, 1.0 / np.sqrt(np.where(dg > 0, dg, 1.0)), 1.0)
        An = sp.diags(d) @ A @ sp.diags(d)
        levels.append({"A": A, "d": d, "An": sp.csr_matrix(An)})
        if k < len(aggs):
            agg = np.asarray(aggs[k], dtype=np.int64)
            nc = int(agg.max()) + 1
            P = sp.csr_matrix((np.ones(len(agg)), (np.arange(len(agg)), agg)), shape=(len(agg), nc))
            levels[-1]["P"] = P
            A = sp.csr_matrix(P.T @ A @ P)
    for k in range(len(aggs)):
        agg = np.asarray(aggs[k], dtype=np.int64)
        levels[k]["w"] = levels[k + 1]["d"][agg] / levels[k]["d"]
    return levels


def unpack(params: np.ndarray, n_layers: int):
    p = np.asarray(params, dtype=np.float64)
    o = 0
    we = p[o:o + 2 * F].reshape(2, F); o += 2 * F
    be = p[o:o + F]; o += F
    layers = []
    for _ in range(n_layers):
        W = p[o:o + 2 * F * F].reshape(2 * F, F); o += 2 * F * F
        b = p[o:o + F]; o += F
        layers.append((W, b))
    wd = p[o:o + F]; o += F
    bd = p[o]
    return we, be, layers, wd, bd


def _act(u, n_lin):
    out = u.copy()
    out[:, n_lin:] = np.maximum(out[:, n_lin:], 0.0)
    return out


def _dact(u, n_lin):
    g = np.ones_like(u)
    g[:, n_lin:] = (u[:, n_lin:] > 0).astype(np.float64)
    return g


def _program(schedule, bottom=None):
    """(kind, level, in, out, extra) ops mirroring the U-shaped schedule; between
    the down and the up half the features are pooled down to ``bottom``."""
    ops, cur, h, nb = [], 0, 0, 1
    skips = {}
    bottom = max(schedule) if bottom is None else bottom
    for t, lev in enumerate(schedule):
        if t == len(schedule) // 2:
            while cur < max(lev, bottom):
                skips[cur] = h
                ops.append(("pool", cur, h, nb, None)); h, cur, nb = nb, cur + 1, nb + 1
        while cur < lev:
            skips[cur] = h
            ops.append(("pool", cur, h, nb, None)); h, cur, nb = nb, cur + 1, nb + 1
        while cur > lev:
            ops.append(("unpool", cur - 1, h, nb, skips.pop(cur - 1))); h, cur, nb = nb, cur - 1, nb + 1
        ops.append(("mp", cur, h, nb, t)); h, nb = nb, nb + 1
    while cur > 0:
        ops.append(("unpool", cur - 1, h, nb, skips.pop(cur - 1))); h, cur, nb = nb, cur - 1, nb + 1
    return ops, h


def forward(levels, params, schedule, n_lin, r, lk, bf16=False, bottom=None):
    """Returns (z, y, s, tape)."""
    rnd = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    we, be, layers, wd, bd = unpack(params, len(schedule))
    d0 = levels[0]["d"]
    n = len(r)
    s = np.sqrt(np.sum((d0 * r) ** 2) / n)
    s = s if s > 0 else 1.0
    x = np.stack([(d0 * r) / s, lk], axis=1)
    if bf16:  # the kernel evaluates the encoder in float32
        x = x.astype(np.float32).astype(np.float64)
    pre0 = x @ we + be
    bufs = {0: rnd(pre0)}
    ops, final = _program(schedule, bottom)
    tape = {"x": x, "pre0": pre0, "ops": ops, "final": final, "bufs": bufs, "msg": {}, "pre": {}}
    for i, (kind, lev, a, b, extra) in enumerate(ops):
        L = levels[lev]
        if kind == "mp":
            W, bias = layers[extra]
            Wr = rnd(W)
            h = bufs[a]
            m = rnd(L["An"] @ h)
            u = np.concatenate([h, m], axis=1) @ Wr + bias
            bufs[b] = rnd(h + _act(u, n_lin))
            tape["msg"][i], tape["pre"][i] = m, u
        elif kind == "pool":
            bufs[b] = rnd(L["P"].T @ (L["w"][:, None] * bufs[a]))
        else:
            bufs[b] = rnd(bufs[extra] + L["w"][:, None] * (L["P"] @ bufs[a]))
    y = bufs[final] @ wd + bd
    z = s * d0 * y
    tape["y"] = y
    return z, y, s, tape


def loss(levels, y, r, s):
    d0, An = levels[0]["d"], levels[0]["An"]
    rt = d0 * r / s
    return (0.5 * y @ (An @ y) - y @ rt) / len(y)


def cosine_loss(levels, y, r, s, e):
    """A^-norm error left after the optimal step along y:
    1 - (y.r~)^2 / ((y.A^y)(e~.r~)), with r~ = d r / s and e~ = e / (d s)."""
    d0, An = levels[0]["d"], levels[0]["An"]
    rt, et = d0 * r / s, e / (d0 * s)
    return 1.0 - (y @ rt) ** 2 / ((y @ (An @ y)) * (et @ rt))


def cosine_grad_y(levels, y, r, s, e):
    d0, An = levels[0]["d"], levels[0]["An"]
    rt, et = d0 * r / s, e / (d0 * s)
    ay = An @ y
    yay, yr, er = y @ ay, y @ rt, et @ rt
    return -2.0 * yr / (yay * er) * (rt - yr / yay * ay)


def backward(levels, params, schedule, n_lin, r, lk, tape, s, bf16=False, gy=None):
    """Gradient of ``loss`` (or, with ``gy``, of any loss whose gradient with
    respect to the output y is ``gy``) w.r.t. the flat parameter vector."""
    rnd = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    we, be, layers, wd, bd = unpack(params, len(schedule))
    d0, An = levels[0]["d"], levels[0]["An"]
    n = len(r)
    y = tape["y"]
    if gy is None:
        gy = (An @ y - d0 * r / s) / n
    bufs, ops = tape["bufs"], tape["ops"]
    g_layers = [None] * len(schedule)
    grads = {tape["final"]: np.outer(gy, wd)}
    g_wd = bufs[tape["final"]].T @ gy
    g_bd = gy.sum()

    def acc(k, g):
        grads[k] = grads[k] + g if k in grads else g

    for i in range(len(ops) - 1, -1, -1):
        kind, lev, a, b, extra = ops[i]
        L = levels[lev]
        g = grads.pop(b)
        if kind == "mp":
            W, bias = layers[extra]
            Wr = rnd(W)
            gu = g * _dact(tape["pre"][i], n_lin)
            X = np.concatenate([bufs[a], tape["msg"][i]], axis=1)
            g_layers[extra] = (X.T @ gu, gu.sum(0))
            gx = gu @ Wr.T
            acc(a, g + gx[:, :F] + L["An"].T @ gx[:, F:])
        elif kind == "pool":
            acc(a, L["w"][:, None] * (L["P"] @ g))
        else:
            acc(extra, g)
            acc(a, L["P"].T @ (L["w"][:, None] * g))
    g0 = grads.pop(0)
    x = tape["x"]
    out = [(x.T @ g0).ravel(), g0.sum(0)]
    for (gW, gb) in g_layers:
        out += [gW.ravel(), gb]
    out += [g_wd, np.array([g_bd])]
    return np.concatenate(out)


def fcg(A: sp.csr_matrix, b: np.ndarray, precond: Callable[[np.ndarray], np.ndarray], rtol: float = 1e-8,
        maxiter: int = 1000, m: int = 1):
    """Flexible CG truncated to the last ``m`` directions (Notay 2000, FCG(m)),
    float64, with the explicit orthogonalisation p = z - sum_j (z.Ap_j)/(p_j.Ap_j) p_j.
    Returns (x, iterations, relres history)."""
    x = np.zeros_like(b)
    r = b.copy()
    bb = np.sqrt(b @ b)
    dirs = []          # (p_j, q_j = A p_j, p_j.q_j), newest last
    hist = [np.sqrt(r @ r) / bb]
    it = 0
    while hist[-1] > rtol and it < maxiter:
        z = precond(r)
        p = z.copy()
        for pj, qj, pqj in dirs:
            p -= (z @ qj) / pqj * pj
        q = A @ p
        pq = p @ q
        alpha = (p @ r) / pq
        x = x + alpha * p
        r = r - alpha * q
        dirs = (dirs + [(p, q, pq)])[-m:]
        it += 1
        hist.append(np.sqrt(r @ r) / bb)
    return x, it, hist
