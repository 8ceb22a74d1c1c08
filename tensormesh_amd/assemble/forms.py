"""Recognition of weak forms that the HIP kernels implement, and the generic path.

The reference dispatches ``forward`` arguments by name
(``tensormesh/assemble/element_assembler.py:332-365``,
``node_assembler.py:341-366``) and evaluates it under nested ``vmap``.  A
P1 kernel can only implement a fixed integrand, so ``forward`` is probed: it
is evaluated on random per-quadrature-point arguments and compared with the
candidate integrands below.  A match (constant ratio, relative spread
<= 1e-12 over three draws) selects the kernel with that ratio as scale:

======================================  =====================  ==============
argument names (besides scalar_data)     integrand               kernel form
======================================  =====================  ==============
``gradu, gradv``                         s * gradu . gradv       LAPLACE
``gradu, gradv, k`` (k nodal, 1-D)       s * k gradu . gradv     KAPPA_LAPLACE
``u, v``                                 s * u v                 MASS
``v``                  (load)            s * v                   LOAD_CONST
``v, f`` (f nodal, 1-D) (load)           s * f v                 LOAD_NODAL
======================================  =====================  ==============

Any other ``forward`` runs through :func:`generic_matrix` /
:func:`generic_load`: the same name-dispatch and ``vmap`` evaluation on the
tensors' own device, with the element integrals scattered by ``index_add_``
into the same CSR slots.  It is the general-form path, not the hot path.
"""
from __future__ import annotations

import inspect
from typing import Callable, Dict, Optional, Tuple

import numpy as np
import torch

from .. import _lib
from ..element import quadrature, shape_val, ref_grad

_TOL = 1e-12


def _probe(fn: Callable, names, build, ref) -> Optional[float]:
    gen = torch.Generator().manual_seed(20240531)
    ratios = []
    try:
        for _ in range(3):
            args = build(gen)
            out = fn(*[args[n] for n in names])
            if not torch.is_tensor(out) or out.numel() != 1:
                return None
            e = float(ref(args))
            ratios.append(float(out.reshape(())) / e)
    except Exception:
        return None
    r0 = ratios[0]
    if not np.isfinite(r0) or r0 == 0.0:
        return None
    if max(abs(r - r0) for r in ratios) > _TOL * abs(r0):
        return None
    return r0


def _scalars(scalar_data):
    return {k: (v if torch.is_tensor(v) else torch.tensor(v, dtype=torch.float64)) for k, v in scalar_data.items()}


def _value_key(v):
    """Hashable stand-in for a value a ``forward`` may read: scalars and small
    tensors / arrays by value, anything else by identity."""
    if isinstance(v, (bool, int, float, str)) or v is None:
        return v
    if torch.is_tensor(v) and v.numel() <= 64:
        return ("t", str(v.dtype), tuple(v.shape), tuple(v.detach().cpu().reshape(-1).tolist()))
    if isinstance(v, np.ndarray) and v.size <= 64:
        return ("a", str(v.dtype), v.shape, tuple(v.reshape(-1).tolist()))
    return ("id", id(v))


def form_cache_key(fn: Callable, point_data: Dict, scalar_data: Dict, element_data: Dict):
    """Key under which a classification of ``fn`` may be reused, or ``None``.

    It covers what decides the result: the function object (for a bound method
    the function and its instance), the instance's scalar and small-tensor
    attributes, the closure's values and the module globals it names (by value,
    so a changed coefficient is probed again), the scalar data and the
    point-data ranks.
    """
    try:
        func = getattr(fn, "__func__", fn)
        owner = getattr(fn, "__self__", None)
        attrs = ()
        if owner is not None:
            attrs = tuple(sorted((k, _value_key(v)) for k, v in vars(owner).items()
                                 if isinstance(v, (bool, int, float, str)) or
                                 (torch.is_tensor(v) and v.numel() <= 64)))
        cells = tuple(_value_key(c.cell_contents) for c in (getattr(func, "__closure__", None) or ()))
        defaults = tuple(_value_key(d) for d in (getattr(func, "__defaults__", None) or ()))
        glb = getattr(func, "__globals__", {})
        names = tuple((n, _value_key(glb[n])) for n in getattr(getattr(func, "__code__", None), "co_names", ())
                      if n in glb)
        key = (func, id(owner), attrs, cells, defaults, names,
               tuple(sorted((k, _value_key(v)) for k, v in scalar_data.items())),
               tuple(sorted((k, v.dim()) for k, v in point_data.items())), tuple(sorted(element_data)))
        hash(key)
        return key
    except Exception:
        return None


def classify_bilinear(fn: Callable, dim: int, point_data: Dict, scalar_data: Dict,
                      element_data: Dict) -> Optional[Tuple[int, float, Optional[str]]]:
    """Return ``(form, scale, coefficient key)`` or ``None``."""
    names = list(inspect.signature(fn).parameters)
    sc = _scalars(scalar_data)
    rest = [n for n in names if n not in sc]
    if any(n in element_data for n in rest):
        return None

    def rnd(gen, shape=()):
        return torch.rand(shape, generator=gen, dtype=torch.float64) + 0.5

    def build(extra):
        def b(gen):
            a = dict(sc)
            for n in rest:
                a[n] = rnd(gen, (dim,)) if n in ("gradu", "gradv") else rnd(gen)
            return a
        return b

    s = set(rest)
    if s == {"gradu", "gradv"}:
        r = _probe(fn, names, build(None), lambda a: torch.dot(a["gradu"], a["gradv"]))
        if r is not None:
            return (_lib.FORM_LAPLACE, r, None)
    if len(rest) == 3 and {"gradu", "gradv"} <= s:
        k = (s - {"gradu", "gradv"}).pop()
        if k in point_data:
            if point_data[k].dim() != 1:
                return None
            r = _probe(fn, names, build(None), lambda a: a[k] * torch.dot(a["gradu"], a["gradv"]))
            return None if r is None else (_lib.FORM_KAPPA_LAPLACE, r, k)
    if s == {"u", "v"}:
        r = _probe(fn, names, build(None), lambda a: a["u"] * a["v"])
        if r is not None:
            return (_lib.FORM_MASS, r, None)
    if s and s <= {"u", "gradu", "v", "gradv"} and s & {"u", "gradu"} and s & {"v", "gradv"}:
        D = probe_general(fn, names, sc, dim)
        if D is not None:
            return (_lib.FORM_GENERAL, D, None)
    return None


def probe_general(fn: Callable, names, sc: Dict, dim: int) -> Optional[np.ndarray]:
    """``D`` with ``fn = U^T D V`` for ``U = (u, gradu)``, ``V = (v, gradv)``, or ``None``.

    ``D_ij`` is ``fn`` at the unit vectors ``U = e_i``, ``V = e_j``; three random
    pairs must then satisfy ``fn(U, V) = U^T D V`` to 1e-12 of ``|U|^T |D| |V|``
    (so ``fn`` is bilinear with constant coefficients).
    """
    dd = dim + 1

    def call(U, V):
        a = dict(sc)
        for n in names:
            if n == "u":
                a[n] = U[0]
            elif n == "gradu":
                a[n] = U[1:]
            elif n == "v":
                a[n] = V[0]
            elif n == "gradv":
                a[n] = V[1:]
        out = fn(*[a[n] for n in names])
        if not torch.is_tensor(out) or out.numel() != 1:
            raise ValueError("not a scalar integrand")
        return float(out.reshape(()))

    try:
        I = torch.eye(dd, dtype=torch.float64)
        D = np.array([[call(I[i], I[j]) for j in range(dd)] for i in range(dd)])
        if not np.isfinite(D).all() or not D.any():
            return None
        gen = torch.Generator().manual_seed(20240601)
        for _ in range(3):
            U = torch.rand(dd, generator=gen, dtype=torch.float64) - 0.5
            V = torch.rand(dd, generator=gen, dtype=torch.float64) - 0.5
            u, v = U.numpy(), V.numpy()
            if abs(call(U, V) - u @ D @ v) > _TOL * (np.abs(u) @ np.abs(D) @ np.abs(v)):
                return None
    except Exception:
        return None
    return D


def classify_linear(fn: Callable, point_data: Dict, scalar_data: Dict) -> Optional[Tuple[int, float, Optional[str]]]:
    names = list(inspect.signature(fn).parameters)
    sc = _scalars(scalar_data)
    rest = [n for n in names if n not in sc]

    def build(gen):
        a = dict(sc)
        for n in rest:
            a[n] = torch.rand((), generator=gen, dtype=torch.float64) + 0.5
        return a

    if rest == ["v"]:
        r = _probe(fn, names, build, lambda a: a["v"])
        return None if r is None else (_lib.LOAD_CONST, r, None)
    if len(rest) == 2 and "v" in rest:
        k = [n for n in rest if n != "v"][0]
        if k in point_data and point_data[k].dim() == 1:
            r = _probe(fn, names, build, lambda a: a[k] * a["v"])
            return None if r is None else (_lib.LOAD_NODAL, r, k)
    return None


# ------------------------------------------------ spatially varying forms ---

_BASIS = ("u", "gradu", "v", "gradv")


def _eval_points(fn, names, args, dims, basis, E: int, Q: int):
    """``fn`` at every (element, quadrature point) with the basis arguments
    replaced by the constants in ``basis``; returns ``[E, Q]`` or ``None`` when
    ``fn`` does not give one scalar per point."""
    a2 = [basis[n] if n in _BASIS else a for n, a in zip(names, args)]
    d2 = [(None, None) if n in _BASIS else (d[0], d[1]) for n, d in zip(names, dims)]
    has_e = any(d[0] is not None for d in d2)
    has_q = any(d[1] is not None for d in d2)
    out = _vmapped(fn, d2, 2)(*a2)
    if not torch.is_tensor(out):
        return None
    want = (E if has_e else None, Q if has_q else None)
    shape = tuple(x for x in want if x is not None)
    if tuple(out.shape) != shape:
        return None
    if not has_e:
        out = out.unsqueeze(0)
    if not has_q:
        out = out.unsqueeze(1)
    return out.expand(E, Q)


def _coefficient_args(fn, points, conn, element_type, order, point_data, element_data, scalar_data):
    names = list(inspect.signature(fn).parameters)
    phi, grad, _ = geometry_torch(points, conn, element_type, order)
    args, dims = _args(names, phi, grad, point_data, element_data, scalar_data, conn, True)
    return names, phi, args, dims


def _chunks(conn, element_data, chunk):
    E = conn.shape[0]
    for s in range(0, E, chunk):
        yield conn[s:s + chunk], {k: v[s:s + chunk] for k, v in element_data.items()}


def probe_bilinear_q(fn: Callable, points: torch.Tensor, conn: torch.Tensor, element_type: str, order: int,
                     point_data: Dict, element_data: Dict, scalar_data: Dict, chunk: int = 1 << 19):
    """Per-element moments of a bilinear integrand whose coefficients vary in
    space, for ``tm_asm_p1_moments``, or ``None``.

    The reference evaluates ``forward`` at every (element, quadrature point,
    a, b) under ``vmap`` (``element_assembler.py:371-466``).  For P1 elements
    the integrand is ``U_a^T D(q) V_b`` whenever ``forward`` is bilinear in
    ``U = (u, gradu)`` and ``V = (v, gradv)``: ``D(q)`` is ``forward`` at the
    unit vectors (``(d+1)^2`` vectorised evaluations over all points), and
    three random ``(U, V)`` pairs must reproduce ``U^T D V`` at every point to
    1e-12.  Returns ``(moments [E, (nb+d)^2] float64, symmetric)``."""
    names = list(inspect.signature(fn).parameters)
    if not (set(names) & {"u", "gradu"} and set(names) & {"v", "gradv"}):
        return None
    d = points.shape[1]
    dd = d + 1
    dt, dev = torch.float64, points.device
    w_np, q_np = quadrature(element_type, order)
    w = torch.as_tensor(w_np, dtype=dt, device=dev)
    gen = torch.Generator().manual_seed(20240602)
    checks = [(torch.rand(dd, generator=gen, dtype=dt) - 0.5, torch.rand(dd, generator=gen, dtype=dt) - 0.5)
              for _ in range(3)]
    out, symmetric = [], True
    try:
        for c, ed in _chunks(conn, element_data, chunk):
            names, phi, args, dims = _coefficient_args(fn, points, c, element_type, order, point_data, ed,
                                                       scalar_data)
            E, Q = c.shape[0], phi.shape[0]

            def basis(U, V):
                U, V = U.to(dev), V.to(dev)
                return {"u": U[0], "gradu": U[1:], "v": V[0], "gradv": V[1:]}

            I = torch.eye(dd, dtype=dt)
            D = torch.empty((E, Q, dd, dd), dtype=dt, device=dev)
            for i in range(dd):
                for j in range(dd):
                    f = _eval_points(fn, names, args, dims, basis(I[i], I[j]), E, Q)
                    if f is None:
                        return None
                    D[:, :, i, j] = f
            if not torch.isfinite(D).all():
                return None
            for U, V in checks:
                f = _eval_points(fn, names, args, dims, basis(U, V), E, Q)
                Ud, Vd = U.to(dev), V.to(dev)
                ref = torch.einsum("i,eqij,j->eq", Ud, D, Vd)
                bound = torch.einsum("i,eqij,j->eq", Ud.abs(), D.abs(), Vd.abs())
                if f is None or ((f.to(dt) - ref).abs() > _TOL * bound.max().clamp_min(1e-300)).any():
                    return None
            symmetric = symmetric and bool(torch.equal(D, D.transpose(-1, -2)))
            ph = phi.to(dt)
            MM = torch.einsum("q,qa,qb,eq->eab", w, ph, ph, D[..., 0, 0])
            P = torch.einsum("q,qa,eqj->eaj", w, ph, D[..., 0, 1:])
            Qm = torch.einsum("q,eqj,qb->ebj", w, D[..., 1:, 0], ph)
            S = torch.einsum("q,eqij->eij", w, D[..., 1:, 1:])
            out.append(torch.cat([MM.reshape(E, -1), P.reshape(E, -1), Qm.reshape(E, -1), S.reshape(E, -1)], 1))
    except ValueError:
        raise
    except Exception:
        return None
    return torch.cat(out, 0).contiguous(), symmetric


def probe_linear_q(fn: Callable, points: torch.Tensor, conn: torch.Tensor, element_type: str, order: int,
                   point_data: Dict, scalar_data: Dict, chunk: int = 1 << 19):
    """Per-element moments of a linear integrand ``F0(q) v + F1(q) . gradv``
    with spatially varying ``F0``, ``F1`` (``node_assembler.py:242-409``; e.g.
    ``func_node_assembler``, ``builtin.py:929-969``) for
    ``tm_asm_p1_load_moments``, or ``None``: ``F0`` and ``F1`` are ``forward``
    at ``(v, gradv) = (1, 0)`` and ``(0, e_j)``, checked on three random pairs
    to 1e-12.  Returns ``moments [E, nb + d]`` float64."""
    names = list(inspect.signature(fn).parameters)
    if not set(names) & {"v", "gradv"} or set(names) & {"u", "gradu"}:
        return None
    d = points.shape[1]
    dt, dev = torch.float64, points.device
    w_np, _ = quadrature(element_type, order)
    w = torch.as_tensor(w_np, dtype=dt, device=dev)
    gen = torch.Generator().manual_seed(20240603)
    checks = [torch.rand(d + 1, generator=gen, dtype=dt) - 0.5 for _ in range(3)]
    out = []
    try:
        for c, _ in _chunks(conn, {}, chunk):
            names, phi, args, dims = _coefficient_args(fn, points, c, element_type, order, point_data, {},
                                                       scalar_data)
            E, Q = c.shape[0], phi.shape[0]

            def basis(V):
                V = V.to(dev)
                return {"v": V[0], "gradv": V[1:]}

            I = torch.eye(d + 1, dtype=dt)
            F = torch.empty((E, Q, d + 1), dtype=dt, device=dev)
            for j in range(d + 1):
                f = _eval_points(fn, names, args, dims, basis(I[j]), E, Q)
                if f is None:
                    return None
                F[:, :, j] = f
            if not torch.isfinite(F).all():
                return None
            for V in checks:
                f = _eval_points(fn, names, args, dims, basis(V), E, Q)
                Vd = V.to(dev)
                ref = torch.einsum("eqj,j->eq", F, Vd)
                bound = torch.einsum("eqj,j->eq", F.abs(), Vd.abs())
                if f is None or ((f.to(dt) - ref).abs() > _TOL * bound.max().clamp_min(1e-300)).any():
                    return None
            ph = phi.to(dt)
            L0 = torch.einsum("q,qa,eq->ea", w, ph, F[..., 0])
            L1 = torch.einsum("q,eqj->ej", w, F[..., 1:])
            out.append(torch.cat([L0, L1], 1))
    except ValueError:
        raise
    except Exception:
        return None
    return torch.cat(out, 0).contiguous()


# ---------------------------------------------------------------- generic path ---

def geometry_torch(points: torch.Tensor, conn: torch.Tensor, element_type: str, order: int):
    """``(shape_val [Q, nb], grad [E, Q, nb, d], jxw [E, Q])`` with torch ops
    (``element.py:958-966``, ``transformation.py:1284-1290``)."""
    w, q = quadrature(element_type, order)
    dt, dev = points.dtype, points.device
    phi = torch.as_tensor(shape_val(element_type, q), dtype=dt, device=dev)
    G = torch.as_tensor(np.broadcast_to(ref_grad(element_type).T[None], (len(w),) + ref_grad(element_type).T.shape).copy(),
                        dtype=dt, device=dev)                               # [Q, d, nb]
    X = points[conn]                                                         # [E, nb, d]
    J = torch.einsum("ebj,qib->eqij", X, G)
    grad = torch.einsum("qib,eqji->eqbj", G, torch.linalg.inv(J))
    jxw = torch.as_tensor(w, dtype=dt, device=dev)[None] * torch.linalg.det(J).abs()
    return phi, grad, jxw


def _args(names, phi, grad, point_data, element_data, scalar_data, conn, with_u: bool):
    """Arguments and in_dims per level (element, quadrature, u, v) like the
    reference's broadcast table (``element_assembler.py:332-346``)."""
    args, dims = [], []
    epd = {}
    for n in names:
        if n in ("u", "v"):
            args.append(phi)
            dims.append((None, 0, 0 if n == "u" else None, 0 if n == "v" else None))
        elif n in ("gradu", "gradv"):
            args.append(grad)
            dims.append((0, 0, 0 if n == "gradu" else None, 0 if n == "gradv" else None))
        elif n in point_data:
            v = epd.setdefault(n, point_data[n][conn])
            args.append(torch.einsum("eb...,qb->eq...", v, phi.to(v.dtype)))
            dims.append((0, 0, None, None))
        elif n.startswith("grad") and n[4:] in point_data:
            v = epd.setdefault(n[4:], point_data[n[4:]][conn])
            args.append(torch.einsum("eb...,eqbd->eq...d", v, grad.to(v.dtype)))
            dims.append((0, 0, None, None))
        elif n in element_data:
            args.append(element_data[n])
            dims.append((0, None, None, None))
        elif n in scalar_data:
            v = scalar_data[n]
            args.append(v if torch.is_tensor(v) else torch.tensor(v))
            dims.append((None, None, None, None))
        else:
            raise ValueError(f"{n} is not supported, please use `u`, `v`, `gradu`, `gradv` or more keys "
                             f"provided by point_data, element_data or scalar_data")
    if not with_u:
        dims = [(e, q, v) for (e, q, u, v) in dims]
    return args, dims


def _vmapped(fn, dims, levels):
    f = fn
    for lv in reversed(range(levels)):
        d = tuple(x[lv] for x in dims)
        if all(x is None for x in d):
            continue
        f = torch.vmap(f, in_dims=d)
    return f


def generic_element_integrals(fn, points, conn, element_type, order, point_data, element_data, scalar_data,
                              bilinear: bool):
    names = list(inspect.signature(fn).parameters)
    phi, grad, jxw = geometry_torch(points, conn, element_type, order)
    args, dims = _args(names, phi, grad, point_data, element_data, scalar_data, conn, bilinear)
    levels = 4 if bilinear else 3
    elem_batched = any(d[0] is not None for d in dims)
    if elem_batched:
        out = _vmapped(fn, dims, levels)(*args)                 # [E, Q, nb(, nb), ...]
        return torch.einsum("eq...,eq->e...", out, jxw.to(out.dtype))
    out = _vmapped(fn, [d[1:] for d in dims], levels - 1)(*args)  # [Q, nb(, nb), ...]
    return torch.einsum("q...,eq->e...", out, jxw.to(out.dtype))


def generic_energy(fn, points, conn, element_type, order, point_data, element_data, scalar_data):
    """``sum_e sum_q jxw psi(...)`` with the reference's energy broadcast rules
    (``element_assembler.py:582-680``): point data and their gradients are
    interpolated per quadrature point, element data is per element (or per
    element and quadrature point when its second dimension equals Q)."""
    names = list(inspect.signature(fn).parameters)
    phi, grad, jxw = geometry_torch(points, conn, element_type, order)
    Q = phi.shape[0]
    args, edims, qdims = [], [], []
    for n in names:
        if n in element_data:
            v = element_data[n]
            args.append(v)
            edims.append(0)
            qdims.append(0 if (v.dim() >= 2 and v.shape[1] == Q) else None)
        elif n in scalar_data:
            v = scalar_data[n]
            args.append(v if torch.is_tensor(v) else torch.tensor(v))
            edims.append(None)
            qdims.append(None)
        elif n in point_data:
            v = point_data[n][conn]
            args.append(torch.einsum("eb...,qb->eq...", v, phi.to(v.dtype)))
            edims.append(0)
            qdims.append(0)
        elif n.startswith("grad") and n[4:] in point_data:
            v = point_data[n[4:]][conn]
            args.append(torch.einsum("eb...,eqbd->eq...d", v, grad.to(v.dtype)))
            edims.append(0)
            qdims.append(0)
        else:
            raise ValueError(f"{n} is not supported for energy calculation.")
    dens = torch.vmap(torch.vmap(fn, in_dims=tuple(qdims)), in_dims=tuple(edims))(*args)
    return (dens * jxw.to(dens.dtype)).sum()
