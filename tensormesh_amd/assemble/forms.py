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


def p1_gradients(X: torch.Tensor) -> torch.Tensor:
    """``grad phi_b`` of P1 simplices ``X [E, nb, d]`` as ``[E, nb, d]``, in closed
    form: the cofactor vectors of the edge matrix over its determinant (the
    rows of ``J^-1``, as the HIP kernels compute them)."""
    d = X.shape[2]
    e = [X[:, k] - X[:, 0] for k in range(1, d + 1)]
    if d == 2:
        det = e[0][:, 0] * e[1][:, 1] - e[0][:, 1] * e[1][:, 0]
        c = [torch.stack([e[1][:, 1], -e[1][:, 0]], 1), torch.stack([-e[0][:, 1], e[0][:, 0]], 1)]
    else:
        c = [torch.linalg.cross(e[1], e[2]), torch.linalg.cross(e[2], e[0]), torch.linalg.cross(e[0], e[1])]
        det = (e[0] * c[0]).sum(1)
    c0 = -sum(c)
    return torch.stack([c0] + c, 1) / det[:, None, None]


def _coefficient_args(fn, points, conn, element_type, order, point_data, element_data, scalar_data):
    """The arguments of ``forward`` at every (element, quadrature point) for the
    probes; the basis arguments are placeholders that the probes replace."""
    names = list(inspect.signature(fn).parameters)
    w, q = quadrature(element_type, order)
    phi = torch.as_tensor(shape_val(element_type, q), dtype=points.dtype, device=points.device)
    grad = p1_gradients(points[conn])                                       # constant per element
    grad = grad[:, None].expand(grad.shape[0], len(w), grad.shape[1], grad.shape[2])
    args, dims = _args(names, phi, grad, point_data, element_data, scalar_data, conn, True)
    return names, phi, args, dims


def _chunks(conn, element_data, chunk):
    E = conn.shape[0]
    for s in range(0, E, chunk):
        yield conn[s:s + chunk], {k: v[s:s + chunk] for k, v in element_data.items()}


def _bilinear_moment_map(w: np.ndarray, phi: np.ndarray, d: int) -> np.ndarray:
    """``T [(d+1)^2 * Q, (nb+d)^2]`` with ``moments = D.reshape(E, -1) @ T`` for
    ``D [E, d+1, d+1, Q]``: ``MM [nb][nb]``, ``P [nb][d]``, ``Qm [nb][d]``, ``S [d][d]``."""
    Q, nb = phi.shape
    dd = d + 1
    T = np.zeros((Q, dd, dd, (nb + d) ** 2))
    o_p, o_q, o_s = nb * nb, nb * nb + nb * d, nb * nb + 2 * nb * d
    for q in range(Q):
        for a in range(nb):
            for b in range(nb):
                T[q, 0, 0, a * nb + b] = w[q] * phi[q, a] * phi[q, b]
            for j in range(d):
                T[q, 0, 1 + j, o_p + a * d + j] = w[q] * phi[q, a]
                T[q, 1 + j, 0, o_q + a * d + j] = w[q] * phi[q, a]
        for i in range(d):
            for j in range(d):
                T[q, 1 + i, 1 + j, o_s + i * d + j] = w[q]
    return T.transpose(1, 2, 0, 3).reshape(dd * dd * Q, -1)


def _chunk_for(points: torch.Tensor, chunk: Optional[int]) -> int:
    """Elements per probe pass: large on the GPU (fewer, larger kernels), small on the host."""
    if chunk is not None:
        return int(chunk)
    return 1 << 21 if points.is_cuda else 1 << 17


def probe_bilinear_q(fn: Callable, points: torch.Tensor, conn: torch.Tensor, element_type: str, order: int,
                     point_data: Dict, element_data: Dict, scalar_data: Dict, chunk: Optional[int] = None):
    """Per-element moments of a bilinear integrand whose coefficients vary in
    space, for ``tm_asm_p1_moments``, or ``None``.

    The reference evaluates ``forward`` at every (element, quadrature point,
    a, b) under ``vmap`` (``element_assembler.py:371-466``).  For P1 elements
    the integrand is ``U_a^T D(q) V_b`` whenever ``forward`` is bilinear in
    ``U = (u, gradu)`` and ``V = (v, gradv)``: ``D(q)`` is ``forward`` at the
    unit vectors (``(d+1)^2`` vectorised evaluations over all points), and
    three random ``(U, V)`` pairs must reproduce ``U^T D V`` at every point to
    1e-12.  The moments are one product of ``D`` with a constant map.  Returns
    ``(moments [E, (nb+d)^2] float64, symmetric)``."""
    names = list(inspect.signature(fn).parameters)
    if not (set(names) & {"u", "gradu"} and set(names) & {"v", "gradv"}):
        return None
    d = points.shape[1]
    dd = d + 1
    dt, dev = torch.float64, points.device
    w_np, q_np = quadrature(element_type, order)
    T = torch.as_tensor(_bilinear_moment_map(w_np, shape_val(element_type, q_np), d), dtype=dt, device=dev)
    gen = torch.Generator().manual_seed(20240602)
    checks = [(torch.rand(dd, generator=gen, dtype=dt) - 0.5, torch.rand(dd, generator=gen, dtype=dt) - 0.5)
              for _ in range(3)]
    I = torch.eye(dd, dtype=dt)
    out, symmetric = [], True
    try:
        for c, ed in _chunks(conn, element_data, _chunk_for(points, chunk)):
            names, phi, args, dims = _coefficient_args(fn, points, c, element_type, order, point_data, ed,
                                                       scalar_data)
            E, Q = c.shape[0], phi.shape[0]

            def at(U, V):
                U, V = U.to(dev), V.to(dev)
                return _eval_points(fn, names, args, dims, {"u": U[0], "gradu": U[1:], "v": V[0], "gradv": V[1:]},
                                    E, Q)

            cols = []
            for i in range(dd):
                for j in range(dd):
                    f = at(I[i], I[j])
                    if f is None:
                        return None
                    cols.append(f.to(dt))
            # [E, dd * dd, Q]: each element's quadrature values stay contiguous
            # (stacking on the last axis writes every 16th double)
            D = torch.stack(cols, 1)
            dmax = D.abs().amax(dim=(0, 2))                                   # per entry of D; NaN propagates
            if not torch.isfinite(dmax).all():
                return None
            for U, V in checks:
                f = at(U, V)
                uv = torch.outer(U, V).reshape(-1).to(dev)
                # elementwise products and a sum over the entries: a matrix-vector
                # product of this tall, narrow shape runs far below memory speed
                ref = (D * uv[:, None]).sum(1)
                bound = (dmax * uv.abs()).sum().clamp_min(1e-300)
                if f is None or ((f.to(dt) - ref).abs() > _TOL * bound).any():
                    return None
            D4 = D.view(E, dd, dd, Q)
            symmetric = symmetric and bool(torch.equal(D4, D4.transpose(1, 2)))
            out.append(D.reshape(E, -1) @ T)
    except ValueError:
        raise
    except Exception:
        return None
    return torch.cat(out, 0).contiguous(), symmetric


def probe_linear_q(fn: Callable, points: torch.Tensor, conn: torch.Tensor, element_type: str, order: int,
                   point_data: Dict, scalar_data: Dict, chunk: Optional[int] = None):
    """Per-element moments of a linear integrand ``F0(q) v + F1(q) . gradv``
    with spatially varying ``F0``, ``F1`` (``node_assembler.py:242-409``; e.g.
    ``func_node_assembler``, ``builtin.py:929-969``) for
    ``tm_asm_p1_load_moments``, or ``None``: ``F0`` and ``F1`` are ``forward``
    at ``(v, gradv) = (1, 0)`` and ``(0, e_j)``, checked on three random pairs
    to 1e-12.  Returns ``moments [E, nb + d]`` float64
    (``L0[a] = sum_q w_q phi_a F0``, ``L1[j] = sum_q w_q F1_j``)."""
    names = list(inspect.signature(fn).parameters)
    if not set(names) & {"v", "gradv"} or set(names) & {"u", "gradu"}:
        return None
    d = points.shape[1]
    dt, dev = torch.float64, points.device
    w_np, q_np = quadrature(element_type, order)
    ph_np = shape_val(element_type, q_np)
    Q, nb = ph_np.shape
    T_np = np.zeros((d + 1, Q, nb + d))                                    # for F [E, d + 1, Q]
    T_np[0, :, :nb] = w_np[:, None] * ph_np
    for j in range(d):
        T_np[1 + j, :, nb + j] = w_np
    T = torch.as_tensor(T_np.reshape((d + 1) * Q, nb + d), dtype=dt, device=dev)
    gen = torch.Generator().manual_seed(20240603)
    checks = [torch.rand(d + 1, generator=gen, dtype=dt) - 0.5 for _ in range(3)]
    I = torch.eye(d + 1, dtype=dt)
    out = []
    try:
        for c, _ in _chunks(conn, {}, _chunk_for(points, chunk)):
            names, phi, args, dims = _coefficient_args(fn, points, c, element_type, order, point_data, {},
                                                       scalar_data)
            E = c.shape[0]

            def at(V):
                V = V.to(dev)
                return _eval_points(fn, names, args, dims, {"v": V[0], "gradv": V[1:]}, E, Q)

            cols = []
            for j in range(d + 1):
                f = at(I[j])
                if f is None:
                    return None
                cols.append(f.to(dt))
            F = torch.stack(cols, 1)                                          # [E, d + 1, Q]
            fmax = F.abs().amax(dim=(0, 2))
            if not torch.isfinite(fmax).all():
                return None
            for V in checks:
                f = at(V)
                Vd = V.to(dev)
                bound = (fmax * Vd.abs()).sum().clamp_min(1e-300)
                if f is None or ((f.to(dt) - (F * Vd[:, None]).sum(1)).abs() > _TOL * bound).any():
                    return None
            out.append(F.reshape(E, -1) @ T)
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
