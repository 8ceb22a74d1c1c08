"""Boundary-facet load vectors (Neumann tractions, Robin terms, surface loads).

Drop-in for ``tensormesh.FacetAssembler``
(``tensormesh/assemble/facet_assembler.py``): ``from_mesh(mesh,
boundary_mask=None, quadrature_order=2, project='reduce')`` (``:517-566``),
``from_elements(points, elements, boundary_mask, quadrature_order=2,
device='cpu', dtype=torch.float32, project='reduce')`` (``:368-515``),
``from_assembler`` (``:338-366``), ``__call__(points=None, func=None,
point_data=None) -> Tensor[N * H]`` (``:123-277``) and ``forward`` evaluated
at one facet quadrature point with the cell basis kept: ``v``/``u`` are the
``[B]`` cell basis values there, ``gradv``/``gradu`` the ``[B, D]`` cell
gradients, ``x`` and point data (and ``grad<key>``) the interpolated fields.

Facets follow the reference's selection rule: facet ``k`` of element ``e``
(the facet opposite local vertex ``k``) is integrated iff every one of its
nodes is flagged in ``boundary_mask`` (``:480-496``).  The measure is the
facet Jacobian times the facet rule's weights (``transformation.py:1051-1112``).

On GPU tensors, a ``forward`` linear in the cell basis (``F0 v + F1 . gradv``
per basis function, scalar or ``[H]``-valued) is tabulated at the facet
quadrature points and integrated by ``tm_facet_load`` (FP64 facet geometry,
quadrature and scatter); any other ``forward``, and CPU tensors, run the
generic path: the same evaluation under ``vmap`` and an ``index_add``.
"""
from __future__ import annotations

import inspect
from typing import Callable, Dict, Optional, Union

import numpy as np
import torch
import torch.nn as nn

from .. import _lib
from ..element import ELEMENT_DIM, ELEMENT_NB, facet_quadrature
from ..mesh import Mesh
from .forms import _TOL, geometry_torch


def select_facets(elements: torch.Tensor, boundary_mask: torch.Tensor):
    """``(cells [F, nb], opposite [F])``: every (element, local facet) whose
    nodes are all flagged; facet ``k`` is the one opposite local vertex ``k``."""
    conn = elements.long()
    nb = conn.shape[1]
    mask = boundary_mask.to(conn.device).bool()
    cells, opp = [], []
    for k in range(nb):
        others = [a for a in range(nb) if a != k]
        sel = mask[conn[:, others]].all(-1)
        cells.append(conn[sel])
        opp.append(torch.full((int(sel.sum()),), k, dtype=torch.int32, device=conn.device))
    return torch.cat(cells, 0).contiguous(), torch.cat(opp, 0).contiguous()


class FacetAssembler(nn.Module):
    #: Deterministic mode of the facet kernel: each entry's contributions are
    #: rounded to a power-of-two grid 2^-44 below the entry's largest one, so
    #: the FP64 sums are exact and repeat bit for bit whatever the atomics' order.
    deterministic: bool = False

    def __init__(self, cells: torch.Tensor, opposite: torch.Tensor, points: torch.Tensor, element_type: str,
                 quadrature_order: int = 2, *args, **kwargs):
        super().__init__()
        self.register_buffer("cells", cells.long())
        self.register_buffer("opposite", opposite.to(torch.int32))
        self.register_buffer("points", points)
        self.element_types = [element_type]
        self.dimension = ELEMENT_DIM[element_type]
        self.n_points = int(points.shape[0])
        self.quadrature_order = int(quadrature_order)
        self.__post_init__(*args, **kwargs)

    def __post_init__(self, *args, **kwargs):
        pass

    def forward(self, *args):
        raise NotImplementedError("forward is not implemented")

    @property
    def n_facets(self) -> int:
        return int(self.cells.shape[0])

    @property
    def device(self) -> torch.device:
        return self.points.device

    @property
    def dtype(self) -> torch.dtype:
        return self.points.dtype

    def type(self, dtype: torch.dtype):
        if dtype == torch.float64:
            self.double()
        elif dtype == torch.float32:
            self.float()
        else:
            raise Exception(f"the dtype {dtype} is not supported")
        return self

    # ---- construction ---------------------------------------------------------
    @classmethod
    def from_elements(cls, points: torch.Tensor, elements: Dict[str, torch.Tensor], boundary_mask: torch.Tensor,
                      quadrature_order: int = 2, device: Union[str, torch.device] = "cpu",
                      dtype: torch.dtype = torch.float32, project: str = "reduce", *args, **kwargs):
        if project not in ("reduce", "sparse"):
            raise ValueError(f"project should be either 'reduce' or 'sparse', but got {project}")
        if len(elements) != 1:
            raise NotImplementedError("mixed element types are not on the P1 facet path")
        (etype, conn), = elements.items()
        if etype not in ELEMENT_NB:
            raise NotImplementedError(f"facet assembly covers P1 triangles and tetrahedra, not {etype}")
        assert boundary_mask.dim() == 1 and boundary_mask.shape[0] == points.shape[0]
        cells, opp = select_facets(torch.as_tensor(conn), torch.as_tensor(boundary_mask))
        asm = cls(cells, opp, points, etype, quadrature_order, *args, **kwargs)
        return asm.type(dtype).to(device)

    @classmethod
    def from_mesh(cls, mesh: Mesh, boundary_mask: Optional[Union[str, torch.Tensor]] = None,
                  quadrature_order: int = 2, project: str = "reduce", *args, **kwargs):
        elements = mesh.elements()
        if isinstance(elements, torch.Tensor):
            elements = {mesh.default_element_type: elements}
        if boundary_mask is None:
            boundary_mask = mesh.boundary_mask
        elif isinstance(boundary_mask, str):
            boundary_mask = mesh.point_data[boundary_mask]
        assert boundary_mask.dim() == 1 and boundary_mask.shape[0] == mesh.points.shape[0]
        return cls.from_elements(mesh.points, dict(elements), boundary_mask, quadrature_order, mesh.device,
                                 mesh.dtype, project, *args, **kwargs)

    @classmethod
    def from_assembler(cls, obj, *args, **kwargs):
        assert isinstance(obj, FacetAssembler), f"the object {obj} should inheritate from FacetAssembler"
        return cls(obj.cells, obj.opposite, obj.points, obj.element_types[0], obj.quadrature_order,
                   *args, **kwargs)

    # ---- facet geometry (torch) ---------------------------------------------------
    def _facet_tables(self, points: torch.Tensor):
        """``(phi [F, Q, nb], fxw [F, Q], grad [F, nb, d], w [Q], bary [Q, d])``."""
        et = self.element_types[0]
        nb, d = ELEMENT_NB[et], self.dimension
        dt, dev = points.dtype, points.device
        w_np, bary_np = facet_quadrature(et, self.quadrature_order)
        w = torch.as_tensor(w_np, dtype=dt, device=dev)
        bary = torch.as_tensor(bary_np, dtype=dt, device=dev)
        table = torch.tensor([[a for a in range(nb) if a != k] for k in range(nb)], device=dev)
        fv = table[self.opposite.long()]                                   # [F, d]
        xf = points[self.cells.gather(1, fv)]                              # [F, d, d]
        if d == 2:
            meas = torch.linalg.vector_norm(xf[:, 1] - xf[:, 0], dim=-1)
        else:
            meas = torch.linalg.vector_norm(torch.cross(xf[:, 1] - xf[:, 0], xf[:, 2] - xf[:, 0], dim=-1), dim=-1)
        F, Q = self.n_facets, w.shape[0]
        phi = torch.zeros((F, Q, nb), dtype=dt, device=dev)
        phi.scatter_(2, fv[:, None, :].expand(F, Q, d), bary[None].expand(F, Q, d))
        fxw = w[None, :] * meas[:, None]
        _, grad, _ = geometry_torch(points, self.cells, et, 1)
        return phi, fxw, grad[:, 0], w, bary

    def _arguments(self, names, points, point_data, phi, grad):
        """Arguments of ``forward`` per name, with their (facet, point) batch dims."""
        F, Q = phi.shape[:2]
        args, dims = [], []
        for n in names:
            if n in ("u", "v"):
                args.append(phi)
                dims.append((0, 0))
            elif n in ("gradu", "gradv"):
                args.append(grad[:, None].expand(F, Q, *grad.shape[1:]))
                dims.append((0, 0))
            elif n in point_data:
                v = point_data[n][self.cells]                              # [F, nb, ...]
                args.append(torch.einsum("fb...,fqb->fq...", v, phi.to(v.dtype)))
                dims.append((0, 0))
            elif n.startswith("grad") and n[4:] in point_data:
                v = point_data[n[4:]][self.cells]
                g = torch.einsum("fb...,fbd->f...d", v, grad.to(v.dtype))
                args.append(g[:, None].expand(F, Q, *g.shape[1:]))
                dims.append((0, 0))
            else:
                raise NotImplementedError(f"key {n} is not implemented")
        return args, dims

    # ---- call ------------------------------------------------------------------
    def __call__(self, points: Optional[torch.Tensor] = None, func: Optional[Callable] = None,
                 point_data: Optional[Dict[str, torch.Tensor]] = None) -> torch.Tensor:
        point_data = dict(point_data or {})
        if points is None:
            points = self.points
        else:
            self.type(points.dtype).to(points.device)
        point_data["x"] = points
        for key, value in point_data.items():
            assert value.shape[0] == self.n_points, \
                f"the shape of {key} should be [n_point, ...], but got {value.shape}"
        fn = self.forward if func is None else func
        names = list(inspect.signature(fn).parameters)
        dev = points.device
        if self.n_facets == 0:
            return torch.zeros(self.n_points, dtype=points.dtype, device=dev)
        phi, fxw, grad, w, bary = self._facet_tables(points)
        if dev.type == "cuda" and not any(torch.is_tensor(v) and v.requires_grad for v in point_data.values()):
            tab = self._probe_linear(fn, names, points, point_data, phi, grad)
            if tab is not None:
                F0, F1 = tab
                H = F0.shape[-1]
                b = torch.zeros(self.n_points * H, dtype=torch.float64, device=dev)
                p64 = points.detach().to(torch.float64).contiguous()
                bary64 = bary.detach().to(torch.float64).contiguous()
                w64 = w.detach().to(torch.float64).contiguous()
                scratch = torch.zeros_like(b) if self.deterministic else None
                _lib.check(_lib.lib().tm_facet_load(
                    ELEMENT_NB[self.element_types[0]], self.dimension, self.n_facets, _lib.ptr(p64),
                    _lib.ptr(self.cells), _lib.ptr(self.opposite), int(w64.numel()), _lib.ptr(bary64),
                    _lib.ptr(w64), int(H), _lib.ptr(F0), _lib.ptr(F1), _lib.ptr(b),
                    _lib.ASM_DETERMINISTIC if self.deterministic else 0, _lib.ptr(scratch), _lib.stream_of(dev)))
                return b if points.dtype != torch.float32 else b.to(torch.float32)
        args, dims = self._arguments(names, points, point_data, phi, grad)
        out = torch.vmap(torch.vmap(fn))(*args)                             # [F, Q, nb, ...]
        integ = torch.einsum("fqb...,fq->fb...", out, fxw.to(out.dtype))
        res = torch.zeros((self.n_points,) + tuple(integ.shape[2:]), dtype=integ.dtype, device=dev)
        res = res.index_add(0, self.cells.reshape(-1), integ.reshape((-1,) + tuple(integ.shape[2:])))
        return res.flatten()

    def _probe_linear(self, fn, names, points, point_data, phi, grad):
        """``(F0 [F, Q, H], F1 [F, Q, H, d] or None)`` with ``forward(v, gradv)[a]
        = F0 v[a] + F1 . gradv[a]`` for every cell basis function ``a``, checked
        on three random ``(v, gradv)`` to 1e-12; ``None`` otherwise."""
        if set(names) & {"u", "gradu"} or not set(names) & {"v", "gradv"}:
            return None
        nb, d = phi.shape[2], grad.shape[2]
        F, Q = phi.shape[:2]
        dt, dev = torch.float64, points.device
        try:
            args, dims = self._arguments(names, points, point_data, phi, grad)

            def ev(v, gv):
                a2 = [v.to(dev) if n == "v" else (gv.to(dev) if n == "gradv" else a)
                      for n, a in zip(names, args)]
                d2 = [None if n in ("v", "gradv") else 0 for n in names]
                out = torch.vmap(torch.vmap(fn, in_dims=tuple(d2)), in_dims=tuple(d2))(*a2)
                if out.dim() == 3:
                    out = out[..., None]
                if out.dim() != 4 or tuple(out.shape[:3]) != (F, Q, nb) or out.shape[3] > 3:
                    raise ValueError("not a per-basis scalar or [H] integrand")
                return out.to(dt)                                             # [F, Q, nb, H]

            ones, zeros = torch.ones(nb, dtype=dt), torch.zeros(nb, dtype=dt)
            base = ev(ones, torch.zeros(nb, d, dtype=dt))
            F0 = base[:, :, 0]
            F1 = torch.zeros(F, Q, F0.shape[-1], d, dtype=dt, device=dev)
            if "gradv" in names:
                for j in range(d):
                    g = torch.zeros(nb, d, dtype=dt)
                    g[:, j] = 1.0
                    F1[..., j] = ev(zeros, g)[:, :, 0]
            gen = torch.Generator().manual_seed(20240604)
            for _ in range(3):
                v = torch.rand(nb, generator=gen, dtype=dt) - 0.5
                gv = torch.rand(nb, d, generator=gen, dtype=dt) - 0.5
                got = ev(v, gv)
                vd, gd = v.to(dev), gv.to(dev)
                ref = F0[:, :, None, :] * vd[None, None, :, None] + torch.einsum("fqhj,aj->fqah", F1, gd)
                bound = (F0.abs()[:, :, None, :] * vd.abs()[None, None, :, None]
                         + torch.einsum("fqhj,aj->fqah", F1.abs(), gd.abs()))
                if ((got - ref).abs() > _TOL * bound.max().clamp_min(1e-300)).any():
                    return None
        except NotImplementedError:
            raise
        except Exception:
            return None
        if not (torch.isfinite(F0).all() and torch.isfinite(F1).all()):
            return None
        return F0.contiguous(), (F1.contiguous() if F1.abs().max() > 0 else None)

    def __repr__(self):
        return (f"{self.__class__.__name__}(element_types={self.element_types}, n_points={self.n_points}, "
                f"n_facets={self.n_facets})")
