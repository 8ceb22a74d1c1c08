"""Linear-form (load vector) assembly.

Drop-in for ``tensormesh.NodeAssembler`` (``tensormesh/assemble/node_assembler.py``):
``from_mesh(mesh, quadrature_order=2, project='reduce')`` (``:700-738``),
``from_assembler``, ``__call__(points=None, func=None, point_data=None,
scalar_data=None, batch_size=1) -> Tensor[N * H]`` (``:242-409``).  On GPU
tensors the recognised loads ``c * v`` and ``f * v`` (nodal ``f``) run the
owner-computes HIP kernel (FP64 accumulation, returned in the points' dtype as the
reference's result is); other loads linear in ``(v, gradv)``
(coefficients from ``x``, point data and their gradients, as
``func_node_assembler`` builds) run the moments kernel on per-element
quadrature moments; anything else runs the generic path.
"""
from __future__ import annotations

from typing import Callable, Dict, Optional

import torch
import torch.nn as nn

from .. import _lib
from ..element import ELEMENT_DIM
from ..mesh import BufferDict, Mesh
from . import forms
from .topology import P1Topology, topology_for


def _in_dtype(b: torch.Tensor, points: torch.Tensor) -> torch.Tensor:
    """The kernels accumulate in FP64; the result takes the points' floating dtype
    (``self.type(points.dtype)``, ``node_assembler.py:290``)."""
    dt = points.dtype if points.dtype in (torch.float32, torch.float64) else torch.float64
    return b if b.dtype == dt else b.to(dt)


class NodeAssembler(nn.Module):
    #: Deterministic mode of the recognised-load kernel (``c * v``, ``f * v``):
    #: the entries of each row are rounded to a power-of-two grid 2^-44 below
    #: the row's largest entry, so the FP64 sums are exact and the result is
    #: bit-identical from call to call whatever the order of the GPU atomics.
    deterministic: bool = False

    def __init__(self, topology: P1Topology, elements: Dict[str, torch.Tensor], points: torch.Tensor,
                 quadrature_order: int = 2, *args, **kwargs):
        super().__init__()
        self.topology = topology
        self.quadrature_order = int(quadrature_order)
        self.elements = BufferDict({k: v.long() for k, v in elements.items()})
        self.register_buffer("points", points)
        self.element_types = list(elements.keys())
        self.dimension = ELEMENT_DIM[self.element_types[0]]
        self.n_points = int(points.shape[0])
        self._quad = topology.quad(self.quadrature_order)
        self.__post_init__(*args, **kwargs)

    def __post_init__(self, *args, **kwargs):
        pass

    def forward(self, *args, **kwargs):
        raise NotImplementedError("forward is not implemented")

    @classmethod
    def from_mesh(cls, mesh: Mesh, quadrature_order: int = 2, project: str = "reduce", *args, **kwargs):
        if project not in ("reduce", "sparse"):
            raise ValueError(f"project should be either 'reduce' or 'sparse', but got {project}")
        elements = mesh.elements()
        if isinstance(elements, torch.Tensor):
            elements = {mesh.default_element_type: elements}
        if len(elements) != 1:
            raise NotImplementedError("mixed element types are not on the P1 assembly path")
        (etype, conn), = elements.items()
        topo = topology_for(mesh.points, conn, etype)
        return cls(topo, elements, mesh.points, quadrature_order, *args, **kwargs)

    @classmethod
    def from_assembler(cls, obj, *args, **kwargs):
        assert isinstance(obj, NodeAssembler), f"the object {obj} should inheritate from NodeAssembler"
        elements = {k: obj.elements[k] for k in obj.elements.keys()}
        return cls(obj.topology, elements, obj.points, obj.quadrature_order, *args, **kwargs)

    def _launch(self, form, points64, f, c) -> torch.Tensor:
        dev = points64.device
        plan = self.topology.device_plan(dev)
        out = torch.empty(self.n_points, dtype=torch.float64, device=dev)
        fdt = _lib.dtype_code(f) if f is not None else _lib.F32
        if f is not None:
            f = f.contiguous()
        if self.deterministic:
            form |= _lib.ASM_DETERMINISTIC
        _lib.check(_lib.lib().tm_asm_p1_load(form, self.topology.nb, self.topology.dim, plan.struct,
                                             _lib.ptr(points64), _lib.ptr(f), fdt, float(c), self._quad,
                                             _lib.ptr(out), _lib.stream_of(dev)))
        return out

    def __call__(self, points: Optional[torch.Tensor] = None, func: Optional[Callable] = None,
                 point_data: Optional[Dict[str, torch.Tensor]] = None,
                 scalar_data: Optional[Dict[str, torch.Tensor]] = None, batch_size: int = 1) -> torch.Tensor:
        point_data = dict(point_data or {})
        scalar_data = dict(scalar_data or {})
        if points is None:
            points = self.points
        else:
            assert points.shape[1] == self.dimension, \
                f"the dimension of the points should be {self.dimension}, but got {points.shape[1]}"
        point_data["x"] = points
        for key, value in point_data.items():
            assert value.shape[0] == points.shape[0], \
                f"the shape of {key} should be [n_point, ...], but got {value.shape}"
        fn = self.forward if func is None else func
        dev = points.device
        if self.points.device != dev:
            self.to(dev)
        etype = self.element_types[0]
        conn = self.elements[etype]
        form = None
        pd = {k: v for k, v in point_data.items() if k != "x"}
        on_kernel = dev.type == "cuda" and not points.requires_grad and not any(
            torch.is_tensor(v) and v.requires_grad for v in list(pd.values()) + list(scalar_data.values()))
        if on_kernel:
            form = forms.classify_linear(fn, pd, scalar_data)
            if form is None:
                # spatially varying load (e.g. func_node_assembler): per-element
                # moments of the integrand, then the moments kernel
                mom = forms.probe_linear_q(fn, points, conn.to(dev), etype, self.quadrature_order, point_data,
                                           scalar_data)
                if mom is not None:
                    p64 = points.detach().to(torch.float64).contiguous()
                    plan = self.topology.device_plan(dev)
                    out = torch.empty(self.n_points, dtype=torch.float64, device=dev)
                    _lib.check(_lib.lib().tm_asm_p1_load_moments(
                        self.topology.nb, self.topology.dim, plan.struct, _lib.ptr(p64), _lib.ptr(mom),
                        _lib.ptr(out), _lib.ASM_DETERMINISTIC if self.deterministic else 0, _lib.stream_of(dev)))
                    return _in_dtype(out, points)
        if form is not None:
            kind, scale, key = form
            p64 = points.detach().to(torch.float64).contiguous()
            if kind == _lib.LOAD_CONST:
                b = self._launch(kind, p64, None, scale)
            else:
                b = self._launch(kind, p64, point_data[key].to(dev), 1.0)
                if scale != 1.0:
                    b = b * scale
            return _in_dtype(b, points)
        integ = forms.generic_element_integrals(fn, points, conn.to(dev), etype, self.quadrature_order,
                                                point_data, {}, scalar_data, bilinear=False)
        out = torch.zeros((self.n_points,) + tuple(integ.shape[2:]), dtype=integ.dtype, device=dev)
        out = out.index_add(0, conn.reshape(-1).to(dev), integ.reshape((-1,) + tuple(integ.shape[2:])))
        return out.flatten()
