"""Bilinear-form assembly into a CSR-ordered :class:`SparseMatrix`.

Drop-in for ``tensormesh.ElementAssembler`` (``tensormesh/assemble/element_assembler.py``):
``from_mesh(mesh, quadrature_order=2, project='reduce', *args, **kwargs)``
(``:789-909``), ``from_assembler`` (``:762-787``), ``__call__(points=None,
func=None, point_data=None, element_data=None, scalar_data=None,
batch_size=-1)`` (``:243-466``), ``forward`` by subclassing, ``__post_init__``
for parameters, ``edges`` ``[2, nnz]`` in canonical CSR order.

On GPU tensors a recognised ``forward`` (see :mod:`.forms`) runs the
owner-computes HIP kernel: FP64 geometry, FP64 on-chip accumulation, values
stored in the points' dtype as the reference's are (``self.type(points.dtype)``,
``element_assembler.py:321``), or in ``value_dtype`` when it is set.  A ``forward`` that
is bilinear in ``(u, gradu)`` and ``(v, gradv)`` with coefficients varying in
space (``x``, point data and their gradients, element data) is probed at the
quadrature points and runs the moments kernel.  Other forms, inputs that
require gradients, and CPU tensors run the generic name-dispatch path.  Unlike the reference, the
matrix's ``row``/``col`` are the assembler's own int64 ``edges`` rows on
every call, so a Condenser's cached layout stays valid across assemblies.
"""
from __future__ import annotations

import inspect
from typing import Callable, Dict, Mapping, Optional, Union

import numpy as np
import torch
import torch.nn as nn

from .. import _lib
from ..element import ELEMENT_DIM
from ..mesh import BufferDict, Mesh
from ..sparse.matrix import SparseMatrix
from ..sparse.ops import CSR
from . import forms
from .topology import P1Topology, topology_for


class _KappaLaplace(torch.autograd.Function):
    """Values of the kappa-Laplace matrix with the hand-written adjoint for kappa."""

    @staticmethod
    def forward(ctx, kappa, asm, points64, scale, out_dtype):
        ctx.asm, ctx.points64, ctx.scale = asm, points64, scale
        ctx.kdtype = kappa.dtype
        return asm._launch_matrix(_lib.FORM_KAPPA_LAPLACE, points64, kappa.detach(), scale, out_dtype)

    @staticmethod
    def backward(ctx, g):
        asm = ctx.asm
        dk = asm._launch_kappa_adjoint(ctx.points64, g) * ctx.scale
        return dk.to(ctx.kdtype), None, None, None, None


class ElementAssembler(nn.Module):
    # dtype of the stored values: None follows the points (float32 or float64)
    value_dtype: Optional[torch.dtype] = None
    # bit-identical values on every call (TM_ASM_DETERMINISTIC: entries rounded
    # to a per-row power-of-two grid, at most 2^-44 of the row's largest entry,
    # so the on-chip sums are exact and order-independent)
    deterministic: bool = False

    def __init__(self, topology: P1Topology, elements: Dict[str, torch.Tensor], points: torch.Tensor,
                 quadrature_order: int = 2, *args, **kwargs):
        super().__init__()
        self.topology = topology
        self.quadrature_order = int(quadrature_order)
        self.elements = BufferDict({k: v.long() for k, v in elements.items()})
        self.register_buffer("edges", topology.edges_on(points.device))
        self.register_buffer("points", points)
        self.element_types = list(elements.keys())
        self.dimension = ELEMENT_DIM[self.element_types[0]]
        self.n_points = int(points.shape[0])
        self._quad = topology.quad(self.quadrature_order)
        self._form_cache = {}
        self.__post_init__(*args, **kwargs)

    def __post_init__(self, *args, **kwargs):
        pass

    def forward(self, *args, **kwargs):
        raise NotImplementedError("forward is not implemented")

    @property
    def device(self) -> torch.device:
        return self.points.device

    @property
    def dtype(self) -> torch.dtype:
        return self.points.dtype

    @property
    def nnz(self) -> int:
        return self.topology.nnz

    @property
    def projector(self):
        """The reference's element->edge indices ``P_E`` (int64, element-major), on the
        assembler's device (built there when the pattern was)."""
        return self.topology.element_slots_on(self.edges.device)

    # ---- construction -----------------------------------------------------------
    @classmethod
    def from_mesh(cls, mesh: Mesh, quadrature_order: int = 2, project: str = "reduce", *args, **kwargs):
        assert project in ["reduce", "sparse"]
        elements = mesh.elements()
        if isinstance(elements, torch.Tensor):
            elements = {mesh.default_element_type: elements}
        if len(elements) != 1:
            raise NotImplementedError("mixed element types are not on the P1 assembly path")
        (etype, conn), = elements.items()
        topo = topology_for(mesh.points, conn, etype)
        return cls(topo, elements, mesh.points, quadrature_order, *args, **kwargs)

    @classmethod
    def from_assembler(cls, obj: "ElementAssembler", *args, **kwargs):
        assert isinstance(obj, ElementAssembler), \
            f"obj must be an instance of ElementAssembler, but got {type(obj)}"
        elements = {k: obj.elements[k] for k in obj.elements.keys()}
        return cls(obj.topology, elements, obj.points, obj.quadrature_order, *args, **kwargs)

    # ---- kernels --------------------------------------------------------------
    def _out_dtype(self, points: torch.Tensor) -> torch.dtype:
        if self.value_dtype is not None:
            return self.value_dtype
        return points.dtype if points.dtype in (torch.float32, torch.float64) else torch.float32

    def _plan(self, device):
        return self.topology.device_plan(device)

    def csr(self, device=None) -> CSR:
        device = torch.device(device or self.device)
        rp, ci = self.topology.csr(device)
        return CSR(rp, ci, self.n_points, self.n_points)

    def _launch_matrix(self, form, points64, coef, scale, out_dtype) -> torch.Tensor:
        dev = points64.device
        # symmetric forms read the plan's edge records (the deterministic mode keeps the 16-byte ones)
        plan = self.topology.device_plan(dev, edge=not self.deterministic)
        values = torch.empty(self.topology.nnz, dtype=out_dtype, device=dev)
        cdt = _lib.dtype_code(coef) if coef is not None else _lib.F32
        if coef is not None:
            coef = coef.contiguous()
        if self.deterministic:
            form |= _lib.ASM_DETERMINISTIC
        _lib.check(_lib.lib().tm_asm_p1_matrix(
            form, self.topology.nb, self.topology.dim, plan.struct, _lib.ptr(points64), _lib.ptr(coef), cdt,
            float(scale), self._quad, _lib.ptr(values), _lib.dtype_code(values), _lib.stream_of(dev)))
        return values

    def _launch_general(self, points64, D: np.ndarray, out_dtype) -> torch.Tensor:
        dev = points64.device
        plan = self._plan(dev)
        values = torch.empty(self.topology.nnz, dtype=out_dtype, device=dev)
        Dc = np.ascontiguousarray(D, dtype=np.float64)
        _lib.check(_lib.lib().tm_asm_p1_general(
            self.topology.nb, self.topology.dim, plan.struct, _lib.ptr(points64), _lib.ptr(Dc), self._quad,
            _lib.ptr(values), _lib.dtype_code(values), _lib.ASM_DETERMINISTIC if self.deterministic else 0,
            _lib.stream_of(dev)))
        return values

    def _launch_moments(self, points64, mom: torch.Tensor, out_dtype) -> torch.Tensor:
        dev = points64.device
        plan = self._plan(dev)
        values = torch.empty(self.topology.nnz, dtype=out_dtype, device=dev)
        _lib.check(_lib.lib().tm_asm_p1_moments(
            self.topology.nb, self.topology.dim, plan.struct, _lib.ptr(points64), _lib.ptr(mom),
            _lib.ptr(values), _lib.dtype_code(values), _lib.ASM_DETERMINISTIC if self.deterministic else 0,
            _lib.stream_of(dev)))
        return values

    def _launch_kappa_adjoint(self, points64, g) -> torch.Tensor:
        dev = points64.device
        plan = self._plan(dev)
        g = g.contiguous()
        dk = torch.zeros(self.n_points, dtype=torch.float64, device=dev)
        _lib.check(_lib.lib().tm_asm_p1_kappa_adjoint(
            self.topology.nb, self.topology.dim, plan.struct, _lib.ptr(points64), _lib.ptr(g),
            _lib.dtype_code(g), self._quad, _lib.ptr(dk), _lib.stream_of(dev)))
        return dk

    def _classify(self, fn, point_data, scalar_data, element_data):
        key = forms.form_cache_key(fn, point_data, scalar_data, element_data)
        if key is not None and key in self._form_cache:
            return self._form_cache[key]
        r = forms.classify_bilinear(fn, self.dimension, point_data, scalar_data, element_data)
        if key is not None:
            self._form_cache[key] = r
        return r

    # ---- call -------------------------------------------------------------------
    def __call__(self, points: Optional[torch.Tensor] = None, func: Optional[Callable] = None,
                 point_data: Optional[Mapping[str, torch.Tensor]] = None,
                 element_data: Optional[Union[Mapping[str, Mapping[str, torch.Tensor]], Mapping[str, torch.Tensor]]] = None,
                 scalar_data: Optional[Mapping[str, torch.Tensor]] = None, batch_size: int = -1) -> SparseMatrix:
        assert isinstance(point_data, dict) or point_data is None, \
            f"point_data should be a dict, but got {type(point_data)}. Please pass  in extra parameter using key-value pairs"
        point_data = dict(point_data or {})
        etype = self.element_types[0]
        conn = self.elements[etype]
        if element_data is None:
            element_data = {}
        else:
            element_data = dict(element_data)
            if element_data and isinstance(next(iter(element_data.values())), dict):
                element_data = {k: v[etype] for k, v in element_data.items()}
            for k, v in element_data.items():
                assert v.shape[0] == conn.shape[0], \
                    f"the shape of {k} should be [{conn.shape[0]}, ...], but got {v.shape[0]}"
        scalar_data = dict(scalar_data or {})
        if points is None:
            points = self.points
        else:
            assert points.shape[1] == self.dimension, \
                f"the dimension of points should be {self.dimension}, but got {points.shape[1]}"
        point_data["x"] = points
        for key, value in point_data.items():
            assert value.shape[0] == points.shape[0], \
                f"the shape of {key} should be [n_point, ...], but got {value.shape}"
        fn = self.forward if func is None else func
        dev = points.device
        if self.edges.device != dev:
            self.to(dev)

        form = None
        on_kernel = dev.type == "cuda" and not points.requires_grad
        if on_kernel:
            pd = {k: v for k, v in point_data.items() if k != "x"}
            form = self._classify(fn, pd, scalar_data, element_data)
        if form is None and on_kernel and not any(
                torch.is_tensor(v) and v.requires_grad
                for v in list(point_data.values()) + list(element_data.values()) + list(scalar_data.values())):
            # coefficients that vary in space: per-element moments of the
            # integrand, then the moments kernel
            q = forms.probe_bilinear_q(fn, points, conn.to(dev), etype, self.quadrature_order, point_data,
                                       element_data, scalar_data)
            if q is not None:
                mom, sym = q
                values = self._launch_moments(points.detach().to(torch.float64).contiguous(), mom,
                                              self._out_dtype(points))
                e = self.topology.edges_on(dev)
                return SparseMatrix._from_csr(values, e[0], e[1], (self.n_points, self.n_points),
                                              self.csr(dev), symmetric=sym)
        if form is not None:
            kind, scale, key = form
            p64 = points.detach().to(torch.float64).contiguous()
            if kind == _lib.FORM_GENERAL:
                values = self._launch_general(p64, scale, self._out_dtype(points))
                e = self.topology.edges_on(dev)
                return SparseMatrix._from_csr(values, e[0], e[1], (self.n_points, self.n_points),
                                              self.csr(dev), symmetric=bool(np.array_equal(scale, scale.T)))
            if kind == _lib.FORM_KAPPA_LAPLACE:
                coef = point_data[key].to(dev)
                if coef.requires_grad:
                    values = _KappaLaplace.apply(coef, self, p64, scale, self._out_dtype(points))
                else:
                    values = self._launch_matrix(kind, p64, coef, scale, self._out_dtype(points))
            else:
                values = self._launch_matrix(kind, p64, None, scale, self._out_dtype(points))
            e = self.topology.edges_on(dev)
            return SparseMatrix._from_csr(values, e[0], e[1], (self.n_points, self.n_points),
                                          self.csr(dev), symmetric=True)
        return self._generic(fn, points, conn, point_data, element_data, scalar_data)

    def energy(self, points: Optional[torch.Tensor] = None, func: Optional[Callable] = None,
               point_data: Optional[Mapping[str, torch.Tensor]] = None, element_data=None,
               scalar_data: Optional[Mapping[str, torch.Tensor]] = None, batch_size: int = -1):
        """Total energy ``sum_e sum_q w_q |J| psi`` (``element_assembler.py:468-680``),
        generic autograd-differentiable path; ``element_energy`` gives psi."""
        assert isinstance(point_data, dict) or point_data is None, "point_data should be a dict"
        point_data = dict(point_data or {})
        etype = self.element_types[0]
        element_data = dict(element_data or {})
        if element_data and isinstance(next(iter(element_data.values())), dict):
            element_data = {k: v[etype] for k, v in element_data.items()}
        points = self.points if points is None else points
        point_data["x"] = points
        fn = self.element_energy if func is None else func
        return forms.generic_energy(fn, points, self.elements[etype].to(points.device), etype,
                                    self.quadrature_order, point_data, element_data, dict(scalar_data or {}))

    def element_energy(self, **kwargs):
        raise NotImplementedError("element_energy is not implemented")

    def _generic(self, fn, points, conn, point_data, element_data, scalar_data) -> SparseMatrix:
        integ = forms.generic_element_integrals(fn, points, conn.to(points.device), self.element_types[0],
                                                self.quadrature_order, point_data, element_data, scalar_data,
                                                bilinear=True)
        slots = self.projector.to(points.device)
        E, nb = conn.shape
        n = self.n_points
        edges = self.topology.edges_on(points.device)
        if integ.dim() == 3:
            vals = torch.zeros(self.topology.nnz, dtype=integ.dtype, device=points.device)
            vals = vals.index_add(0, slots, integ.reshape(-1))
            return SparseMatrix._from_csr(vals, edges[0], edges[1], (n, n), self.csr(points.device)) \
                if points.is_cuda else SparseMatrix(vals, edges[0], edges[1], (n, n))
        if integ.dim() == 5:
            H = integ.shape[-1]
            vals = torch.zeros((self.topology.nnz, H, H), dtype=integ.dtype, device=points.device)
            vals = vals.index_add(0, slots, integ.reshape(E * nb * nb, H, H))
            return SparseMatrix.from_block_coo(vals, edges[0], edges[1], (n, n))
        raise Exception(f"the shape of integral is supposed to be  1D or 3D, but got {integ.shape}")

    def __repr__(self):
        return (f"{self.__class__.__name__}(element_types={self.element_types}, n_points={self.n_points}, "
                f"n_elements={self.topology.n_elements}, nnz={self.topology.nnz})")
