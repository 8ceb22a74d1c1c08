"""Mesh topology shared by the assemblers: CSR pattern, element slots, block plans.

The pattern is the canonical CSR of all element node pairs (the reference's
``edges``, ``tensormesh/assemble/element_assembler.py:838-857``), built on the
host by the native library.  A *block plan* is the precomputed form of the
reference's projector ``P_E`` (``:858-884``) for the owner-computes kernels:
nodes are grouped into compact blocks by recursive coordinate bisection and
each block lists the elements touching its rows with the in-row position of
every element entry.  Plans depend only on the connectivity (the
coordinates steer the partition, never the result), so one topology serves
every assembler built on the same mesh.
"""
from __future__ import annotations

import ctypes
import os
from typing import Dict, Optional, Tuple

import numpy as np
import torch

from .. import _lib
from ..element import ELEMENT_DIM, ELEMENT_NB, quadrature_moments

# Plan kinds: (accumulator bytes per CSR entry, LDS bytes per halo node,
# LDS bytes per owned row).  "scalar" serves the matrix, load and adjoint
# kernels of scalar problems; "vector3" the 3x3-block kernels that accumulate
# CSR blocks (neo-Hookean tangent and residual, linear-elastic stiffness);
# "vector3a" the matrix-free tangent action, which stages coordinates,
# displacements and the multiplier per halo node (72 bytes) and three doubles
# per owned row, and keeps no CSR accumulators.
PLAN_KINDS = {"scalar": (8, 36, 8), "vector3": (72, 56, 32), "vector3a": (0, 80, 32)}


def _default_plan_params(kind: str, dim: int) -> Tuple[int, int, int]:
    """(max owned rows per block, LDS budget in bytes, lanes per block).

    3D scalar: 80 rows in 20 KiB of LDS on 256-lane workgroups measured
    fastest on MI355X for the 128^3 Kuhn kappa-Laplace matrix (233.5 / 234.0
    us against 242.6 / 238.1 us for 160 rows in 40 KiB on 512 lanes, the
    earlier default; 64 / 72 / 88 / 96 / 120 rows: 235-275 us; 224-512 rows
    and 48-80 KiB: 250-430 us).  Eight such blocks fit one CU: blocks in
    different phases (gathering their nodes, evaluating elements, writing
    rows) overlap, which outweighs the 9 % more element records.

    3x3 blocks (64^3 neo-Hookean): 64 rows in 64 KiB; the tangent / residual
    and the linear-elastic stiffness ask for plans of at most 56 rows in
    40 KiB explicitly (faster there).  The matrix-free tangent action without
    CSR accumulators: 80 rows in 24 KiB (128.5 us at 64^3; 40 / 48 / 56 / 64
    rows in 14-20 KiB: 138.2 / 135.0 / 135.1 / 130.8 us; 96 / 128 / 192 / 256
    rows: 130.9 / 131.3 / 132.3 / 138.2 us; 153.9 us on the 64-row 3x3-block
    plan it used before): its 178 VGPRs allow two workgroups per CU whatever
    their LDS, and larger blocks leave fewer of them for the end of the grid.
    """
    threads = os.environ.get("TM_ASM_THREADS")
    if kind == "vector3a":
        return (int(os.environ.get("TM_ASM_MAX_ROWS_APPLY", 80)), 24 * 1024, 256)
    if kind == "vector3":
        return (int(os.environ.get("TM_ASM_MAX_ROWS_VEC", 64)),
                int(os.environ.get("TM_ASM_LDS_BUDGET", 64 * 1024 - 64)), int(threads or 512))
    if dim == 3:
        return (int(os.environ.get("TM_ASM_MAX_ROWS", 80)),
                int(os.environ.get("TM_ASM_LDS_BUDGET", 20 * 1024)), int(threads or 256))
    return (int(os.environ.get("TM_ASM_MAX_ROWS_2D", 640)),
            int(os.environ.get("TM_ASM_LDS_BUDGET", 64 * 1024 - 64)), int(threads or 512))


class DevicePlan:
    """Device copy of an exported block plan (keeps the tensors alive)."""

    def __init__(self, host: Dict[str, np.ndarray], info: np.ndarray, device: torch.device, threads: int = 512,
                 entry_bytes: int = 0):
        self.tensors = {k: torch.from_numpy(v).to(device) for k, v in host.items()}
        self.info = info
        self.edge_refused: Optional[str] = None   # why edge records were not built (limits)
        t = self.tensors
        self.struct = _lib.PlanDev(
            t["blk"].data_ptr(), t["halo"].data_ptr(), t["meta"].data_ptr(), t["recs"].data_ptr(),
            t["rec_elem"].data_ptr(), t["runs"].data_ptr(), int(info[0]), int(info[4]), int(threads), int(entry_bytes))

    @property
    def n_blocks(self) -> int:
        return int(self.info[0])

    @property
    def has_edge(self) -> bool:
        return "erecs" in self.tensors

    def attach_edge(self, host: Dict[str, np.ndarray], info: np.ndarray) -> None:
        """Adds the plan's edge records (:func:`edge_plan`): the matrix kernel
        of the symmetric scalar forms then reads them instead of the 16-byte
        records."""
        dev = self.tensors["blk"].device
        for k in ("erecs", "eblk", "emir", "erows"):
            self.tensors[k] = torch.from_numpy(host[k]).to(dev)
        self.edge_info = info
        t = self.tensors
        self.struct.erecs = t["erecs"].data_ptr()
        self.struct.eblk = t["eblk"].data_ptr()
        self.struct.emir = t["emir"].data_ptr()
        self.struct.erows = t["erows"].data_ptr()
        self.struct.erec_words = int(info[0])
        self.struct.elds_bytes = int(info[4])

    def bytes(self) -> int:
        return sum(v.numel() * v.element_size() for v in self.tensors.values())


def edge_plan(nb: int, arrs: Dict[str, np.ndarray], info: np.ndarray):
    """Edge records and mirror lists of an exported scalar plan
    (``tm_plan_edge_build``): ``({"erecs", "eblk", "emir", "erows"}, info)``
    with info = (words per record, mirror entries, max per block, max rank,
    LDS bytes per block, node-word bits, owned rows, 0)."""
    L = _lib.lib()
    nblk, nrec = int(info[0]), int(info[2])
    h = ctypes.c_void_p()
    _lib.check(L.tm_plan_edge_build(nb, nblk, _lib.ptr(arrs["blk"]), _lib.ptr(arrs["meta"]), _lib.ptr(arrs["recs"]),
                                    _lib.ptr(arrs["runs"]), 0, ctypes.byref(h)))
    try:
        einfo = np.zeros(8, dtype=np.int64)
        _lib.check(L.tm_plan_edge_info(h, _lib.ptr(einfo)))
        out = dict(erecs=np.empty(nrec * int(einfo[0]), np.uint32), eblk=np.empty(nblk * 4, np.int32),
                   emir=np.empty(max(1, int(einfo[1])), np.uint32),
                   erows=np.empty(max(1, 2 * int(einfo[6])), np.int32))
        _lib.check(L.tm_plan_edge_export(h, _lib.ptr(out["erecs"]), _lib.ptr(out["eblk"]), _lib.ptr(out["emir"]),
                                         _lib.ptr(out["erows"])))
    finally:
        L.tm_plan_edge_destroy(h)
    return out, einfo


def edge_records_enabled() -> bool:
    """Edge records for the symmetric matrix kernels (``TM_ASM_EDGE=0`` keeps
    the 16-byte records, e.g. to compare the two)."""
    return os.environ.get("TM_ASM_EDGE", "1") != "0"


class P1Topology:
    """Pattern and plans of one P1 simplex connectivity (``triangle`` or ``tetra``).

    With ``device_conn`` (the connectivity on a GPU) the canonical pattern is
    built on that GPU (:func:`device_pattern`) and the matrices' ``edges`` are
    formed there; the host receives a copy of ``row_ptr`` / ``col_idx`` only
    when a block plan is built (the plan builder runs on the host)."""

    def __init__(self, points: np.ndarray, conn: np.ndarray, n_points: int, element_type: str,
                 device_conn: Optional[torch.Tensor] = None):
        if element_type not in ELEMENT_NB:
            raise NotImplementedError(f"element type {element_type!r}: the HIP path covers P1 "
                                      f"'triangle' and 'tetra' meshes")
        self.element_type = element_type
        self.nb = ELEMENT_NB[element_type]
        self.dim = ELEMENT_DIM[element_type]
        self.n_points = int(n_points)
        self.conn = np.ascontiguousarray(conn, dtype=np.int64)
        assert self.conn.ndim == 2 and self.conn.shape[1] == self.nb, \
            f"connectivity must be [n_elements, {self.nb}], got {self.conn.shape}"
        self.n_elements = self.conn.shape[0]
        self.points = np.ascontiguousarray(points, dtype=np.float64)
        self._host_plans: Dict[tuple, tuple] = {}
        self._dev_plans: Dict[tuple, DevicePlan] = {}
        self._dev_csr: Dict[torch.device, Tuple[torch.Tensor, torch.Tensor]] = {}
        self._dev_edges: Dict[torch.device, torch.Tensor] = {}
        self._slots: Optional[np.ndarray] = None
        self._dev_slots: Dict[torch.device, torch.Tensor] = {}
        self._row_ptr = self._col_idx = self._edges = None
        self.pattern_device: Optional[torch.device] = None
        if device_conn is not None and device_conn.is_cuda:
            rp_d, ci_d, work = device_pattern(device_conn, self.n_points)
            dev = device_conn.device
            self.pattern_device = dev
            self._dev_csr[dev] = (rp_d, ci_d)
            self._dev_conn, self._dev_work = device_conn.detach().to(torch.int64).contiguous(), work
            self.nnz = int(ci_d.numel())
            return
        L = _lib.lib()
        rp = np.empty(self.n_points + 1, dtype=np.int32)
        _lib.check(L.tm_pattern_rowptr(_lib.ptr(self.conn), self.n_elements, self.nb, self.n_points,
                                       _lib.ptr(rp)))
        ci = np.empty(int(rp[-1]), dtype=np.int32)
        _lib.check(L.tm_pattern_colidx(_lib.ptr(self.conn), self.n_elements, self.nb, self.n_points,
                                       _lib.ptr(rp), _lib.ptr(ci)))
        self._row_ptr, self._col_idx = rp, ci
        self.nnz = int(rp[-1])

    @property
    def row_ptr(self) -> np.ndarray:
        if self._row_ptr is None:
            self._row_ptr = self._dev_csr[self.pattern_device][0].cpu().numpy()
        return self._row_ptr

    @property
    def col_idx(self) -> np.ndarray:
        if self._col_idx is None:
            self._col_idx = self._dev_csr[self.pattern_device][1].cpu().numpy()
        return self._col_idx

    @property
    def edges(self) -> torch.Tensor:
        """int64 ``[2, nnz]`` edges on the host."""
        if self._edges is None:
            if self.pattern_device is not None:
                self._edges = self.edges_on(self.pattern_device).cpu()
            else:
                row = np.repeat(np.arange(self.n_points, dtype=np.int64), np.diff(self.row_ptr))
                self._edges = torch.from_numpy(np.stack([row, self.col_idx.astype(np.int64)], 0))
        return self._edges

    # ---- pattern ------------------------------------------------------------
    def edges_on(self, device) -> torch.Tensor:
        """int64 ``[2, nnz]`` edges on ``device``, one shared tensor per device so every
        matrix assembled on this mesh carries the same row/col storage."""
        device = torch.device(device)
        if device.type == "cpu":
            return self.edges
        if device not in self._dev_edges:
            if device == self.pattern_device:
                rp, ci = self._dev_csr[device]
                row = torch.repeat_interleave(torch.arange(self.n_points, device=device),
                                              (rp[1:] - rp[:-1]).long(), output_size=self.nnz)
                self._dev_edges[device] = torch.stack([row, ci.long()], 0)
            else:
                self._dev_edges[device] = self.edges.to(device)
        return self._dev_edges[device]

    def csr(self, device) -> Tuple[torch.Tensor, torch.Tensor]:
        device = torch.device(device)
        if device not in self._dev_csr:
            self._dev_csr[device] = (torch.from_numpy(self.row_ptr).to(device),
                                     torch.from_numpy(self.col_idx).to(device))
        return self._dev_csr[device]

    def element_slots_on(self, device) -> torch.Tensor:
        """``P_E`` on ``device`` (built there when the pattern was built on it)."""
        device = torch.device(device)
        if device == self.pattern_device:
            if device not in self._dev_slots:
                rp, ci = self._dev_csr[device]
                self._dev_slots[device] = device_element_slots(self._dev_conn, rp, ci, self._dev_work,
                                                               self.n_points)
            return self._dev_slots[device]
        return torch.from_numpy(self.element_slots()).to(device)

    def element_slots(self) -> np.ndarray:
        """The reference's ``P_E``: int64 ``[E * nb * nb]``, element-major then (i, j)."""
        if self._slots is None and self.pattern_device is not None:
            self._slots = self.element_slots_on(self.pattern_device).cpu().numpy()
        if self._slots is None:
            s = np.empty(self.n_elements * self.nb * self.nb, dtype=np.int64)
            _lib.check(_lib.lib().tm_element_slots(_lib.ptr(self.conn), self.n_elements, self.nb,
                                                   _lib.ptr(self.row_ptr), _lib.ptr(self.col_idx),
                                                   _lib.ptr(s)))
            self._slots = s
        return self._slots

    # ---- plans ----------------------------------------------------------------
    def host_plan(self, kind: str = "scalar", max_rows: Optional[int] = None,
                  lds_budget: Optional[int] = None):
        rows0, budget0, _ = _default_plan_params(kind, self.dim)
        rows = int(max_rows or rows0)
        budget = int(lds_budget or budget0)
        key = (kind, rows, budget)
        if key not in self._host_plans:
            eb, nbytes, ob = PLAN_KINDS[kind]
            L = _lib.lib()
            h = ctypes.c_void_p()
            _lib.check(L.tm_plan_build(_lib.ptr(self.points), self.dim, _lib.ptr(self.conn), self.n_elements,
                                       self.nb, self.n_points, _lib.ptr(self.row_ptr), _lib.ptr(self.col_idx),
                                       rows, budget, eb, nbytes, ob, 0, ctypes.byref(h)))
            try:
                info = np.zeros(8, dtype=np.int64)
                _lib.check(L.tm_plan_info(h, _lib.ptr(info)))
                nblk, nh, nr, nrun = (int(x) for x in info[:4])
                arrs = dict(blk=np.empty(nblk * 8, np.int32), halo=np.empty(nh, np.int32),
                            meta=np.empty(nh, np.uint32), recs=np.empty(nr * 4, np.uint32),
                            rec_elem=np.empty(nr, np.int32), runs=np.empty(nrun * 4, np.int32))
                _lib.check(L.tm_plan_export(h, *(_lib.ptr(arrs[k]) for k in
                                                  ("blk", "halo", "meta", "recs", "rec_elem", "runs"))))
            finally:
                L.tm_plan_destroy(h)
            self._host_plans[key] = (arrs, info)
        return self._host_plans[key]

    def device_plan(self, device, kind: str = "scalar", max_rows: Optional[int] = None,
                    lds_budget: Optional[int] = None, edge: bool = False) -> DevicePlan:
        """Device copy of a plan; ``edge=True`` (scalar plans) also attaches
        its edge records when :func:`edge_records_enabled`."""
        device = torch.device(device)
        rows0, budget0, threads = _default_plan_params(kind, self.dim)
        key = (device, kind, int(max_rows or rows0), int(lds_budget or budget0), threads)
        if key not in self._dev_plans:
            arrs, info = self.host_plan(kind, max_rows, lds_budget)
            # the plan states its accumulator bytes per entry (-1: none), which the 3x3-block
            # tangent and stiffness calls check
            self._dev_plans[key] = DevicePlan(arrs, info, device, threads, PLAN_KINDS[kind][0] or -1)
        plan = self._dev_plans[key]
        if edge and kind == "scalar" and not plan.has_edge and edge_records_enabled() and not plan.edge_refused:
            arrs, info = self.host_plan(kind, max_rows, lds_budget)
            try:
                plan.attach_edge(*edge_plan(self.nb, arrs, info))
            except _lib.TMError as err:
                if err.code != _lib.TM_ERR_LIMIT:
                    raise
                # a block outside the edge records' limits (accumulators, records or LDS):
                # the matrix kernel keeps the 16-byte records
                plan.edge_refused = str(err)
        return plan

    def quad(self, order: int) -> _lib.Quad:
        wsum, W, M = quadrature_moments(self.element_type, order)
        q = _lib.Quad()
        q.wsum = wsum
        for i, v in enumerate(W):
            q.W[i] = float(v)
        nb = self.nb
        for a in range(nb):
            for b in range(nb):
                q.M[a * nb + b] = float(M[a, b])
        return q


_KEEP: Dict[tuple, Tuple[P1Topology, torch.Tensor]] = {}


def topology_for(points: torch.Tensor, conn: torch.Tensor, element_type: str) -> P1Topology:
    """Topology of a connectivity tensor, shared by every assembler built on it."""
    key = (conn.data_ptr(), tuple(conn.shape), str(conn.device), conn._version, points.shape[0])
    hit = _KEEP.get(key)
    if hit is None:
        topo = P1Topology(points.detach().cpu().double().numpy(), conn.detach().cpu().numpy(),
                          points.shape[0], element_type, device_conn=conn if conn.is_cuda else None)
        # the cache holds the connectivity tensor so its address stays unique
        _KEEP[key] = (topo, conn)
        return topo
    return hit[0]


def clear_topology_cache():
    _KEEP.clear()


def device_pattern(conn: torch.Tensor, n_points: int):
    """Canonical CSR pattern of device connectivity ``conn [E, nb]`` built on
    the GPU (``tm_pattern_dev_*``): ``(row_ptr, col_idx, work)`` with int32
    ``row_ptr [N+1]`` / ``col_idx [nnz]``, bit-identical to the host pattern
    (``element_assembler.py:838-857``); ``work`` serves
    :func:`device_element_slots`.  One host synchronisation (the size of
    ``col_idx``)."""
    _lib.require_gpu(conn, "connectivity")
    c = conn.detach().to(torch.int64).contiguous()
    E, nb = int(c.shape[0]), int(c.shape[1])
    L = _lib.lib()
    dev, st = c.device, _lib.stream_of(c.device)
    work = torch.empty(max(1, int(L.tm_pattern_dev_work_size(E, nb, n_points))), dtype=torch.uint8, device=dev)
    row_ptr = torch.empty(n_points + 1, dtype=torch.int32, device=dev)
    _lib.check(L.tm_pattern_dev_rows(_lib.ptr(c), E, nb, n_points, _lib.ptr(work), _lib.ptr(row_ptr), st))
    nnz = int(row_ptr[n_points].item())
    col_buf = torch.empty(max(1, nnz), dtype=torch.int32, device=dev)
    _lib.check(L.tm_pattern_dev_cols(_lib.ptr(c), E, nb, n_points, _lib.ptr(work), _lib.ptr(row_ptr),
                                     _lib.ptr(col_buf), st))
    return row_ptr, col_buf[:nnz], work


def device_element_slots(conn: torch.Tensor, row_ptr: torch.Tensor, col_idx: torch.Tensor, work: torch.Tensor,
                         n_points: int) -> torch.Tensor:
    """``P_E`` on the GPU (``element_assembler.py:858-866``): int64
    ``[E * nb * nb]`` positions of every element pair in the pattern."""
    c = conn.detach().to(torch.int64).contiguous()
    E, nb = int(c.shape[0]), int(c.shape[1])
    slots = torch.empty(max(1, E * nb * nb), dtype=torch.int64, device=c.device)
    _lib.check(_lib.lib().tm_element_slots_dev(_lib.ptr(c), E, nb, _lib.ptr(row_ptr), _lib.ptr(col_idx),
                                               _lib.ptr(work), n_points, _lib.ptr(slots), _lib.stream_of(c.device)))
    return slots[:E * nb * nb]
