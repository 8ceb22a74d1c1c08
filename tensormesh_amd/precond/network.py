"""U-shaped edge message-passing network used as a preconditioner.

Inputs per level-0 node: the Jacobi-scaled residual ``d r`` (normalised by
its RMS ``s``) and a log-coefficient feature ``lk``; output the correction
``z = s d y``.  Layers (``schedule`` lists the level of each of the six
message-passing layers, default ``(1, 2, 3, 2, 1, 0)``):

* linear encoder ``h = [d r / s, lk] We + be`` at level 0;
* before a layer on a coarser level the features are pooled level by level
  (``h_c[I] = sum_i w_i h[i]``, ``w_i = sqrt(A_ii / Ac_II)``), and the
  features left behind are kept as that level's skip;
* before a layer on a finer level they are unpooled level by level and added
  to the skip (``h[i] = skip[i] + w_i h_c[agg_i]``);
* message passing ``h' = h + act([h | A^ h] W + b)`` on the level's
  normalised Galerkin operator ``A^``;
* decoder ``y = h wd + bd`` at level 0.

When the schedule starts below level 0 and ends on it, inference fuses the
finest level: encode + first pooling in one kernel, and unpool (with the
encoder output as the skip) + the level-0 layer + decode in another.  With
the linear encoder the level-0 messages reduce to two scalar products over
the CSR row plus ``B = A^ P^`` applied to coarse features, so no 64-wide
level-0 tensor is read or written.  Training runs the unfused kernels (which
save what the backward pass needs); the two agree up to BF16 rounding of the
level-0 features.

Every kernel is a HIP kernel of ``libtensormesh_amd.so`` (forward, backward,
loss and Adam); torch only holds the buffers.  Weights live as an FP32 master
vector; the message-passing layers read BF16 copies.
"""
from __future__ import annotations

import ctypes
from typing import Dict, List, Optional, Sequence

import numpy as np
import torch

from .. import _lib
from .._lib import check, ptr
from .hierarchy import Hierarchy

F = 64


def _addr(x) -> ctypes.c_void_p:
    """Device address of a tensor, a raw address (int) or NULL."""
    if x is None:
        return ctypes.c_void_p(None)
    if isinstance(x, int):
        return ctypes.c_void_p(x)
    return ctypes.c_void_p(x.data_ptr())


class GNNPreconditioner:
    """Weights and kernels of the network (no torch autograd involved)."""

    def __init__(self, schedule: Sequence[int] = (1, 2, 3, 3, 2, 1), n_linear: int = 64,
                 device: torch.device = torch.device("cuda:0"), seed: int = 0, homogeneous: bool = True,
                 symmetric: bool = True, bottom: Optional[int] = None):
        self.schedule = tuple(int(s) for s in schedule)
        # levels below the deepest layer that still contribute a pooled term
        # (between the down and the up path, at layer index ``turn``)
        self.bottom = max(self.schedule) if bottom is None else int(bottom)
        self.turn = len(self.schedule) // 2
        self.n_linear = int(n_linear)
        self.symmetric = bool(symmetric)
        self.homogeneous = bool(homogeneous) or self.symmetric
        if self.symmetric:
            L = len(self.schedule)
            assert L % 2 == 0 and self.schedule == self.schedule[::-1] and self.n_linear == F, \
                "a symmetric network needs a palindromic schedule of even length and identity activations"
        self.device = torch.device(device)
        nl = len(self.schedule)
        self.sizes = [3 * F] + [2 * F * F + F] * nl + [F + 1]
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)]).astype(np.int64)
        self.n_params = int(self.offsets[-1])
        self.params = torch.zeros(self.n_params, dtype=torch.float32, device=self.device)
        self.mask = self._mask()
        self.init_weights(seed)
        self._adam = None
        self._step = 0
        self._part = None

    # ---- parameters --------------------------------------------------------
    def _view(self, k: int) -> torch.Tensor:
        return self.params[int(self.offsets[k]):int(self.offsets[k + 1])]

    def enc(self):
        v = self._view(0)
        return v[:2 * F], v[2 * F:]

    def layer(self, t: int):
        v = self._view(1 + t)
        return v[:2 * F * F].view(2 * F, F), v[2 * F * F:]

    def dec(self):
        v = self._view(1 + len(self.schedule))
        return v[:F], v[F:]

    @staticmethod
    def block_t(W: torch.Tensor) -> torch.Tensor:
        """Weights of the adjoint layer: [W_h; W_n] -> [W_h^T; W_n^T]
        (the layer h -> h + h W_h + (A^ h) W_n with A^ symmetric)."""
        return torch.cat([W[:F].t(), W[F:].t()], 0)

    def mirror(self, t: int) -> int:
        """Free layer whose adjoint layer ``t`` is (``t`` itself if it is free)."""
        L = len(self.schedule)
        return L - 1 - t if (self.symmetric and t >= L // 2) else t

    def _mask(self) -> torch.Tensor:
        """1 for trainable entries.  A homogeneous network keeps every bias and
        the log-coefficient input weights at zero: then M(a r) = a M(r), and with
        identity activations M is linear in r for a fixed matrix (the
        coefficient still enters through the normalised matrix entries), which
        is what flexible CG converges best with."""
        m = torch.ones(self.n_params, dtype=torch.float32)
        if self.homogeneous:
            o = self.offsets
            m[int(o[0]) + F:int(o[1])] = 0.0                      # We1 (log kappa) and be
            for t in range(len(self.schedule)):
                m[int(o[1 + t]) + 2 * F * F:int(o[2 + t])] = 0.0   # layer biases
            m[int(o[-1]) - 1] = 0.0                                # decoder bias
        if self.symmetric:
            L = len(self.schedule)
            for t in range(L // 2, L):                             # adjoint layers: derived
                m[int(o[1 + t]):int(o[2 + t])] = 0.0
            m[int(o[-2]):int(o[-1])] = 0.0                         # decoder = encoder^T
        return m.to(self.device)

    def init_weights(self, seed: int = 0, noise: float = 0.05):
        """Additive multilevel start.  Channel 0 carries the scaled residual
        through every layer untouched (encoder weight 1, zero weights into it)
        and the decoder reads it, so the untrained network applies
        ``I + P0 P0^T + P0 P1 P1^T P0^T + ...`` in normalised coordinates (a
        BPX-type additive aggregation preconditioner, symmetric positive
        definite).  The other channels start random (``noise``) and only reach
        the output once training moves the weights into channel 0."""
        g = np.random.default_rng(seed)
        p = np.zeros(self.n_params, dtype=np.float32)
        o = self.offsets
        we = g.standard_normal((2, F)) * 0.5
        we[:, 0] = 0.0
        we[0, 0] = 1.0
        p[o[0]:o[0] + 2 * F] = we.ravel()
        for t in range(len(self.schedule)):
            a = int(o[1 + t])
            W = g.standard_normal((2 * F, F)) * (noise / np.sqrt(2 * F))
            W[:, 0] = 0.0
            p[a:a + 2 * F * F] = W.ravel()
        a = int(o[1 + len(self.schedule)])
        p[a] = 1.0
        self.params.copy_(torch.from_numpy(p).to(self.device) * self.mask)
        self.refresh()

    def refresh(self):
        """Derived weights (adjoint layers, tied decoder) and the BF16 copies for
        the kernels (W^T for the forward, W for the backward).

        A symmetric network applies M = sum_l R_l^T R_l with R_0 = E (encoder)
        and R_l = L_l P_{l-1}^T R_{l-1}: the second half of the schedule runs
        the adjoints of the first half's layers, pooling and unpooling are
        adjoint, and the decoder is the encoder's transpose, so M is symmetric
        positive definite whatever the weights."""
        if self.symmetric:
            with torch.no_grad():
                for t in range(len(self.schedule) // 2, len(self.schedule)):
                    W, b = self.layer(t)
                    W.copy_(self.block_t(self.layer(self.mirror(t))[0]))
                    b.zero_()
                wd, bd = self.dec()
                wd.copy_(self.enc()[0][:F])
                bd.zero_()
        self.wt, self.w = [], []
        for t in range(len(self.schedule)):
            W, _ = self.layer(t)
            wb = W.to(torch.bfloat16)
            self.w.append(wb.contiguous())
            self.wt.append(wb.t().contiguous())
        # level-0 layer of a symmetric network on the rank-one encoder output:
        # L0(x (x) we) = x (x) a + (A^x) (x) b
        self.sym0 = None
        if self.symmetric and self.schedule[0] == 0:
            we = self.enc()[0][:F].double()
            W0 = self.w[0].double()
            a = we + we @ W0[:F]
            b = we @ W0[F:]
            self.sym0 = {"ab": torch.cat([a, b]).float().contiguous(),
                         "zero": torch.zeros(F, dtype=torch.float32, device=self.device),
                         "c3": torch.stack([a @ a, a @ b, b @ b]).float().contiguous()}

    def fold(self, grad: torch.Tensor):
        """Move the gradient of derived weights onto the weights they derive from."""
        if not self.symmetric:
            return
        o = self.offsets
        view = lambda k: grad[int(o[k]):int(o[k + 1])]
        for t in range(len(self.schedule) // 2, len(self.schedule)):
            gt = view(1 + t)[:2 * F * F].view(2 * F, F)
            view(1 + self.mirror(t))[:2 * F * F].view(2 * F, F).add_(self.block_t(gt))
        view(0)[:F].add_(view(1 + len(self.schedule))[:F])

    def state_dict(self) -> Dict[str, np.ndarray]:
        return {"params": self.params.detach().cpu().numpy(), "schedule": np.array(self.schedule),
                "n_linear": np.array(self.n_linear), "homogeneous": np.array(int(self.homogeneous)),
                "symmetric": np.array(int(self.symmetric)), "bottom": np.array(self.bottom)}

    def save(self, path: str):
        np.savez(path, **self.state_dict())

    @classmethod
    def load(cls, path: str, device: torch.device = torch.device("cuda:0")) -> "GNNPreconditioner":
        z = np.load(path, allow_pickle=False)
        hom = bool(int(z["homogeneous"])) if "homogeneous" in z.files else False
        sym = bool(int(z["symmetric"])) if "symmetric" in z.files else False
        bottom = int(z["bottom"]) if "bottom" in z.files else None
        net = cls(tuple(int(s) for s in z["schedule"]), int(z["n_linear"]), device, homogeneous=hom, symmetric=sym,
                  bottom=bottom)
        net.params.copy_(torch.from_numpy(z["params"]).to(torch.float32))
        net.refresh()
        return net

    @property
    def n_levels(self) -> int:
        return max(max(self.schedule), self.bottom) + 1

    # ---- forward -----------------------------------------------------------
    def plan(self, hier: Hierarchy) -> "Program":
        assert hier.n_levels >= self.n_levels, \
            f"network needs {self.n_levels} levels, hierarchy has {hier.n_levels}"
        return Program(self, hier)


class Program:
    """Buffers and op list of the network on one hierarchy (reused every call,
    so a sequence of calls can be captured in a hipGraph)."""

    def __init__(self, net: GNNPreconditioner, hier: Hierarchy):
        self.net, self.hier = net, hier
        dev = hier.device
        self.n0 = hier.levels[0].n
        bf = lambda lev: torch.zeros((hier.levels[lev].n, F), dtype=torch.bfloat16, device=dev)
        # op list: (kind, level, in_id, out_id, extra)
        self.bufs: List[torch.Tensor] = []
        self.levels: List[int] = []
        self.ops = []

        def new(lev):
            self.bufs.append(bf(lev))
            self.levels.append(lev)
            return len(self.bufs) - 1

        cur, h = 0, new(0)
        self.enc_out = h
        skips: Dict[int, int] = {}
        for t, lev in enumerate(net.schedule):
            if t == net.turn:
                lev_down = max(lev, net.bottom)  # pooled-only levels below the last layer
                while cur < lev_down:
                    skips[cur] = h
                    o = new(cur + 1)
                    self.ops.append(("pool", cur, h, o, None))
                    h, cur = o, cur + 1
            while cur < lev:
                skips[cur] = h
                o = new(cur + 1)
                self.ops.append(("pool", cur, h, o, None))
                h, cur = o, cur + 1
            while cur > lev:
                o = new(cur - 1)
                self.ops.append(("unpool", cur - 1, h, o, skips.pop(cur - 1)))
                h, cur = o, cur - 1
            o = new(cur)
            self.ops.append(("mp", cur, h, o, t))
            h = o
        while cur > 0:
            o = new(cur - 1)
            self.ops.append(("unpool", cur - 1, h, o, skips.pop(cur - 1)))
            h, cur = o, cur - 1
        self.final = h
        ops = self.ops
        uses_enc = sum(1 for op in ops if op[2] == self.enc_out or (op[0] == "unpool" and op[4] == self.enc_out))
        self.fusable = (len(ops) >= 3 and ops[0][0] == "pool" and ops[0][1] == 0 and ops[0][2] == self.enc_out
                        and ops[-2][0] == "unpool" and ops[-2][1] == 0 and ops[-2][4] == self.enc_out
                        and ops[-1][0] == "mp" and ops[-1][1] == 0 and ops[-1][2] == ops[-2][3]
                        and ops[-1][3] == self.final and uses_enc == 2
                        and getattr(hier.levels[0], "b_row_ptr", None) is not None)
        # finest level with no layer: encode+pool ... unpool+decode
        self.fusable_plain = (len(ops) >= 2 and ops[0][0] == "pool" and ops[0][1] == 0
                              and ops[0][2] == self.enc_out and ops[-1][0] == "unpool" and ops[-1][1] == 0
                              and ops[-1][4] == self.enc_out and ops[-1][3] == self.final and uses_enc == 2)
        self.saves = {}
        self.y = torch.zeros(self.n0, dtype=torch.float32, device=dev)
        self.z = torch.zeros(self.n0, dtype=torch.float64, device=dev)
        self.scale = torch.zeros(1, dtype=torch.float64, device=dev)
        self.ss_part = torch.zeros(1024, dtype=torch.float64, device=dev)
        # symmetric network with its first and last layer on level 0 (rank-one encoder)
        self.fusable_sym0 = (net.symmetric and net.homogeneous and len(ops) >= 4 and ops[0][0] == "mp"
                             and ops[0][1] == 0 and ops[0][2] == self.enc_out and ops[1][0] == "pool"
                             and ops[1][1] == 0 and ops[1][2] == ops[0][3] and ops[-2][0] == "unpool"
                             and ops[-2][1] == 0 and ops[-2][4] == ops[0][3] and ops[-1][0] == "mp"
                             and ops[-1][1] == 0 and ops[-1][3] == self.final)
        self.xs = (torch.zeros((self.n0, 2), dtype=torch.float32, device=dev)
                   if (self.fusable or self.fusable_plain or self.fusable_sym0) else None)
        if self.fusable_sym0:
            n1 = hier.levels[1].n
            self.sym_ua = torch.zeros(n1, dtype=torch.float32, device=dev)
            self.sym_ub = torch.zeros(n1, dtype=torch.float32, device=dev)
            self.sym_t = torch.zeros(self.n0, dtype=torch.float32, device=dev)
            self.sym_dr = torch.zeros(self.n0, dtype=torch.float32, device=dev)
        self.coarse_dot = (torch.zeros(hier.levels[1].n, dtype=torch.float32, device=dev)
                           if self.fusable_plain else None)

    def alloc_training(self):
        dev = self.hier.device
        for i, op in enumerate(self.ops):
            if op[0] == "mp" and i not in self.saves:
                n = self.hier.levels[op[1]].n
                self.saves[i] = (torch.zeros((n, F), dtype=torch.bfloat16, device=dev),
                                 torch.zeros((n, F), dtype=torch.bfloat16, device=dev))

    def forward(self, r, lk: torch.Tensor, *, ss_part=None, nparts: int = 0, z=None, scale=None, stop=None,
                train: bool = False, want_y: bool = False, fused: bool = True, dr=None):
        """z = M(r).  Without ``ss_part`` the RMS partials are computed here.
        ``r``, ``z``, ``scale``, ``ss_part`` and ``stop`` may be tensors or raw
        device addresses (the flexible-CG work buffer)."""
        L = _lib.lib()
        net, hier = self.net, self.hier
        st = _lib.stream_of(hier.device)
        lv0 = hier.levels[0]
        stop_p = _addr(stop)
        if ss_part is None:
            npc = ctypes.c_int(0)
            check(L.tm_gnn_sumsq(self.n0, _addr(r), ptr(lv0.d), ptr(self.ss_part), ctypes.byref(npc), st))
            ss_ptr, nparts = ptr(self.ss_part), npc.value
        else:
            ss_ptr = _addr(ss_part)
        scale_p = ptr(self.scale) if scale is None else _addr(scale)
        we, be = net.enc()
        zt = self.z if z is None else z
        sym0_path = fused and self.fusable_sym0 and not train and not want_y
        if not sym0_path and not hier.level0_ready:
            raise RuntimeError("this network path needs the level-0 normalised values: call "
                               "Hierarchy.set_matrix(values) with level0_normalized=True")
        if sym0_path:
            if nparts > 0:   # scale from the partials (the flexible-CG kernels store it themselves)
                check(L.tm_gnn_scale(self.n0, ss_ptr, nparts, scale_p, st))
            lv1 = hier.levels[1]
            s0 = net.sym0
            # A's own values with d-scaled vectors (A^ = D A D): the same value array as the solver's SpMV
            check(L.tm_gnn_sym0_down(self.n0, ptr(lv0.row_ptr), ptr(lv0.col), ptr(lv0.values), _addr(r), ptr(lv0.d),
                                     ptr(self.sym_dr) if dr is None else _addr(dr), 0 if dr is None else 1,
                                     scale_p, ptr(self.xs), stop_p, st))
            check(L.tm_gnn_encode_pool(self.n0, lv1.n, ptr(lv1.child_ptr), ptr(lv1.child), ptr(lv0.w), _addr(None),
                                       ptr(lv0.d), _addr(None), _addr(None), 0, ptr(s0["ab"]), ptr(s0["zero"]),
                                       ptr(self.bufs[self.ops[1][3]]), ptr(self.xs), scale_p, stop_p, st))
            self._run_ops(self.ops[2:-2], 2, train, stop_p, st)
            check(L.tm_gnn_sym0_up(self.n0, lv1.n, ptr(lv0.row_ptr), ptr(lv0.col), ptr(lv0.values), ptr(self.xs),
                                   ptr(lv0.agg), ptr(lv0.w), ptr(self.bufs[self.ops[-2][2]]), ptr(s0["ab"]),
                                   ptr(s0["ab"][F:]), ptr(s0["c3"]), ptr(lv0.d), scale_p, ptr(self.sym_ua),
                                   ptr(self.sym_ub), ptr(self.sym_t), _addr(zt), stop_p, st))
            return zt
        if fused and self.fusable_plain and not train and not want_y:
            first, last = self.ops[0], self.ops[-1]
            lv1 = hier.levels[1]
            check(L.tm_gnn_encode_pool(self.n0, lv1.n, ptr(lv1.child_ptr), ptr(lv1.child), ptr(lv0.w), _addr(r),
                                       ptr(lv0.d), _addr(None) if net.homogeneous else ptr(lk), ss_ptr, nparts,
                                       ptr(we), ptr(be), ptr(self.bufs[first[3]]), ptr(self.xs), scale_p, stop_p,
                                       st))
            self._run_ops(self.ops[1:-1], 1, train, stop_p, st)
            wd, bd = net.dec()
            check(L.tm_gnn_unpool_decode(self.n0, hier.levels[1].n, ptr(self.xs), ptr(lv0.agg), ptr(lv0.w),
                                         ptr(self.bufs[last[2]]), ptr(we), ptr(be), ptr(wd), ptr(bd), ptr(lv0.d),
                                         scale_p, ptr(self.coarse_dot), _addr(zt), stop_p, st))
            return zt
        if fused and self.fusable and not train and not want_y:
            first, unp, last = self.ops[0], self.ops[-2], self.ops[-1]
            lv1 = hier.levels[1]
            check(L.tm_gnn_encode_pool(self.n0, lv1.n, ptr(lv1.child_ptr), ptr(lv1.child), ptr(lv0.w), _addr(r),
                                       ptr(lv0.d), ptr(lk), ss_ptr, nparts, ptr(we), ptr(be),
                                       ptr(self.bufs[first[3]]), ptr(self.xs), scale_p, stop_p, st))
            self._run_ops(self.ops[1:-2], 1, train, stop_p, st)
            _, bias = net.layer(last[4])
            wd, bd = net.dec()
            check(L.tm_gnn_final(self.n0, ptr(lv0.row_ptr), ptr(lv0.col), ptr(lv0.vn), ptr(self.xs),
                                 ptr(lv0.row_sums), ptr(lv0.agg), ptr(lv0.w), ptr(self.bufs[unp[2]]),
                                 ptr(lv0.b_row_ptr), ptr(lv0.b_col), ptr(lv0.b_values), ptr(we), ptr(be),
                                 ptr(net.wt[last[4]]), ptr(bias), net.n_linear, ptr(wd), ptr(bd), ptr(lv0.d), scale_p,
                                 _addr(zt), stop_p, st))
            return zt
        check(L.tm_gnn_encode(self.n0, _addr(r), ptr(lv0.d), ptr(lk), ss_ptr, nparts, ptr(we), ptr(be),
                              ptr(self.bufs[self.enc_out]), scale_p, stop_p, st))
        self._run_ops(self.ops, 0, train, stop_p, st)
        wd, bd = net.dec()
        check(L.tm_gnn_decode(self.n0, ptr(self.bufs[self.final]), ptr(wd), ptr(bd), ptr(lv0.d), scale_p,
                              ptr(self.y) if (train or want_y) else _addr(None), _addr(zt), stop_p, st))
        return zt

    def _run_ops(self, ops, first_index, train, stop_p, st):
        L = _lib.lib()
        net, hier = self.net, self.hier
        for i, (kind, lev, a, b, extra) in enumerate(ops, start=first_index):
            lv = hier.levels[lev]
            if kind == "mp":
                _, bias = net.layer(extra)
                sm, su = self.saves[i] if train else (None, None)
                check(L.tm_gnn_mp_forward(lv.n, ptr(lv.row_ptr), ptr(lv.col), ptr(lv.vn), ptr(self.bufs[a]),
                                          ptr(net.wt[extra]), ptr(bias), net.n_linear, ptr(self.bufs[b]),
                                          ptr(sm), ptr(su), stop_p, st))
            elif kind == "pool":
                nx = hier.levels[lev + 1]
                check(L.tm_gnn_pool(nx.n, ptr(nx.child_ptr), ptr(nx.child), ptr(lv.w), ptr(self.bufs[a]),
                                    ptr(self.bufs[b]), _lib.BF16, stop_p, st))
            else:  # unpool into level lev from lev + 1
                check(L.tm_gnn_unpool(lv.n, ptr(lv.agg), ptr(lv.w), ptr(self.bufs[a]), ptr(self.bufs[extra]),
                                      ptr(self.bufs[b]), _lib.BF16, stop_p, st))

    # ---- training ------------------------------------------------------------
    def loss_and_grad(self, r: torch.Tensor, lk: torch.Tensor, grad: torch.Tensor,
                      e: Optional[torch.Tensor] = None, gy_scale=None, loss: str = "energy") -> torch.Tensor:
        """Forward, loss and backward for one residual; accumulates the
        parameter gradient of ``gy_scale * L`` into ``grad`` (``gy_scale`` a
        float or a device scalar; default 1).  ``loss``: ``"energy"``
        ``(y.A^y/2 - y.r~)/n`` or ``"cosine"`` ``1 - (y.r~)^2/((y.A^y)(e~.r~))``
        (needs ``e``).  Returns the device sums [y.A^y, y.r~, e~.r~]."""
        L = _lib.lib()
        net, hier = self.net, self.hier
        st = _lib.stream_of(hier.device)
        self.alloc_training()
        self.forward(r, lk, train=True)
        lv0 = hier.levels[0]
        gy = torch.empty(self.n0, dtype=torch.float32, device=hier.device)
        part = torch.zeros(3 * 1024, dtype=torch.float64, device=hier.device)
        npc = ctypes.c_int(0)
        check(L.tm_gnn_energy_loss(self.n0, ptr(lv0.row_ptr), ptr(lv0.col), ptr(lv0.vn), ptr(self.y), ptr(r),
                                   ptr(lv0.d), ptr(self.scale), ptr(e), ptr(gy), ptr(part), ctypes.byref(npc), st))
        sums = part.view(3, 1024)[:, :npc.value].sum(1)
        if loss == "cosine":
            assert e is not None, "the cosine loss needs the error"
            check(L.tm_gnn_cosine_grad(self.n0, ptr(r), ptr(lv0.d), ptr(self.scale), ptr(part), npc.value, ptr(gy),
                                       st))
        else:
            assert loss == "energy", f"unknown loss {loss}"
        if gy_scale is not None:
            gy.mul_(gy_scale)
        if net._part is None:
            net._part = torch.empty(int(L.tm_gnn_part_floats()), dtype=torch.float32, device=hier.device)
        gview = lambda k: grad[int(net.offsets[k]):int(net.offsets[k + 1])]
        wd, _ = net.dec()
        grads: Dict[int, torch.Tensor] = {}
        g = torch.empty((self.n0, F), dtype=torch.float32, device=hier.device)
        check(L.tm_gnn_decode_backward(self.n0, ptr(self.bufs[self.final]), ptr(wd), ptr(gy), ptr(g), ptr(net._part),
                                       ptr(gview(1 + len(net.schedule))), st))
        grads[self.final] = g
        for i in range(len(self.ops) - 1, -1, -1):
            kind, lev, a, b, extra = self.ops[i]
            lv = hier.levels[lev]
            gout = grads.pop(b)
            if kind == "mp":
                sm, su = self.saves[i]
                gin = torch.empty_like(gout)
                gmsg = torch.empty_like(gout)
                check(L.tm_gnn_mp_backward(lv.n, ptr(lv.row_ptr), ptr(lv.col), ptr(lv.vn), ptr(gout),
                                           ptr(self.bufs[a]), ptr(sm), ptr(su), ptr(net.w[extra]), net.n_linear,
                                           ptr(gin), ptr(gmsg), ptr(net._part), ptr(gview(1 + extra)), st))
                self._accum(grads, a, gin)
            elif kind == "pool":
                # h_c = P^T h  ->  g_h (+)= P g_c
                nxt = hier.levels[lev + 1]
                if a in grads:
                    check(L.tm_gnn_unpool(lv.n, ptr(lv.agg), ptr(lv.w), ptr(gout), ptr(grads[a]), ptr(grads[a]),
                                          _lib.F32, None, st))
                else:
                    gin = torch.empty((lv.n, F), dtype=torch.float32, device=hier.device)
                    check(L.tm_gnn_unpool(lv.n, ptr(lv.agg), ptr(lv.w), ptr(gout), None, ptr(gin), _lib.F32,
                                          None, st))
                    grads[a] = gin
            else:
                # out = skip + P hc  ->  g_skip (+)= g_out, g_hc (+)= P^T g_out
                coarse = hier.levels[lev + 1]
                ghc = torch.empty((coarse.n, F), dtype=torch.float32, device=hier.device)
                check(L.tm_gnn_pool(coarse.n, ptr(coarse.child_ptr), ptr(coarse.child), ptr(lv.w), ptr(gout),
                                    ptr(ghc), _lib.F32, None, st))
                self._accum(grads, a, ghc)
                self._accum(grads, extra, gout)
        genc = grads.pop(self.enc_out)
        we, be = net.enc()
        check(L.tm_gnn_encode_backward(self.n0, ptr(r), ptr(lv0.d), ptr(lk), ptr(self.scale), ptr(genc),
                                       ptr(net._part), ptr(gview(0)), st))
        return sums

    @staticmethod
    def _accum(grads: Dict[int, torch.Tensor], k: int, g: torch.Tensor):
        if k in grads:
            grads[k].add_(g)
        else:
            grads[k] = g


class Adam:
    """Adam on the FP32 master vector (HIP kernel)."""

    def __init__(self, net: GNNPreconditioner, lr: float = 1e-3, betas=(0.9, 0.999), eps: float = 1e-8):
        self.net, self.lr, self.b1, self.b2, self.eps = net, lr, betas[0], betas[1], eps
        self.m = torch.zeros_like(net.params)
        self.v = torch.zeros_like(net.params)
        self.step_count = 0

    def step(self, grad: torch.Tensor, lr: Optional[float] = None):
        self.step_count += 1
        L = _lib.lib()
        self.net.fold(grad)
        check(L.tm_gnn_adam(self.net.n_params, ptr(self.net.params), ptr(grad), ptr(self.m), ptr(self.v),
                            ptr(self.net.mask), float(self.lr if lr is None else lr), self.b1, self.b2, self.eps,
                            self.step_count, _lib.stream_of(self.net.device)))
        self.net.refresh()
