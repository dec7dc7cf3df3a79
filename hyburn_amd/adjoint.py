"""Gradients through the batched integration path (hb_adjoint_* in
include/hyburn_hip.h).

`AdjointPlan` uploads a ragged set of instances with their losses, runs the
forward (checkpoints + coupling log + loss) and the hand-written discrete
adjoint on the GPU. `NetworkLoss` wraps it as a torch.autograd.Function: the
differentiable inputs are the global coupling G (coupling_params[0]) of each
instance, one model parameter per node, and the initial state.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

from . import _lib as L
from .instances import Instance, _Marshal


@dataclass
class Loss:
    """kind: L.LOSS_FC (target [N, N]), L.LOSS_BOLD_MSE (target [n_bold, N]
    native BOLD rows) or L.LOSS_ORDER (target [n_steps], Kuramoto)."""
    kind: int
    target: np.ndarray
    grad_param: int = -1     # model parameter index differentiated per node


@dataclass
class Gradients:
    loss: float
    grad_coupling: float
    grad_param: Optional[np.ndarray]
    grad_x0: np.ndarray
    final_state: np.ndarray


class AdjointPlan:
    def __init__(self, insts: Sequence[Instance], losses: Sequence[Loss], device: int = 0,
                 checkpoint_every: int = 0, blocking: bool = True, bold_helpers: bool = True):
        if len(insts) != len(losses):
            raise ValueError("one loss per instance")
        self.insts = list(insts)
        # the BOLD monitor is part of the differentiated graph whenever it is configured
        self._m = _Marshal([dataclasses.replace(I, record_bold=I.bold_period > 0) for I in self.insts])
        n = len(self.insts)
        self._spec = (L.HbAdjointSpec * n)()
        self._keep = []
        self.loss = np.zeros(n, np.float32)
        self.grad_coupling = np.zeros(n, np.float32)
        self.grad_param = [None] * n
        self.grad_x0 = []
        self.final_state = []
        for i, (I, Ls) in enumerate(zip(self.insts, losses)):
            s = self._spec[i]
            tgt = np.ascontiguousarray(Ls.target, np.float32).ravel()
            self._keep.append(tgt)
            s.loss_kind = Ls.kind
            s.target = tgt.ctypes.data_as(L.fp)
            s.grad_param = Ls.grad_param
            s.loss = C.cast(self.loss.ctypes.data + 4 * i, L.fp)
            s.grad_coupling = C.cast(self.grad_coupling.ctypes.data + 4 * i, L.fp)
            if Ls.grad_param >= 0:
                g = np.zeros(I.n_nodes, np.float32)
                self.grad_param[i] = g
                s.grad_param_out = g.ctypes.data_as(L.fp)
            gx = np.zeros((I.n_nodes, I.nvar), np.float32)
            fs = np.zeros((I.n_nodes, I.nvar), np.float32)
            self.grad_x0.append(gx)
            self.final_state.append(fs)
            s.grad_x0 = gx.ctypes.data_as(L.fp)
            s.final_state = fs.ctypes.data_as(L.fp)
        o = L.HbAdjointOpts()
        o.device, o.checkpoint_every = device, int(checkpoint_every)
        # bold_helpers=False: the BOLD monitor's forward as a kernel of its own, not helper waves
        o.flags = (0 if blocking else L.FLAG_NO_BLOCKING) | (0 if bold_helpers else L.FLAG_SEPARATE_BOLD)
        self._h = C.c_void_p()
        L.check(L.lib().hb_adjoint_plan_create(self._m.arr, self._spec, n, C.byref(o), C.byref(self._h)))
        self.stats = L.HbAdjointStats()

    def forward(self) -> np.ndarray:
        L.check(L.lib().hb_adjoint_forward(self._h, C.byref(self.stats)))
        L.check(L.lib().hb_adjoint_download(self._h))
        return self.loss.copy()

    def backward(self) -> list:
        L.check(L.lib().hb_adjoint_backward(self._h, C.byref(self.stats)))
        L.check(L.lib().hb_adjoint_download(self._h))
        return [Gradients(float(self.loss[i]), float(self.grad_coupling[i]),
                          None if self.grad_param[i] is None else self.grad_param[i].copy(),
                          self.grad_x0[i].copy(), self.final_state[i].copy()) for i in range(len(self.insts))]

    def close(self):
        if self._h:
            L.lib().hb_adjoint_plan_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def _network_loss_fn():
    import torch

    class NetworkLoss(torch.autograd.Function):
        """losses[b] = loss_b(simulate(instance_b with G = G[b], params[grad_param]
        per node = node_param[b], x0 = x0[b])); backward runs the HIP adjoint.
        Tensors are CPU float32 (host-side plumbing; the work runs on the GPU)."""

        @staticmethod
        def forward(ctx, G, node_param, x0, insts, losses, checkpoint_every=0, device=0):
            built = []
            for b, I in enumerate(insts):
                cp = list(I.cparams) if len(I.cparams) else [1.0]
                cp[0] = float(G[b])
                kw = dict(cparams=tuple(cp), x0=x0[b].detach().numpy().astype(np.float32))
                gp = losses[b].grad_param
                if node_param is not None and gp >= 0:
                    kw["node_param_idx"] = gp
                    kw["node_param"] = node_param[b].detach().numpy().astype(np.float32)
                built.append(dataclasses.replace(I, **kw))
            plan = AdjointPlan(built, losses, device=device, checkpoint_every=checkpoint_every)
            loss = plan.forward()
            ctx.plan = plan
            ctx.has_np = node_param is not None
            return torch.from_numpy(loss.copy())

        @staticmethod
        def backward(ctx, gl):
            grads = ctx.plan.backward()
            ctx.plan.close()
            gG = torch.tensor([g.grad_coupling for g in grads], dtype=torch.float32) * gl
            gnp = None
            if ctx.has_np:
                gnp = torch.stack([torch.from_numpy(g.grad_param) for g in grads]) * gl[:, None]
            gx0 = torch.stack([torch.from_numpy(g.grad_x0) for g in grads]) * gl[:, None, None]
            return gG, gnp, gx0, None, None, None, None

    return NetworkLoss


def network_loss(G, node_param, x0, insts: Sequence[Instance], losses: Sequence[Loss],
                 checkpoint_every: int = 0, device: int = 0):
    """Differentiable per-instance losses (torch tensors in, [B] tensor out)."""
    return _network_loss_fn().apply(G, node_param, x0, list(insts), list(losses), checkpoint_every, device)
