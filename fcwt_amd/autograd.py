"""Differentiable CWT on device tensors: a ``torch.autograd.Function`` over
the plan's forward (fcwt_hip_cwt_device / fcwt_hip_ccwt_device) and its
hand-written adjoint (fcwt_hip_adjoint_device).

The reference has no derivative; the forward it differentiates is
FCWT::cwt (src/fcwt/fcwt.cpp:447-548).  For a real loss L of the output W,
PyTorch hands ``backward`` the gradient G = dL/dRe W + i dL/dIm W, and the
input gradient is

    xbar[t] = c Re sum_{q<N/2} A[q] e^{2 pi i q t / N},
    A[q]    = sum_i w_i(q) FFT_N(G_i)[q]

(the complex value itself for complex input), computed on the GPU; the signal
spectrum is not stored between forward and backward.

Usage::

    op = CWT(n, scales, fb=2.0)        # one plan, reused
    W = op(x)                          # x: (n,) or (B, n) float32/complex64 cuda tensor
    loss = (w * W.abs() ** 2).sum(); loss.backward()
"""
from __future__ import annotations

import numpy as np
import torch

from .api import Batch, Plan


def _stream(t: torch.Tensor) -> int:
    return torch.cuda.current_stream(t.device).cuda_stream


class _CWTFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x: torch.Tensor, op: "CWT"):
        if not x.is_cuda:
            raise ValueError("CWT: the input must be a CUDA (HIP) tensor")
        plan = op.plan
        cplx = x.is_complex()
        xin = x.detach().to(torch.complex64 if cplx else torch.float32).contiguous()
        batch = xin.shape[:-1]
        rows = xin.reshape(-1, plan.n)
        B = rows.shape[0]
        out = torch.empty((B, plan.fn, plan.n), dtype=torch.complex64, device=x.device)
        st = _stream(x)
        if B > 1 and not cplx:
            # one set of launches for the whole batch
            op.batch_for(B).execute_device([rows[b].data_ptr() for b in range(B)],
                                           [out[b].data_ptr() for b in range(B)], st)
        else:
            for b in range(B):
                plan.execute_device(rows[b].data_ptr(), out[b].data_ptr(), st, complex_input=cplx)
        if op.check_errors:
            op.check()
        ctx.plan = plan
        ctx.op = op
        ctx.cplx = cplx
        ctx.in_dtype = x.dtype
        return out.reshape(*batch, plan.fn, plan.n)

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        plan, cplx = ctx.plan, ctx.cplx
        g = grad_out.to(torch.complex64).contiguous().reshape(-1, plan.fn, plan.n)
        B = g.shape[0]
        xbar = torch.empty((B, plan.n), dtype=torch.complex64 if cplx else torch.float32, device=grad_out.device)
        st = _stream(grad_out)
        for b0 in range(0, B, 65535):
            cnt = min(65535, B - b0)
            plan.adjoint_batch_device(g[b0].data_ptr(), xbar[b0].data_ptr(), cnt, st, complex_input=cplx)
        if ctx.op.check_errors:
            ctx.op.check()
        shape = grad_out.shape[:-2] + (plan.n,)
        return xbar.reshape(shape).to(ctx.in_dtype), None


class CWT(torch.nn.Module):
    """The Morlet CWT of signals of length ``n`` at ``scales`` as a
    differentiable module; output (..., fn, n) complex64.

    The launches are asynchronous.  If an in-kernel synchronisation of the
    wide-band kernel fails, that launch's wide-band rows come out as NaN and
    ``check()`` raises ``FcwtError`` (once); later calls run normally.
    ``check_errors=True`` runs ``check()`` after every forward and backward
    (it waits for the device)."""

    def __init__(self, n: int, scales, fb: float = 2.0, normalize: bool = True, threads: int = 1,
                 check_errors: bool = False):
        super().__init__()
        self._scales = np.ascontiguousarray(scales, np.float32)
        self._args = (float(fb), int(threads), bool(normalize))
        self.plan = Plan(int(n), self._scales, float(fb), threads, normalize)
        self._batches = {}
        self.check_errors = bool(check_errors)

    def check(self) -> None:
        """Waits for the device; raises FcwtError if a launch of this
        module's plans failed since the last check."""
        self.plan.check()
        for b in self._batches.values():
            b.check()

    def batch_for(self, B: int) -> Batch:
        """The batch plan for B signals of this length (built once per B)."""
        if B not in self._batches:
            fb, threads, normalize = self._args
            self._batches[B] = Batch([self.plan.n] * B, self._scales, fb, threads, normalize)
        return self._batches[B]

    def forward(self, x: torch.Tensor) -> torch.Tensor:
        if x.shape[-1] != self.plan.n:
            raise ValueError(f"CWT: signal length {x.shape[-1]} != planned n = {self.plan.n}")
        return _CWTFunction.apply(x, self)


def cwt(x: torch.Tensor, op: CWT) -> torch.Tensor:
    """Functional form of ``op(x)``."""
    return _CWTFunction.apply(x, op)
