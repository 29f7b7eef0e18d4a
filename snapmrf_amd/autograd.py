"""torch.autograd binding of the EPG-SSFP differentiation entry points (a
thin test binding over the C ABI: the forward pass is mrf_epg_jacobian's
signal output, the backward pass mrf_epg_adjoint).

    s = epg_ssfp_signal(theta, fa, tr, phase, te, nstates)

theta: [n, 4] float32 on the GPU (T1 ms, T2 ms, B0 Hz, B1); fa, tr, phase,
te: [T] raw CSV columns (any device; FA and TR are differentiable).  s is the
[n, T] complex64 unnormalised readout of the epg_ssfp program
(src/functions.cu:936-1019) as a continuous function of its inputs.  PyTorch's
complex gradient for a real loss is dL/dRe + i dL/dIm, which is exactly the
cotangent mrf_epg_adjoint takes.
"""
from __future__ import annotations

import numpy as np
import torch

from . import epg_adjoint, epg_jacobian

__all__ = ["EpgSsfpSignal", "epg_ssfp_signal"]


def _table(fa, phase, tr, te) -> np.ndarray:
    cols = [c.detach().to("cpu", torch.float32).numpy() for c in (fa, phase, tr, te)]
    return np.ascontiguousarray(np.stack(cols))


class EpgSsfpSignal(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, fa, tr, phase, te, nstates: int):
        if not theta.is_cuda or theta.dtype != torch.float32:
            raise TypeError("theta must be a float32 GPU tensor [n, 4]")
        theta = theta.contiguous()
        h = _table(fa, phase, tr, te)
        n, T = theta.shape[0], h.shape[1]
        sig = torch.empty((n, T), dtype=torch.complex64, device=theta.device)
        epg_jacobian(h, theta, n, nstates, d_sig=sig,
                     stream=torch.cuda.current_stream(theta.device))
        ctx.save_for_backward(theta)
        ctx.h = h
        ctx.nstates = nstates
        ctx.seq_meta = (fa.device, fa.dtype, tr.device, tr.dtype)
        return sig

    @staticmethod
    def backward(ctx, grad_s):
        (theta,) = ctx.saved_tensors
        n, T = theta.shape[0], ctx.h.shape[1]
        cot = grad_s.to(torch.complex64).contiguous()
        g_theta = torch.empty((n, 4), dtype=torch.float32, device=theta.device)
        g_seq = torch.zeros((2, T), dtype=torch.float32, device=theta.device)
        epg_adjoint(ctx.h, theta, n, ctx.nstates, cot, d_grad_theta=g_theta, d_grad_seq=g_seq,
                    stream=torch.cuda.current_stream(theta.device))
        fdev, fdt, tdev, tdt = ctx.seq_meta
        return (g_theta, g_seq[0].to(fdev, fdt), g_seq[1].to(tdev, tdt), None, None, None)


def epg_ssfp_signal(theta, fa, tr, phase, te, nstates: int = 30):
    return EpgSsfpSignal.apply(theta, fa, tr, phase, te, nstates)
