"""The NCE objective of the NCVis epoch loop as a differentiable torch op.

The reference has no loss function: its update (src/ncvis.cpp:304-338) applies, per sample, the
gradient of

    s(i, j) = log Ph(d2) - Q - log n,   Ph = 1 / (1 + a d2^b),   d2 = |y_i - y_j|^2
    L = sum_edges [ softplus(-s(i, j)) + sum_{m=1..n} softplus(s(i, neg_m)) ]

with Y += -step * dL/dY (clipped to +-4) and Q -= alpha_Q * dL/dQ. ``NCELoss`` evaluates L with
the HIP forward kernel (FP64 reduction) and returns dL/dY, dL/dQ from the hand-written adjoint
kernel -- no autograd tape through the loss.
"""
from __future__ import annotations

import torch

from . import abi


class NCELoss(torch.autograd.Function):
    """loss = NCELoss.apply(Y, Q, row_ptr, col, neg, a, b)

    Y: [N, d] float32 (cuda), Q: 0-dim float32 tensor, row_ptr: [N+1] int64 and col: [E] int32
    (the symmetrized graph in CSR / build_edges order), neg: [E, n] int32 negatives (n >= 1).
    Returns a float64 0-dim tensor.
    """

    @staticmethod
    def forward(ctx, Y, Q, row_ptr, col, neg, a, b):
        if not (Y.is_cuda and Y.dtype == torch.float32 and Y.is_contiguous()):
            raise ValueError("Y must be a contiguous float32 CUDA tensor")
        if row_ptr.dtype != torch.int64 or col.dtype != torch.int32 or neg.dtype != torch.int32:
            raise ValueError("row_ptr must be int64, col and neg int32")
        N, d = Y.shape
        E, n = neg.shape
        if col.numel() != E or row_ptr.numel() != N + 1:
            raise ValueError("graph shapes do not match Y / neg")
        loss = torch.zeros((), dtype=torch.float64, device=Y.device)
        st = torch.cuda.current_stream(Y.device).cuda_stream
        abi.nce_loss(N, d, row_ptr.data_ptr(), col.data_ptr(), E, neg.data_ptr(), n, Y.data_ptr(),
                     float(Q), float(a), float(b), loss.data_ptr(), None, None, st)
        ctx.save_for_backward(Y, row_ptr, col, neg)
        ctx.q, ctx.a, ctx.b = float(Q), float(a), float(b)
        return loss

    @staticmethod
    def backward(ctx, grad_loss):
        Y, row_ptr, col, neg = ctx.saved_tensors
        N, d = Y.shape
        E, n = neg.shape
        gY = torch.zeros_like(Y)
        gQ = torch.zeros((), dtype=torch.float64, device=Y.device)
        scratch = torch.zeros((), dtype=torch.float64, device=Y.device)
        st = torch.cuda.current_stream(Y.device).cuda_stream
        abi.nce_loss(N, d, row_ptr.data_ptr(), col.data_ptr(), E, neg.data_ptr(), n, Y.data_ptr(),
                     ctx.q, ctx.a, ctx.b, scratch.data_ptr(), gY.data_ptr(), gQ.data_ptr(), st)
        g = grad_loss.to(torch.float32)
        return gY * g, (gQ * grad_loss).to(torch.float32), None, None, None, None, None


def nce_loss(Y, Q, row_ptr, col, neg, a, b):
    """Functional form of :class:`NCELoss`."""
    if not torch.is_tensor(Q):
        Q = torch.tensor(float(Q), dtype=torch.float32, device=Y.device)
    return NCELoss.apply(Y, Q, row_ptr, col, neg, a, b)
