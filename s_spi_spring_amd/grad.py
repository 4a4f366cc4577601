"""Differentiable masked Fourier-magnitude loss for PyTorch-ROCm.

FourierLoss is a torch.autograd.Function over the hand-written HIP kernels
behind spr_loss_grad_noalloc (include/spring_hip.h), queued on PyTorch's current
stream without allocation or synchronisation: the per-field losses (FP64 sums)
stay on the device. When rho needs a gradient, forward also computes it through
the adjoint FFT and backward scales the saved gradient by the incoming one;
when it does not (torch.no_grad, or a tensor without requires_grad), forward
runs the loss passes only. The entry point's fixed scratch is shared by all its
calls on one GPU, so losses evaluated concurrently on several streams of one
GPU must be ordered by the caller. Losses (kind):

  "poisson"    L = sum_{A>=0} |F|^2 - I log(|F|^2 + eps),  I = A^2 (counts)
  "amplitude"  L = sum_{A>=0} (|F| - A)^2

with F = FT(rho) (unnormalised, exp(+2 pi i k.x/N)) and A the corner-origin
amplitude of the backend. The gradient follows PyTorch's convention for
complex tensors (dL/dRe + i dL/dIm), so torch.optim works on rho directly.
"""
from __future__ import annotations

import torch

from . import hipcore

KINDS = {"poisson": 0, "amplitude": 1}


class FourierLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rho, backend, kind="poisson", eps=1e-3):
        if not rho.is_cuda or rho.dtype != torch.complex64:
            raise TypeError("rho must be a complex64 tensor on the GPU")
        if rho.device.index != backend.device:
            raise ValueError(f"rho is on cuda:{rho.device.index}, the backend's amplitude on cuda:{backend.device}")
        rho = rho.contiguous()
        if rho.dim() == 2:
            rho = rho.unsqueeze(0)
        P = rho.shape[0]
        if tuple(rho.shape[1:]) != (backend.n, backend.n):
            raise ValueError("field size differs from the backend's pattern")
        want_grad = bool(ctx.needs_input_grad[0])
        grad = torch.empty_like(rho)          # the gradient, or the passes' workspace
        loss = torch.empty(P, dtype=torch.float64, device=rho.device)
        amp = backend.device_buffers()[2]
        with torch.cuda.device(rho.device):
            stream = torch.cuda.current_stream(rho.device).cuda_stream
            hipcore.loss_grad_noalloc(backend.n, P, amp, rho.data_ptr(), grad.data_ptr(), loss.data_ptr(),
                                      KINDS[kind], eps, want_grad, stream)
        if want_grad:
            ctx.save_for_backward(grad)
        return loss

    @staticmethod
    def backward(ctx, grad_out):
        (grad,) = ctx.saved_tensors
        return grad * grad_out.to(torch.float32)[:, None, None], None, None, None


def fourier_loss(rho, backend, kind="poisson", eps=1e-3):
    """Per-field loss (float64 tensor of length P) of rho [P, N, N] (or [N, N])."""
    return FourierLoss.apply(rho, backend, kind, eps)


def refine(rho, backend, steps=20, lr=1e-2, kind="poisson", eps=1e-3):
    """Adam refinement of rho (complex64 [P, N, N] on the GPU) on the loss;
    returns the refined tensor and the loss history (list of float64 arrays)."""
    x = rho.detach().clone().requires_grad_(True)
    opt = torch.optim.Adam([x], lr=lr)
    hist = []
    for _ in range(steps):
        opt.zero_grad()
        loss = fourier_loss(x, backend, kind, eps)
        loss.sum().backward()
        opt.step()
        hist.append(loss.detach().cpu().numpy())
    return x.detach(), hist


def refine_support(rho, support, backend, steps=20, lr_rel=1e-2, kind="poisson", eps=1e-3):
    """Support-projected Adam refinement (the GRAD operation of the memetic loop)
    written with torch.optim.Adam: the densities set to 0 outside their
    supports, then one parameter group per field with step
    lr_rel * RMS(rho on its support), and the gradient masked by the support.
    rho: complex64 [P, N, N] on the GPU; support: 0/1 [P, N, N]."""
    S = torch.as_tensor(support, device=rho.device).to(torch.float32)
    x = rho.detach().clone() * S
    params, groups = [], []
    for p in range(x.shape[0]):
        sp = S[p] > 0
        cnt = int(sp.sum().item())
        e = float((x[p][sp].abs().double() ** 2).sum().item()) if cnt else 0.0
        lr = lr_rel * (e / cnt) ** 0.5 if cnt else 0.0
        t = x[p].clone().requires_grad_(True)
        t.register_hook(lambda g, m=S[p]: g * m)
        params.append(t)
        groups.append({"params": [t], "lr": lr})
    opt = torch.optim.Adam(groups)
    for _ in range(steps):
        opt.zero_grad()
        loss = fourier_loss(torch.stack(params), backend, kind, eps)
        loss.sum().backward()
        opt.step()
    return torch.stack([t.detach() for t in params])
