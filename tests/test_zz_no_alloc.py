"""spr_loss_grad_noalloc, the entry point that allocates nothing, against
spr_loss_grad (the handle's entry point with the same kernels).

GPU: two sizes and populations in a row (256^2 with 5 fields, then the
smaller 128^2 with 2, then 256^2 with 3 again), both loss kinds, gradients and
FP64 losses bit-identical to spr_loss_grad's; the losses alone (want_grad = 0)
identical too. Outputs start poisoned with NaN so that a pass that skipped a
field shows. CPU: argument errors are reported before any device call.
"""
import ctypes

import numpy as np
import pytest

from helpers import case


def _run_pair(hip, torch, n, P, seed):
    amp, rho, _ = case(n, P, seed=seed)
    be = hip.Backend(amp, capacity=P)
    _, _, amp_ptr = be.device_buffers()
    stream = torch.cuda.current_stream().cuda_stream
    x = torch.from_numpy(rho).cuda()
    for kind in (0, 1):
        g_ref = torch.empty_like(x)
        loss_ref = be.loss_grad(P, x.data_ptr(), g_ref.data_ptr(), kind, 1e-3, stream)
        g = torch.full_like(x, complex(float("nan"), float("nan")))
        loss = torch.full((P,), float("nan"), dtype=torch.float64, device=x.device)
        print(f"noalloc n={n} P={P} kind={kind}", flush=True)
        hip.loss_grad_noalloc(n, P, amp_ptr, x.data_ptr(), g.data_ptr(), loss.data_ptr(), kind, 1e-3, True, stream)
        torch.cuda.synchronize()
        assert torch.equal(g, g_ref)
        assert np.array_equal(loss.cpu().numpy(), loss_ref)
        loss2 = torch.full((P,), float("nan"), dtype=torch.float64, device=x.device)
        hip.loss_grad_noalloc(n, P, amp_ptr, x.data_ptr(), g.data_ptr(), loss2.data_ptr(), kind, 1e-3, False, stream)
        torch.cuda.synchronize()
        assert np.array_equal(loss2.cpu().numpy(), loss_ref)
    be.close()


@pytest.mark.gpu
def test_noalloc_matches_loss_grad_on_calls_in_a_row(hip):
    import torch
    _run_pair(hip, torch, 256, 5, seed=31)
    _run_pair(hip, torch, 128, 2, seed=32)     # smaller input on the second call
    _run_pair(hip, torch, 256, 3, seed=33)


def test_noalloc_rejects_bad_arguments_before_any_device_call(hip):
    lib = hip.load()
    buf = np.zeros(16, np.float64)
    p = buf.ctypes.data_as(ctypes.c_void_p)
    q = np.zeros(16, np.float64).ctypes.data_as(ctypes.c_void_p)

    def call(n, P, amp, rho, grad, kind=0):
        return lib.spr_loss_grad_noalloc(n, P, amp, rho, grad, None, kind, 1e-3, 1, None)

    assert call(500, 1, p, p, q) == -1                      # unsupported size
    assert b"size" in lib.spr_last_error(None)
    assert call(256, 0, p, p, q) == -1                      # no fields
    assert call(256, 1, None, p, q) == -1                   # NULL amplitude
    assert call(256, 1, p, p, q, kind=7) == -1              # unknown loss kind
    assert call(2048, 257, p, p, q) == -1                   # beyond the fixed scratch (256 fields of 2048^2)
    assert b"fixed scratch" in lib.spr_last_error(None)
    assert call(1024, 1025, p, p, q) == -1                  # 1024 fields of 1024^2 at most
    assert call(256, 1, p, p, p) == -1                      # rho and grad overlap
    assert b"overlap" in lib.spr_last_error(None)
