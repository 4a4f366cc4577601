"""Gradient-refinement operator: the HIP loss/gradient kernels against an FP64
reference gradient (numpy, closed form 2 FT^H[f'(|F|^2) F]) and against central
finite differences of the FP64 loss (step 1e-3, 64 seeded random pixels, real
and imaginary parts); the torch.autograd.Function against the direct call; the
on-device Adam against torch.optim.Adam.

Tolerances: gradient relative L2 <= 1e-4 against FP64 (the weight
1 - I/(|F|^2+eps) amplifies FP32 rounding of |F|^2); finite differences within
1e-3 relative of the gradient scale; loss relative error <= 1e-6.
"""
import numpy as np
import pytest

from conftest import rel_l2
from helpers import case

pytestmark = pytest.mark.gpu


def np_loss_grad(rho, amp, kind, eps):
    n = rho.shape[-1]
    F = np.fft.ifft2(rho.astype(np.complex128)) * n * n
    A = amp.astype(np.float64)
    k = A >= 0
    u = np.abs(F) ** 2
    if kind == 0:
        I = A ** 2
        L = np.sum((u - I * np.log(u + eps))[k])
        w = np.where(k, 1 - I / (u + eps), 0)
    else:
        m = np.sqrt(u)
        L = np.sum(((m - A) ** 2)[k])
        w = np.where(k & (m > 0), 1 - A / np.where(m > 0, m, 1), 0)
    g = 2 * np.fft.fft2(w * F)
    return L, g


@pytest.mark.parametrize("kind", [0, 1])
@pytest.mark.parametrize("n", [128, 384, 512])
def test_gradient_vs_fp64_and_fd(hip, n, kind):
    import torch
    amp, rho, supp = case(n, 2, seed=21)
    be = hip.Backend(amp, capacity=2)
    x = torch.from_numpy(rho).cuda()
    g = torch.empty_like(x)
    loss = be.loss_grad(2, x.data_ptr(), g.data_ptr(), kind, 1e-3, torch.cuda.current_stream().cuda_stream)
    g = g.cpu().numpy()
    rng = np.random.default_rng(5)
    for i in range(2):
        L, gref = np_loss_grad(rho[i], amp, kind, 1e-3)
        assert abs(loss[i] - L) <= 1e-6 * abs(L)
        e = rel_l2(g[i], gref)
        print(f"n={n} kind={kind}: gradient rel-L2 vs FP64 {e:.2e}")
        assert e < 1e-4
        # central finite differences of the FP64 loss at 64 random pixels
        h = 1e-3
        scale = np.abs(gref).max()
        for p in rng.choice(n * n, 64, replace=False):
            y, xx = divmod(int(p), n)
            for d in (1.0, 1j):
                r1 = rho[i].astype(np.complex128).copy()
                r2 = r1.copy()
                r1[y, xx] += h * d
                r2[y, xx] -= h * d
                fd = (np_loss_grad(r1, amp, kind, 1e-3)[0] - np_loss_grad(r2, amp, kind, 1e-3)[0]) / (2 * h)
                got = g[i][y, xx].real if d == 1.0 else g[i][y, xx].imag
                assert abs(fd - got) <= 1e-3 * scale, (p, d, fd, got)


def test_autograd_function(hip):
    import torch
    from s_spi_spring_amd.grad import fourier_loss
    amp, rho, supp = case(128, 3, seed=22)
    be = hip.Backend(amp, capacity=3)
    x = torch.from_numpy(rho).cuda().requires_grad_(True)
    loss = fourier_loss(x, be, "poisson")
    (loss * torch.tensor([1.0, 2.0, 0.5], dtype=torch.float64, device="cuda")).sum().backward()
    for i, w in enumerate([1.0, 2.0, 0.5]):
        L, gref = np_loss_grad(rho[i], amp, 0, 1e-3)
        assert rel_l2(x.grad[i].cpu().numpy(), w * gref) < 1e-4


def test_device_adam_matches_torch_adam(hip):
    import torch
    from s_spi_spring_amd.grad import refine
    amp, rho, supp = case(128, 2, seed=23)
    be = hip.Backend(amp, capacity=2)
    ref, hist = refine(torch.from_numpy(rho).cuda(), be, steps=20, lr=1e-2)
    be.set_population(rho, supp)
    loss = be.refine(20, lr=1e-2)
    got = be.get_population(support=False)
    assert rel_l2(got, ref.cpu().numpy()) < 1e-4
    assert np.all(loss < hist[0])   # the refinement lowers the loss


def test_support_refinement_matches_torch_adam(hip):
    """spr_refine_support (the GRAD operation) = torch.optim.Adam with one group
    per field, lr = lr_rel x RMS density on the support, gradient masked by it."""
    import torch
    from s_spi_spring_amd.grad import refine_support
    amp, rho, supp = case(128, 3, seed=24)
    supp[2] = 0                                     # an empty support: nothing moves
    be = hip.Backend(amp, capacity=3)
    ref = refine_support(torch.from_numpy(rho).cuda(), supp, be, steps=10, lr_rel=2e-2).cpu().numpy()
    be.set_population(rho, supp)
    be.refine_support(10, lr_rel=2e-2)
    got = be.get_population(support=False)
    assert rel_l2(got[:2], ref[:2]) < 1e-4
    assert np.array_equal(got[supp == 0], rho[supp == 0])      # outside the supports: unchanged
    assert not np.array_equal(got[0][supp[0] == 1], rho[0][supp[0] == 1])
