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
    # without a gradient to compute, forward runs the loss passes only: same losses
    with torch.no_grad():
        loss2 = fourier_loss(x, be, "poisson")
    assert torch.equal(loss2, loss.detach())
    assert fourier_loss(x.detach(), be, "poisson").grad_fn is None


@pytest.mark.parametrize("n", [128, 512])
def test_device_adam_matches_torch_adam(hip, n):
    """spr_refine = torch.optim.Adam on the gradient of the same loss pass
    (FourierLoss). The two Adam updates order their operations differently, and
    at 1024^2 twenty steps at this step size amplify that rounding beyond the
    bar (Adam's first steps move every component by about lr whatever the
    gradient's size, so a component whose gradient is near zero can take either
    sign); there the step is checked component by component instead
    (test_refinement_step_follows_the_gradient)."""
    import torch
    from s_spi_spring_amd.grad import refine
    print(f"refinement vs torch Adam at {n}^2", flush=True)
    amp, rho, supp = case(n, 2, seed=23)
    be = hip.Backend(amp, capacity=2)
    ref, hist = refine(torch.from_numpy(rho).cuda(), be, steps=20, lr=1e-2)
    be.set_population(rho, supp)
    loss = be.refine(20, lr=1e-2)
    got = be.get_population(support=False)
    assert rel_l2(got, ref.cpu().numpy()) < 1e-4
    assert np.all(loss < hist[0])   # the refinement lowers the loss


@pytest.mark.parametrize("kind", [0, 1])
@pytest.mark.parametrize("n", [512, 1024])
def test_refinement_step_follows_the_gradient(hip, n, kind):
    """One spr_refine step. Adam's first step is x -= lr * g / (|g| + 1e-8) per
    real component (m/(1-b1) = g, v/(1-b2) = g^2), so every component whose
    FP64 gradient (np_loss_grad) is not negligible moves by lr against its
    sign: this checks the refinement's gradient pass (without loss partials)
    and its Adam row pass component by component."""
    amp, rho, supp = case(n, 2, seed=25)
    lr = 1e-3
    be = hip.Backend(amp, capacity=2)
    be.set_population(rho, supp)
    print(f"one refinement step at {n}^2, loss kind {kind}", flush=True)
    be.refine(1, lr=lr, kind=kind)
    got = be.get_population(support=False)
    be.close()
    for i in range(2):
        _, g = np_loss_grad(rho[i], amp, kind, 1e-3)
        d = (got[i] - rho[i]).astype(np.complex128)
        for comp, gc, dc in (("re", g.real, d.real), ("im", g.imag, d.imag)):
            rms = np.sqrt(np.mean(gc ** 2))
            sel = np.abs(gc) > max(1e-3 * rms, 1e-4)   # and |g| >> Adam's 1e-8
            wrong = np.count_nonzero(np.sign(dc[sel]) != -np.sign(gc[sel]))
            size = np.abs(np.abs(dc[sel]) - lr).max() / lr
            print(f"  individual {i} {comp}: {sel.sum()} of {gc.size} components selected (rms {rms:.2e}), {wrong} moved the wrong way, "
                  f"largest step-size error {size:.1e}")
            assert wrong == 0
            assert size < 1e-3


def test_refinement_gradient_pass_in_a_same_plan_variant_is_bit_identical(hip):
    """The refinement's gradient pass runs in the handle's column variant when
    that variant holds the default radix plan (at 512^2 the tuned variant 6:
    8 columns in a skewed tile); its results equal the default geometry's bit
    for bit."""
    amp, rho, supp = case(512, 3, seed=27)
    out = []
    for cv in (0, 6):
        be = hip.Backend(amp, capacity=3)
        be.set_variants(cv, 0)
        be.set_population(rho, supp)
        be.refine_support(5, lr_rel=2e-2)
        out.append(be.get_population(support=False))
        be.close()
    assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("n", [128, 512])
def test_support_refinement_matches_torch_adam(hip, n):
    """spr_refine_support (the GRAD operation) = torch.optim.Adam with one group
    per field, lr = lr_rel x RMS density on the support, gradient masked by it,
    from the densities set to 0 outside their supports."""
    import torch
    from s_spi_spring_amd.grad import refine_support
    print(f"support refinement vs torch Adam at {n}^2", flush=True)
    amp, rho, supp = case(n, 3, seed=24)
    supp[2] = 0                                     # an empty support: nothing moves
    be = hip.Backend(amp, capacity=3)
    ref = refine_support(torch.from_numpy(rho).cuda(), supp, be, steps=10, lr_rel=2e-2).cpu().numpy()
    be.set_population(rho, supp)
    be.refine_support(10, lr_rel=2e-2)
    got = be.get_population(support=False)
    assert rel_l2(got[:2], ref[:2]) < 1e-4
    assert not np.any(got[supp == 0])                          # outside the supports: 0
    assert not np.array_equal(got[0][supp[0] == 1], rho[0][supp[0] == 1])


def test_refinement_clears_its_moments_beyond_2g_elements(hip):
    """The Adam moments of a refinement are cleared at every call, also when the
    moment arrays hold more than 2^31 32-bit words (1024^2 with 1100 fields:
    2 N^2 P = 2.3e9). Three individuals replicated on the device over 1100
    slots are refined twice from the same start (the first call leaves
    non-zero moments behind); the second call's results equal those of a
    fresh three-field handle bit for bit (the per-field arithmetic does not
    depend on the population size)."""
    n, P = 1024, 1100
    amp, rho, supp = case(n, 3, seed=41)
    print(f"refinement of {P} fields of {n}^2 (moments {2 * n * n * P} words)", flush=True)
    big = hip.Backend(amp, capacity=P + 3)
    big.put_individuals([P, P + 1, P + 2], list(rho), list(supp))
    src = [P + (i % 3) for i in range(P)] + [-1, -1, -1]
    big.gather(src)
    big.set_active(P)
    big.refine_support(3, lr_rel=2e-2)
    big.gather(src)
    big.set_active(P)
    big.refine_support(3, lr_rel=2e-2)
    got = big.fetch_individuals([0, 1, 2, P - 1], support=False)
    big.close()
    small = hip.Backend(amp, capacity=3)
    small.set_population(rho, supp)
    small.refine_support(3, lr_rel=2e-2)
    want = small.get_population(support=False)
    small.close()
    for i in range(3):
        assert np.array_equal(got[i], want[i]), i
    assert np.array_equal(got[3], want[(P - 1) % 3])


def test_low_count_refinement_lowers_the_poisson_loss(hip):
    """The differentiation workload's step at its size: 1024^2 truncated
    octahedron with 1e8 photons, after a RAAR block, 20 support-projected Adam
    steps (step 1e-2 of the RMS density) lower every individual's Poisson
    likelihood (FP64, numpy) below that of its projection onto the support,
    where they start, and leave 0 outside the supports."""
    from s_spi_spring_amd import synth
    n, P = 1024, 2
    I, _ = synth.octahedron_pattern(seed=3, n=n, photons=1e8)
    amp = synth.corner_amplitude(I)
    supp = synth.disk_support(n, 2 * (150 * n // 1024 + 8))
    rho = synth.random_start(n, P, supp, seed=4, scale=float(np.sqrt(np.maximum(I, 0).sum()) / supp.sum()))
    supps = np.broadcast_to(supp, rho.shape).astype(np.int32).copy()
    be = hip.Backend(amp, capacity=P)
    be.set_population(rho, supps)
    be.iterate("RAAR", 25, 0.87, 0.0)
    before = be.get_population(support=False)
    be.refine_support(20, lr_rel=1e-2, kind=0)
    after = be.get_population(support=False)
    be.close()
    for i in range(P):
        l0 = np_loss_grad(before[i] * supps[i], amp, 0, 1e-3)[0]
        l1 = np_loss_grad(after[i], amp, 0, 1e-3)[0]
        print(f"individual {i}: Poisson loss {l0:.6e} -> {l1:.6e}")
        assert l1 < l0
        assert not np.any(after[i][supps[i] == 0])


def test_differentiation_workload_runs(hip):
    """Two generations of the differentiation workload's memetic loop (RAAR
    blocks alternating with 20 Adam steps) at 1024^2 with 1e8 photons, with a
    small population: every generation's best Fourier error is finite and well
    below 1 (0.20 measured; the noise floor of this pattern is 0.23), and the
    second generation's best is below the first's."""
    from s_spi_spring_amd import workloads
    r = workloads.differentiation_run(0, generations=1, popsize=4)
    print(r["history"], r["times"], r["real_space_error"], r["noise_floor"])
    assert len(r["history"]) == 2 and all(np.isfinite(r["history"]))
    assert max(r["history"]) < 0.5
    assert r["history"][-1] < r["history"][0]
    assert r["adam_steps_per_generation"] == 160 and r["iterations_per_generation"] == 250
