"""Adjoint (gradient) path: hand-written HIP adjoints for image and kernel,
through the torch.autograd wrapper, against FP64 central finite differences
of the loss and FP64 reference gradients.

Workload (differentiation config): 256x256x128 (x,y,z) blob image, 21^3
Gaussian PSF, loss 0.5*||conv(img, psf) - target||^2 with target a seeded
convolution of a perturbed image.  FD tolerance: 1e-3 relative (per entry,
against max(|fd|, rms of the gradient)); full gradients vs the FP64
reference: rel-L2 <= 1e-4 (the FP32 residual conv - target loses a few bits
to cancellation).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_grad_kernel_small_vs_direct(fc, oracle):
    import torch
    from fourierconvolutioncudalib_amd.autograd import grad_kernel
    rng = np.random.default_rng(5)
    img = rng.random((14, 18, 22), dtype=np.float32)
    g = rng.standard_normal((14, 18, 22)).astype(np.float32)
    for kshape in [(3, 5, 7), (1, 1, 1), (14, 3, 22), (5, 18, 1)]:
        dk = grad_kernel(torch.from_numpy(g).cuda(), torch.from_numpy(img).cuda(), kshape)
        ref = oracle.conv_grad_kernel(g.astype(np.float64), img, kshape)
        assert oracle.rel_l2(dk.cpu().numpy(), ref) < 1e-5, kshape


@pytest.mark.parametrize("shape,kshapes", [
    ((9, 12, 35), [(3, 5, 7), (9, 12, 35), (4, 6, 2)]),    # odd nx: no Nyquist bin
    ((6, 10, 74), [(3, 3, 9), (5, 4, 74)]),                # Bluestein nx (37 > 31)
    ((8, 16, 256), [(5, 7, 21), (3, 16, 256), (1, 1, 1)]),
])
def test_grad_kernel_row_c2r_shapes(fc, oracle, shape, kshapes):
    """PSF gradient against FP64 where the support rows' x-C2R is evaluated
    directly at the kernel's x positions (odd and even nx, full and
    one-tap x extents) and where a Bluestein x keeps the x inverse pass."""
    import torch
    from fourierconvolutioncudalib_amd.autograd import grad_kernel
    rng = np.random.default_rng(sum(shape))
    img = rng.random(shape, dtype=np.float32)
    g = rng.standard_normal(shape).astype(np.float32)
    for kshape in kshapes:
        print("grad_kernel", shape, kshape, flush=True)
        dk = grad_kernel(torch.from_numpy(g).cuda(), torch.from_numpy(img).cuda(), kshape)
        ref = oracle.conv_grad_kernel(g.astype(np.float64), img, kshape)
        assert oracle.rel_l2(dk.cpu().numpy(), ref) < 1e-5, (shape, kshape)


def _fft64(a):
    return np.fft.rfftn(np.asarray(a, np.float64))


def _corr_grad_kernel64(r, img, kshape):
    """FP64 dk[j] = sum_i r[i] img[i - (j - c)] via FFT cross-correlation."""
    u = np.fft.irfftn(_fft64(r) * np.conj(_fft64(img)), s=r.shape, axes=(0, 1, 2))
    idx = [(np.arange(k) - k // 2) % n for k, n in zip(kshape, r.shape)]
    return u[np.ix_(*idx)]


def test_differentiation_256x256x128_k21(fc, oracle):
    import torch
    from fourierconvolutioncudalib_amd import synth
    from fourierconvolutioncudalib_amd.autograd import fft_conv3d

    shape = (128, 256, 256)
    img = synth.blobs(shape, 2004, count=400, poisson=False)
    pert = img + 0.1 * synth.blobs(shape, 2005, count=400, background=0.0, poisson=False)
    psf = synth.gaussian_psf((21, 21, 21), (2.0, 2.5, 3.0))
    target = oracle.numpy_fft_conv(pert, psf).astype(np.float32)

    ti = torch.from_numpy(img).cuda().requires_grad_(True)
    tk = torch.from_numpy(psf).cuda().requires_grad_(True)
    tt = torch.from_numpy(target).cuda()
    out = fft_conv3d(ti, tk)
    loss = 0.5 * ((out - tt) ** 2).sum()
    loss.backward()
    gi = ti.grad.cpu().numpy().astype(np.float64)
    gk = tk.grad.cpu().numpy().astype(np.float64)

    # FP64 reference gradients
    r0 = oracle.numpy_fft_conv(img, psf) - target.astype(np.float64)
    gi_ref = oracle.numpy_fft_conv(r0, psf, conj_kernel=True)
    gk_ref = _corr_grad_kernel64(r0, img, psf.shape)
    assert oracle.rel_l2(gi, gi_ref) < 1e-4
    assert oracle.rel_l2(gk, gk_ref) < 1e-4

    # FP64 central differences of the loss on 200 random entries
    rng = np.random.default_rng(2006)
    L0 = 0.5 * np.sum(r0 * r0)
    kz, ky, kx = psf.shape
    c = (kz // 2, ky // 2, kx // 2)
    psf64 = psf.astype(np.float64)
    img64 = img.astype(np.float64)
    scale_i = np.sqrt(np.mean(gi_ref ** 2))
    scale_k = np.sqrt(np.mean(gk_ref ** 2))
    h = 1e-2
    # image entries: the perturbation h*e_i changes the residual by h * psf
    # placed at i (tap j at i + j - c), so L(+-h) is re-evaluated exactly on
    # that support
    for _ in range(100):
        i = tuple(int(rng.integers(0, n)) for n in shape)
        pos = np.ix_(*[(i[a] + np.arange(psf.shape[a]) - c[a]) % shape[a] for a in range(3)])
        rs = r0[pos]
        lp = L0 + 0.5 * np.sum((rs + h * psf64) ** 2 - rs ** 2)
        lm = L0 + 0.5 * np.sum((rs - h * psf64) ** 2 - rs ** 2)
        fd = (lp - lm) / (2 * h)
        assert abs(gi[i] - fd) <= 1e-3 * max(abs(fd), scale_i), (i, gi[i], fd)
    # kernel entries: h*e_j shifts the whole image into the residual
    for _ in range(100):
        j = tuple(int(rng.integers(0, k)) for k in psf.shape)
        sh = np.roll(img64, tuple(j[a] - c[a] for a in range(3)), axis=(0, 1, 2))
        lp = 0.5 * np.sum((r0 + h * sh) ** 2)
        lm = 0.5 * np.sum((r0 - h * sh) ** 2)
        fd = (lp - lm) / (2 * h)
        assert abs(gk[j] - fd) <= 1e-3 * max(abs(fd), scale_k), (j, gk[j], fd)


def test_autograd_gradcheck_small(fc):
    """torch.autograd.gradcheck on a tiny volume (the adjoints are exact, so
    FP32 finite differences agree within a loose tolerance)."""
    import torch
    from fourierconvolutioncudalib_amd.autograd import fft_conv3d
    torch.manual_seed(0)
    img = torch.rand(6, 7, 8, device="cuda", requires_grad=True)
    k = torch.rand(3, 3, 5, device="cuda", requires_grad=True)
    w = torch.randn(6, 7, 8, device="cuda")

    def f(a, b):
        return (fft_conv3d(a, b) * w).sum()

    f(img, k).backward()
    eps = 1e-2
    for t, grad in ((img, img.grad), (k, k.grad)):
        flat = t.detach().clone().view(-1)
        for idx in range(0, flat.numel(), max(1, flat.numel() // 17)):
            a = flat.clone()
            a[idx] += eps
            b = flat.clone()
            b[idx] -= eps
            with torch.no_grad():
                if t is img:
                    fp = f(a.view_as(img), k.detach())
                    fm = f(b.view_as(img), k.detach())
                else:
                    fp = f(img.detach(), a.view_as(k))
                    fm = f(img.detach(), b.view_as(k))
            fd = float((fp - fm) / (2 * eps))
            assert abs(grad.view(-1)[idx].item() - fd) <= 2e-2 * max(1.0, abs(fd))


@pytest.mark.parametrize("shape,kshape", [((6, 20, 4099), (3, 5, 7)),    # chirp-z x (one tile)
                                          ((2, 3, 20000), (1, 3, 9)),    # four-step x
                                          ((4099, 4, 6), (5, 3, 3))],    # chirp-z z
                         ids=lambda v: "x".join(map(str, v)))
def test_long_axis_gradients(fc, oracle, shape, kshape):
    """Both adjoints on shapes of the long-axis path: the PSF gradient by the
    unfused cross-correlation and the support gather, the image gradient by
    the correlation, through torch autograd, against FP64."""
    import torch
    from fourierconvolutioncudalib_amd import synth
    from fourierconvolutioncudalib_amd.autograd import fft_conv3d, grad_kernel
    print("long-axis gradients", shape, kshape, flush=True)
    rng = np.random.default_rng(sum(shape))
    img = rng.random(shape, dtype=np.float32)
    g = rng.standard_normal(shape).astype(np.float32)
    psf = synth.random_psf(kshape, 7)
    dk = grad_kernel(torch.from_numpy(g).cuda(), torch.from_numpy(img).cuda(), kshape)
    ref = oracle.fft_conv_grad_kernel(g, img, kshape)
    assert oracle.rel_l2(dk.cpu().numpy(), ref) < 1e-5
    ti = torch.from_numpy(img).cuda().requires_grad_(True)
    tk = torch.from_numpy(psf).cuda().requires_grad_(True)
    (fft_conv3d(ti, tk) * torch.from_numpy(g).cuda()).sum().backward()
    assert oracle.rel_l2(ti.grad.cpu().numpy(),
                         oracle.numpy_fft_conv(g, psf, conj_kernel=True)) < 1e-5
    assert oracle.rel_l2(tk.grad.cpu().numpy(), ref) < 1e-5
