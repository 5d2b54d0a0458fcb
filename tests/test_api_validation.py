"""Tensor checks of the torch-facing wrappers: a wrong dtype, a host tensor,
a wrong shape or a non-contiguous output is rejected before any library
call, so no such pointer reaches a device kernel.  The library handle is
replaced by a sentinel that fails the test if anything calls it."""
import pytest


class _NoLib:
    def __getattr__(self, name):
        raise AssertionError(f"library entry point {name} called before validation")


@pytest.fixture
def nolib(monkeypatch):
    import fourierconvolutioncudalib_amd as fc
    from fourierconvolutioncudalib_amd import autograd, batched, rl
    sentinel = lambda: _NoLib()  # noqa: E731
    for mod in (fc, batched, rl, autograd):
        monkeypatch.setattr(mod, "lib", sentinel, raising=False)
    monkeypatch.setattr(fc._lib, "lib", sentinel)
    return fc


def test_host_tensors_rejected(nolib):
    import torch
    fc = nolib
    from fourierconvolutioncudalib_amd.autograd import grad_kernel
    from fourierconvolutioncudalib_amd.batched import BatchedConvolution
    from fourierconvolutioncudalib_amd.rl import richardson_lucy
    a = torch.zeros(4, 5, 6)
    k = torch.zeros(3, 3, 3)
    with pytest.raises(TypeError, match="CUDA"):
        fc.convolve3d(a, k)
    with pytest.raises(TypeError, match="CUDA"):
        fc.convolve3d_boundary(a, k, "zero")
    with pytest.raises(TypeError, match="CUDA"):
        fc.ConvolutionPlan(a.shape, k)
    with pytest.raises(TypeError, match="CUDA"):
        BatchedConvolution([a], [k])
    with pytest.raises(TypeError, match="CUDA"):
        richardson_lucy([a], [k], iterations=1)
    with pytest.raises(TypeError, match="CUDA"):
        grad_kernel(a, a, k.shape)


def test_wrong_dtype_rejected(nolib):
    import torch
    fc = nolib
    with pytest.raises(TypeError, match="float32"):
        fc.convolve3d(torch.zeros(4, 5, 6, dtype=torch.float16), torch.zeros(3, 3, 3))
    with pytest.raises(TypeError, match="float32"):
        fc.convolve3d(torch.zeros(4, 5, 6, dtype=torch.float64), torch.zeros(3, 3, 3))
    with pytest.raises(TypeError, match="torch tensor"):
        fc.convolve3d([[[0.0]]], torch.zeros(1, 1, 1))


@pytest.mark.gpu
def test_device_outputs_rejected(nolib):
    """CUDA inputs with a bad output: half precision, host memory, wrong
    shape, non-contiguous; also a bad psi and bad batched outputs."""
    import torch
    fc = nolib
    from fourierconvolutioncudalib_amd.batched import BatchedConvolution
    from fourierconvolutioncudalib_amd.rl import richardson_lucy
    a = torch.rand(8, 10, 12, device="cuda")
    k = torch.rand(3, 3, 3, device="cuda")
    bad = [
        (TypeError, torch.empty(8, 10, 12, dtype=torch.float16, device="cuda")),
        (TypeError, torch.empty(8, 10, 12)),
        (ValueError, torch.empty(8, 10, 11, device="cuda")),
        (ValueError, torch.empty(12, 10, 8, device="cuda").permute(2, 1, 0)),
    ]
    for exc, out in bad:
        with pytest.raises(exc):
            fc.convolve3d(a, k, out=out)
        with pytest.raises(exc):
            fc.convolve3d_boundary(a, k, "reflect", out=out)
        with pytest.raises(exc):
            BatchedConvolution([a], [k], outs=[out])
        with pytest.raises(exc):
            richardson_lucy([a], [k], iterations=1, psi=out)
