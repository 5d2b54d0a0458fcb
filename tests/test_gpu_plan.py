"""Convolution plans (kernel rows/planes transformed once): equal to the
plain device call for every image, after a kernel change, for the
correlation, on compile-time and runtime lengths, with a caller workspace,
and in place; on the long-axis path (axes past one tile, large prime
factors) the plan keeps the kernel's whole spectrum and still equals the
plain call."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape,kshape", [((64, 96, 80), (9, 7, 5)), ((13, 50, 36), (3, 5, 9)),
                                          ((256, 128, 96), (15, 15, 15)), ((1, 40, 48), (1, 9, 7)),
                                          # wide planes: both directions split; forward
                                          # passes with the split inverse
                                          ((32, 256, 256), (9, 9, 9)),
                                          ((128, 256, 256), (21, 21, 21)),
                                          # long-axis path: four-step x, one-tile
                                          # chirp-z x, four-step chirp-z y
                                          ((2, 3, 20000), (1, 3, 9)),
                                          ((8, 64, 4099), (3, 5, 7)),
                                          ((2, 10247, 3), (1, 9, 3))])
def test_plan_equals_plain_calls(fc, oracle, shape, kshape):
    import torch
    from fourierconvolutioncudalib_amd import synth
    k = torch.from_numpy(synth.random_psf(kshape, 11)).cuda()
    plan = fc.ConvolutionPlan(shape, k)
    for i in range(3):
        a = torch.from_numpy(synth.uniform(shape, 20 + i)).cuda()
        got = plan.run(a)
        ref = fc.convolve3d(a, k)
        torch.cuda.synchronize()
        assert torch.equal(got, ref), i
    a = synth.uniform(shape, 31)
    assert oracle.rel_l2(plan.run(torch.from_numpy(a).cuda()).cpu().numpy(),
                         oracle.numpy_fft_conv(a, k.cpu().numpy())) < 1e-5
    plan.close()


def test_plan_kernel_change_correlate_workspace_inplace(fc):
    import torch
    from fourierconvolutioncudalib_amd import synth
    shape, kshape = (40, 64, 72), (7, 9, 5)
    k1 = torch.from_numpy(synth.random_psf(kshape, 1)).cuda()
    k2 = torch.from_numpy(synth.random_psf(kshape, 2)).cuda()
    ws = torch.empty(fc.workspace_bytes(shape, kshape), dtype=torch.uint8, device="cuda")
    plan = fc.ConvolutionPlan(shape, k1, op=fc.CORRELATE, workspace=ws)
    a = torch.from_numpy(synth.uniform(shape, 5)).cuda()
    assert torch.equal(plan.run(a), fc.convolve3d(a, k1, op=fc.CORRELATE))
    plan.set_kernel(k2)
    b = a.clone()
    plan.run(b, out=b)
    assert torch.equal(b, fc.convolve3d(a, k2, op=fc.CORRELATE))
    small = torch.empty(ws.numel() - 256, dtype=torch.uint8, device="cuda")
    with pytest.raises(ValueError, match="workspace"):  # the wrapper's check
        fc.ConvolutionPlan(shape, k1, workspace=small)
    from fourierconvolutioncudalib_amd._lib import ints
    assert not fc.lib().fcamd_conv_plan_create(  # the C entry point's own check
        ints(shape[::-1]), ints(k1.shape[::-1]), k1.data_ptr(), 0, small.data_ptr(),
        small.numel(), None)
    assert fc.lib().fcamd_last_error() == 1
    with pytest.raises(ValueError):
        plan.run(torch.zeros((40, 64, 71), device="cuda"))
    plan.close()


def test_long_axis_plan_kernel_change_correlate_inplace(fc, oracle):
    """A long-axis plan (x = 4099, chirp-z) after set_kernel, as a
    correlation, run in place on a caller workspace."""
    import torch
    from fourierconvolutioncudalib_amd import synth
    shape, kshape = (6, 20, 4099), (3, 5, 7)
    k1 = torch.from_numpy(synth.random_psf(kshape, 3)).cuda()
    k2 = torch.from_numpy(synth.random_psf(kshape, 4)).cuda()
    ws = torch.empty(fc.workspace_bytes(shape, kshape), dtype=torch.uint8, device="cuda")
    plan = fc.ConvolutionPlan(shape, k1, op=fc.CORRELATE, workspace=ws)
    a_np = synth.uniform(shape, 6)
    a = torch.from_numpy(a_np).cuda()
    assert torch.equal(plan.run(a), fc.convolve3d(a, k1, op=fc.CORRELATE))
    plan.set_kernel(k2)
    b = a.clone()
    plan.run(b, out=b)
    assert torch.equal(b, fc.convolve3d(a, k2, op=fc.CORRELATE))
    ref = oracle.numpy_fft_conv(a_np, k2.cpu().numpy(), conj_kernel=True)
    assert oracle.rel_l2(b.cpu().numpy(), ref) < 1e-5
    plan.close()
