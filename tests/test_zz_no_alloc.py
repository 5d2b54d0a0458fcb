"""fcamd_convolve3d_device_noalloc, the allocation-free entry point, through
the C ABI against fcamd_convolve3d_device (the entry point it mirrors) and
the FP64 oracle: calls in a row with shrinking inputs (the shared scratch
keeps the last call's contents), in place, correlation, compile-time and
runtime lengths, the workspace limit, and first use of new lengths recorded
into a CUDA graph (where any allocation would fail the capture)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _noalloc(fc, out, img, k, op=0, stream=None):
    from fourierconvolutioncudalib_amd._lib import ints
    s = ctypes.c_void_p(stream.cuda_stream) if stream is not None else None
    return fc.lib().fcamd_convolve3d_device_noalloc(
        out.data_ptr(), img.data_ptr(), ints(img.shape[::-1]), k.data_ptr(),
        ints(k.shape[::-1]), int(op), s)


def test_noalloc_matches_device_entry_point(fc, oracle):
    import torch
    from fourierconvolutioncudalib_amd import synth
    cases = [((64, 128, 128), (15, 15, 15), 0),  # whole-plane x/y kernels
             ((32, 256, 256), (9, 9, 9), 0),      # wide planes, half-length tables
             ((64, 256, 256), (21, 21, 21), 1),
             ((128, 256, 256), (21, 21, 21), 0),  # x and y forward passes, wide inverse
             ((64, 96, 128), (9, 7, 5), 0),   # compile-time lengths
             ((40, 48, 64), (5, 5, 3), 1),    # smaller, correlation
             ((30, 22, 26), (5, 3, 3), 0),    # runtime lengths, smaller again
             ((12, 10, 14), (3, 3, 5), 0)]
    for i, (shape, kshape, op) in enumerate(cases):
        print("noalloc", shape, kshape, op, flush=True)
        a = torch.from_numpy(synth.uniform(shape, 900 + i)).cuda()
        k = torch.from_numpy(synth.random_psf(kshape, 910 + i)).cuda()
        out = torch.full(shape, -1.0, device="cuda")
        assert _noalloc(fc, out, a, k, op) == 0, fc.lib().fcamd_last_error_message()
        ref_dev = fc.convolve3d(a, k, op=op)
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        assert oracle.rel_l2(got, ref_dev.cpu().numpy()) < 1e-6, shape
        ref = oracle.numpy_fft_conv(a.cpu().numpy(), k.cpu().numpy(), conj_kernel=op == 1)
        assert oracle.rel_l2(got, ref) < TOL, shape
        # in place
        b = a.clone()
        assert _noalloc(fc, b, b, k, op) == 0
        torch.cuda.synchronize()
        assert oracle.rel_l2(b.cpu().numpy(), ref) < TOL, shape


def test_noalloc_workspace_limit(fc):
    """A convolution whose workspace exceeds FCAMD_NOALLOC_MAX_WORKSPACE is
    refused before any launch; the output is untouched."""
    import torch
    shape, kshape = (256, 512, 512), (3, 3, 3)
    assert fc.workspace_bytes(shape, kshape) > (128 << 20)
    a = torch.zeros(shape, device="cuda")
    k = torch.ones(kshape, device="cuda")
    out = torch.full(shape, 7.0, device="cuda")
    rc = _noalloc(fc, out, a, k)
    assert rc == 2 and b"allocation-free" in fc.lib().fcamd_last_error_message()
    torch.cuda.synchronize()
    assert bool((out == 7.0).all())
    assert _noalloc(fc, out, a, k, op=5) == 1  # bad op


def test_noalloc_first_use_in_cuda_graph(fc, oracle):
    """Lengths first used inside a capture: the call records the table
    generation and the passes, allocates nothing, and every replay (new data
    and eager work of the program between replays) matches the eager
    result."""
    import torch
    from fourierconvolutioncudalib_amd import synth
    warm = torch.rand(8, 8, 8, device="cuda")
    kw = torch.rand(3, 3, 3, device="cuda")
    assert _noalloc(fc, torch.empty_like(warm), warm, kw) == 0  # device set-up outside
    torch.cuda.synchronize()
    shape, kshape = (27, 33, 35), (5, 3, 7)
    k = torch.from_numpy(synth.random_psf(kshape, 77)).cuda()
    src = torch.empty(shape, device="cuda")
    out = torch.empty(shape, device="cuda")
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    rcs = []
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            rcs.append(_noalloc(fc, out, src, k, 0, stream=s))
    assert rcs == [0], fc.lib().fcamd_last_error_message()
    torch.cuda.synchronize()
    for i in range(3):
        d = torch.from_numpy(synth.uniform(shape, 50 + i)).cuda()
        src.copy_(d)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        eager = fc.convolve3d(d, k).cpu().numpy()
        ref = oracle.numpy_fft_conv(d.cpu().numpy(), k.cpu().numpy())
        assert oracle.rel_l2(got, eager) < 1e-6 and oracle.rel_l2(got, ref) < TOL, i
        src.mul_(0.5)  # eager work between replays
        torch.cuda.synchronize()
