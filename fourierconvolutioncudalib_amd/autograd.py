"""Differentiable 3D FFT convolution for torch (thin wrapper over the C ABI).

forward:  out = conv(image, kernel)            (fcamd_convolve3d_device)
backward: d image  = corr(g, kernel)           (same entry point, FCAMD_CORRELATE:
                                                the conjugated kernel spectrum)
          d kernel = sum_i g[i] image[i - (j - c)]  (fcamd_conv_grad_kernel_device)

Both adjoints are hand-written HIP passes; torch only provides the tensors,
the stream and the autograd graph.  The reference has no gradient path; its
forward semantics are src/convolution3Dfft.cu:399-578.
"""
from __future__ import annotations

import torch

from . import CONVOLVE, CORRELATE, convolve3d
from ._lib import check, ints, lib


def grad_kernel(grad_out, image, kernel_shape, workspace=None, stream=None):
    """dk[j] = sum_i g[i] * image[(i - (j - k//2)) mod n] on a kernel of
    C-order shape ``kernel_shape`` (torch CUDA float32)."""
    from . import _check_tensor, _stream_handle, _workspace
    _check_tensor(image, "image", contiguous=False)
    _check_tensor(grad_out, "grad_out", image.shape, image.device, contiguous=False)
    if image.dim() != 3 or len(kernel_shape) != 3:
        raise ValueError("image and kernel must be 3D")
    grad_out = grad_out.contiguous()
    image = image.contiguous()
    dk = torch.empty(tuple(kernel_shape), dtype=torch.float32, device=image.device)
    need = grad_kernel_workspace_bytes(image.shape, kernel_shape)
    wptr, wbytes, _keep = _workspace(need, workspace, image.device, stream)
    with torch.cuda.device(image.device):
        check(lib().fcamd_conv_grad_kernel_device(
            dk.data_ptr(), grad_out.data_ptr(), image.data_ptr(), ints(image.shape[::-1]),
            ints(tuple(kernel_shape)[::-1]), wptr, wbytes, _stream_handle(stream)))
    return dk


def grad_kernel_workspace_bytes(image_shape, kernel_shape):
    n = lib().fcamd_grad_kernel_workspace_bytes(ints(tuple(image_shape)[::-1]),
                                                ints(tuple(kernel_shape)[::-1]))
    if n < 0:
        check(lib().fcamd_last_error())
    return int(n)


class FFTConv3dFunction(torch.autograd.Function):
    """conv(image, kernel) with hand-written HIP adjoints for both inputs."""

    @staticmethod
    def forward(ctx, image, kernel):
        ctx.save_for_backward(image, kernel)
        return convolve3d(image.detach(), kernel.detach(), op=CONVOLVE)

    @staticmethod
    def backward(ctx, g):
        image, kernel = ctx.saved_tensors
        g = g.contiguous()
        gi = gk = None
        if ctx.needs_input_grad[0]:
            gi = convolve3d(g, kernel.detach(), op=CORRELATE)
        if ctx.needs_input_grad[1]:
            gk = grad_kernel(g, image.detach(), kernel.shape)
        return gi, gk


def fft_conv3d(image, kernel):
    """Differentiable circular convolution (same semantics as
    convolution3DfftCUDAInPlace): image [nz, ny, nx], kernel [kz, ky, kx]."""
    return FFTConv3dFunction.apply(image, kernel)
