"""Multiview Richardson-Lucy deconvolution on one device (torch CUDA tensors).

    psi <- psi * (1/V) sum_v corr(phi_v / max(conv(psi, P_v), eps), P_v)

conv is the convolution of convolution3DfftCUDAInPlace, corr its adjoint
(the flipped PSF).  The library runs it as a fused pass sequence (one z pass
fanning the estimate out to all views, one x pass computing every view's
ratio without writing it to memory, one z pass summing the views'
correlations in the spectral domain) when all three axis lengths have
compile-time kernels, else as per-view convolutions.
"""
from __future__ import annotations

import ctypes

import torch

from . import _check_tensor, _stream_handle, _workspace
from ._lib import check, ints, lib


def _args(views, kernels):
    if not views or len(views) != len(kernels):
        raise ValueError("need one PSF per view")
    device = views[0].device
    shape = tuple(views[0].shape)
    for i, (v, k) in enumerate(zip(views, kernels)):
        _check_tensor(v, f"views[{i}]", shape, device, contiguous=False)
        _check_tensor(k, f"kernels[{i}]", device=device, contiguous=False)
        if v.dim() != 3 or k.dim() != 3:
            raise ValueError(f"view {i}: view and PSF must be 3D")
    views = [v.contiguous() for v in views]
    kernels = [k.contiguous() for k in kernels]
    imdim = ints(shape[::-1])
    kdims = ints([d for k in kernels for d in k.shape[::-1]])
    return views, kernels, shape, imdim, kdims


def fused_available(views, kernels):
    _, _, _, imdim, kdims = _args(views, kernels)
    return bool(lib().fcamd_rl_fused_available(len(views), imdim, kdims))


def richardson_lucy(views, kernels, iterations=10, eps=1e-6, psi=None, fused=True, stream=None):
    """Deconvolve ``views`` (list of [nz, ny, nx] tensors) with their PSFs;
    starts from the mean of the views unless ``psi`` is given (updated in
    place).  Returns the estimate."""
    views, kernels, shape, imdim, kdims = _args(views, kernels)
    n = len(views)
    init = psi is None
    if init:
        psi = torch.empty(shape, dtype=torch.float32, device=views[0].device)
    else:
        _check_tensor(psi, "psi", shape, views[0].device)
    vptr = (ctypes.c_void_p * n)(*[v.data_ptr() for v in views])
    kptr = (ctypes.c_void_p * n)(*[k.data_ptr() for k in kernels])
    with torch.cuda.device(psi.device):
        nbytes = lib().fcamd_rl_workspace_bytes(n, imdim, kdims, 1 if fused else 0)
        if nbytes < 0:
            check(lib().fcamd_last_error())
        wptr, wbytes, _keep = _workspace(nbytes, None, psi.device, stream)
        check(lib().fcamd_rl_multiview_device(psi.data_ptr(), vptr, n, imdim, kptr, kdims,
                                              int(iterations), float(eps), 1 if init else 0,
                                              1 if fused else 0, wptr, wbytes,
                                              _stream_handle(stream)))
    return psi
