"""Batched convolution of many uneven blocks (torch CUDA tensors).

Each block has its own image shape and kernel.  By default the library
deals the blocks to a few streams ("lanes", balanced by estimated cost),
each lane running its blocks' pass sequences in its own workspace slot;
FCAMD_BATCH_LANES=0 selects the length-group schedule instead (one
compile-time kernel launch per transform length and pass, length groups on
forked streams).  With graph replay, the first call captures the plan
(one graph per lane, or one graph for the length groups) and every call
launches the captured graphs.  Buffers are bound when the plan is created.
"""
from __future__ import annotations

import ctypes

import torch

from . import CONVOLVE, _check_tensor, _stream_handle
from ._lib import FCError, check, ints, lib


class BatchedConvolution:
    """Plan over lists of image / kernel tensors ([nz, ny, nx] / [kz, ky, kx],
    contiguous float32 on one device).  ``outs`` defaults to new tensors;
    pass the images themselves for in-place results."""

    def __init__(self, images, kernels, outs=None, op=CONVOLVE):
        if len(images) != len(kernels) or not images:
            raise ValueError("need one kernel per image")
        if outs is not None and len(outs) != len(images):
            raise ValueError("need one output per image")
        self.device = images[0].device
        for i, (a, k) in enumerate(zip(images, kernels)):
            _check_tensor(a, f"images[{i}]", device=self.device, contiguous=False)
            _check_tensor(k, f"kernels[{i}]", device=self.device, contiguous=False)
            if a.dim() != 3 or k.dim() != 3:
                raise ValueError(f"block {i}: image and kernel must be 3D")
        self.images = [t.contiguous() for t in images]
        self.kernels = [t.contiguous() for t in kernels]
        self.outs = outs if outs is not None else [torch.empty_like(t) for t in self.images]
        for i, (o, a) in enumerate(zip(self.outs, self.images)):
            _check_tensor(o, f"outs[{i}]", a.shape, self.device)
        n = len(self.images)
        imdims = ints([d for t in self.images for d in t.shape[::-1]])
        kdims = ints([d for t in self.kernels for d in t.shape[::-1]])
        ptrs = lambda ts: (ctypes.c_void_p * n)(*[t.data_ptr() for t in ts])  # noqa: E731
        with torch.cuda.device(self.device):
            self._plan = lib().fcamd_batch_create(n, imdims, kdims, ptrs(self.outs),
                                                  ptrs(self.images), ptrs(self.kernels), int(op),
                                                  None, 0)
        if not self._plan:
            check(lib().fcamd_last_error() or 6)
            raise FCError(6, "batch plan creation failed")

    def run(self, stream=None, graph=True):
        """Enqueue all blocks on ``stream`` (default: torch's current stream);
        ``graph`` launches the captured hipGraphs (captured on the first
        such call)."""
        if self._plan is None:
            raise RuntimeError("plan closed")
        with torch.cuda.device(self.device):
            h = _stream_handle(stream)
            fn = lib().fcamd_batch_graph_launch if graph else lib().fcamd_batch_execute
            check(fn(self._plan, h))
        return self.outs

    def close(self):
        if self._plan is not None:
            lib().fcamd_batch_destroy(self._plan)
            self._plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
