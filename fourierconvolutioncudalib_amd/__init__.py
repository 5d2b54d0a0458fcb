"""MI355X-native 3D FFT convolution (drop-in for libFourierConvolutionCUDALib).

Two Python faces over the same C ABI (``include/*.h``):

* a host mirror of the reference interface (``src/convolution3Dfft.h``):
  ``convolution3DfftCUDAInPlace(im, imDim, kernel, kernelDim, devCUDA)`` and
  friends, on numpy float32 buffers, with the reference's argument meaning
  (dims ``{x, y, z}``, x contiguous).  Failures raise :class:`FCError`
  (the reference prints and calls ``exit``, ``src/book.h:22-30``);
* a device API on torch CUDA tensors (``convolve3d``), asynchronous on a
  stream, used by the benchmark, the autograd wrapper and the batched /
  deconvolution drivers.

Volumes are C-order ``[nz, ny, nx]`` arrays.
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import FCError, check, ints, lib

__all__ = [
    "FCError", "lib",
    "convolution3DfftCUDAInPlace", "convolution3DfftCUDA", "convolution3DfftCUDA_test",
    "selectDeviceWithHighestComputeCapability", "getCUDAcomputeCapabilityMajorVersion",
    "getCUDAcomputeCapabilityMinorVersion", "getNumDevicesCUDA", "getNameDeviceCUDA",
    "getMemDeviceCUDA", "cuda_version", "gpu_mem_needed_mb",
    "convolve", "convolve3d", "workspace_bytes", "radix_plan",
    "CONVOLVE", "CORRELATE",
]

CONVOLVE = 0
CORRELATE = 1

_fp = ctypes.POINTER(ctypes.c_float)


def _f32_buffer(a, name):
    if not isinstance(a, np.ndarray) or a.dtype != np.float32 or not a.flags.c_contiguous:
        raise TypeError(f"{name} must be a C-contiguous float32 numpy array")
    return a.ctypes.data_as(_fp)


def _raise_if_failed():
    code = lib().fcamd_last_error()
    if code != 0:
        check(code)


# ---------------------------------------------------------------------------
# reference interface (host pointers)
# ---------------------------------------------------------------------------
def convolution3DfftCUDAInPlace(im, imDim, kernel, kernelDim, devCUDA=0):
    """src/convolution3Dfft.h:56 -- overwrite ``im`` with its circular
    convolution by ``kernel`` (zero-padded, centred at floor(k/2))."""
    if int(np.prod(imDim)) != im.size or int(np.prod(kernelDim)) != kernel.size:
        raise ValueError("dims do not match buffer sizes")
    lib().convolution3DfftCUDAInPlace(_f32_buffer(im, "im"), ints(imDim),
                                      _f32_buffer(kernel, "kernel"), ints(kernelDim),
                                      int(devCUDA))
    _raise_if_failed()


def _take_result(ptr, n):
    if not ptr:
        _raise_if_failed()
        raise FCError(6, "no result returned")
    try:
        return np.ctypeslib.as_array(ptr, shape=(n,)).copy()
    finally:
        lib().fcamd_free_result(ptr)


def convolution3DfftCUDA(im, imDim, kernel, kernelDim, devCUDA=0):
    """src/convolution3Dfft.h:41-45 -- dims z-fastest (imDim[2] contiguous);
    returns a new flat float32 array, ``im`` untouched."""
    if int(np.prod(imDim)) != im.size or int(np.prod(kernelDim)) != kernel.size:
        raise ValueError("dims do not match buffer sizes")
    ptr = lib().convolution3DfftCUDA(_f32_buffer(im, "im"), ints(imDim),
                                     _f32_buffer(kernel, "kernel"), ints(kernelDim),
                                     int(devCUDA))
    return _take_result(ptr, im.size)


def convolution3DfftCUDA_test(im, imDim, kernel, devCUDA=0):
    """src/convolution3Dfft.h:29-32 -- image-sized, unshifted kernel."""
    if int(np.prod(imDim)) != im.size or kernel.size != im.size:
        raise ValueError("dims do not match buffer sizes")
    ptr = lib().convolution3DfftCUDA_test(_f32_buffer(im, "im"), ints(imDim),
                                          _f32_buffer(kernel, "kernel"), int(devCUDA))
    return _take_result(ptr, im.size)


def selectDeviceWithHighestComputeCapability():
    return lib().selectDeviceWithHighestComputeCapability()


def getCUDAcomputeCapabilityMajorVersion(dev):
    return lib().getCUDAcomputeCapabilityMajorVersion(int(dev))


def getCUDAcomputeCapabilityMinorVersion(dev):
    return lib().getCUDAcomputeCapabilityMinorVersion(int(dev))


def getNumDevicesCUDA():
    return lib().getNumDevicesCUDA()


def getNameDeviceCUDA(dev):
    buf = ctypes.create_string_buffer(256)
    lib().getNameDeviceCUDA(int(dev), buf)
    return buf.value.decode(errors="replace")


def getMemDeviceCUDA(dev):
    return lib().getMemDeviceCUDA(int(dev))


def cuda_version():
    return lib().cuda_version()


def gpu_mem_needed_mb(shape):
    return lib().gpu_mem_needed_mb(ints(shape), len(shape))


# ---------------------------------------------------------------------------
# conveniences
# ---------------------------------------------------------------------------
def convolve(image, kernel, op=CONVOLVE, dev=0):
    """Circular convolution of a 3D numpy volume (host path); returns a new
    float32 array.  ``op=CORRELATE`` gives the adjoint (flipped kernel)."""
    im = np.array(image, dtype=np.float32, order="C", copy=True)
    k = np.ascontiguousarray(kernel, dtype=np.float32)
    if im.ndim != 3 or k.ndim != 3:
        raise ValueError("image and kernel must be 3D")
    check(lib().fcamd_convolve3d_host(_f32_buffer(im, "image"), ints(im.shape[::-1]),
                                      _f32_buffer(k, "kernel"), ints(k.shape[::-1]),
                                      int(op), int(dev)))
    return im


def workspace_bytes(image_shape, kernel_shape):
    """Device workspace for one convolution of C-order shapes (z, y, x)."""
    n = lib().fcamd_workspace_bytes(ints(tuple(image_shape)[::-1]), ints(tuple(kernel_shape)[::-1]))
    if n < 0:
        _raise_if_failed()
    return int(n)


def radix_plan(length):
    buf = (ctypes.c_int * 64)()
    n = lib().fcamd_radix_plan(int(length), buf, 64)
    return [buf[i] for i in range(n)]


def static_lengths():
    """Transform lengths served by compile-time (register-resident) kernels."""
    buf = (ctypes.c_int * 256)()
    n = lib().fcamd_static_lengths(buf, 256)
    return [buf[i] for i in range(min(n, 256))]


def use_static_kernels(enable: bool) -> None:
    """False routes every pass through the runtime-length kernels."""
    lib().fcamd_use_static_kernels(1 if enable else 0)


def autotune(image_shape, kernel_shape, reps=5, stream=None):
    """Time every compile-time kernel variant for this convolution shape
    (C-order) on the current device and keep the fastest per pass."""
    check(lib().fcamd_autotune(ints(tuple(image_shape)[::-1]), ints(tuple(kernel_shape)[::-1]),
                               int(reps), _stream_handle(stream)))


def selected_variants(image_shape):
    """{pass name: variant name} currently selected for a C-order shape."""
    nz, ny, nx = image_shape
    out = {}
    for p, (name, L) in enumerate([("x_forward", nx), ("x_inverse", nx), ("y_forward", ny),
                                   ("z", nz), ("y_inverse", ny)]):
        i = lib().fcamd_selected_static_variant(p, L)
        if i >= 0:
            out[name] = lib().fcamd_static_variant_name(p, L, i).decode()
    return out


def _check_tensor(t, name, shape=None, device=None, contiguous=True):
    """Raises before any library call unless ``t`` is a float32 CUDA tensor
    (on ``device``, of exactly ``shape``, contiguous when asked): a pointer
    of another dtype, a host pointer or a short buffer would otherwise reach
    a device kernel."""
    import torch
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch tensor, got {type(t).__name__}")
    if t.dtype != torch.float32:
        raise TypeError(f"{name} must be float32, got {t.dtype}")
    if not t.is_cuda:
        raise TypeError(f"{name} must be a CUDA tensor (there is no CPU path)")
    if device is not None and t.device != torch.device(device):
        raise ValueError(f"{name} is on {t.device}, expected {device}")
    if shape is not None and tuple(t.shape) != tuple(int(d) for d in shape):
        raise ValueError(f"{name} has shape {tuple(t.shape)}, expected {tuple(shape)}")
    if contiguous and not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    return t


def _workspace(nbytes, workspace, device, stream):
    """(pointer, bytes, keep-alive tensor) of the workspace for one call:
    the caller's, else a new one from torch's caching allocator (no
    synchronisation, usable under graph capture).  A new workspace used on a
    stream other than torch's current one is recorded on that stream so the
    allocator does not hand it out again before the call's kernels ran."""
    import torch
    if workspace is not None:
        if not isinstance(workspace, torch.Tensor) or not workspace.is_cuda:
            raise TypeError("workspace must be a CUDA tensor")
        if workspace.device != device or not workspace.is_contiguous():
            raise ValueError("workspace must be a contiguous tensor on the image's device")
        have = workspace.numel() * workspace.element_size()
        if have < nbytes:
            raise ValueError(f"workspace holds {have} bytes, {nbytes} needed")
        return workspace.data_ptr(), have, workspace
    ws = torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device)
    if stream is not None:
        if isinstance(stream, int):
            stream = torch.cuda.ExternalStream(stream, device=device)
        if stream.cuda_stream != torch.cuda.current_stream(device).cuda_stream:
            ws.record_stream(stream)
    return ws.data_ptr(), int(nbytes), ws


def _stream_handle(stream):
    if stream is None:
        import torch
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    if isinstance(stream, int):
        return ctypes.c_void_p(stream)
    return ctypes.c_void_p(stream.cuda_stream)


class ConvolutionPlan:
    """One kernel, many images of one shape: the kernel's rows and planes
    are transformed once (fcamd_conv_plan_create) and every run() executes
    only the image's passes.  Results equal convolve3d with the same kernel
    and op.  Runs of one plan share its workspace: keep them on one stream
    (or synchronise between streams)."""

    def __init__(self, image_shape, kernel, op=CONVOLVE, stream=None, workspace=None):
        import torch
        _check_tensor(kernel, "kernel", contiguous=False)
        if kernel.dim() != 3:
            raise ValueError("kernel must be 3D")
        self.image_shape = tuple(int(d) for d in image_shape)
        self.kernel_shape = tuple(kernel.shape)
        self.device = kernel.device
        self._kernel = kernel.contiguous()
        self._workspace = workspace
        wptr, wbytes = None, 0
        if workspace is not None:
            wptr, wbytes, _ = _workspace(workspace_bytes(self.image_shape, self.kernel_shape),
                                         workspace, self.device, stream)
        with torch.cuda.device(self.device):
            self._plan = lib().fcamd_conv_plan_create(
                ints(self.image_shape[::-1]), ints(self.kernel_shape[::-1]),
                self._kernel.data_ptr(), int(op), wptr, wbytes, _stream_handle(stream))
        if not self._plan:
            check(lib().fcamd_last_error())

    def set_kernel(self, kernel, stream=None):
        """Replace the kernel (same dims); transformed on ``stream``."""
        _check_tensor(kernel, "kernel", self.kernel_shape, self.device, contiguous=False)
        import torch
        self._kernel = kernel.contiguous()
        with torch.cuda.device(self.device):
            check(lib().fcamd_conv_plan_set_kernel(self._plan, self._kernel.data_ptr(),
                                                   _stream_handle(stream)))

    def run(self, image, out=None, stream=None):
        """Convolve one image of the plan's shape; ``out`` may be ``image``."""
        import torch
        _check_tensor(image, "image", self.image_shape, self.device, contiguous=False)
        image = image.contiguous()
        if out is None:
            out = torch.empty_like(image)
        _check_tensor(out, "out", self.image_shape, self.device)
        with torch.cuda.device(self.device):
            check(lib().fcamd_conv_plan_execute(self._plan, out.data_ptr(), image.data_ptr(),
                                                _stream_handle(stream)))
        return out

    def close(self):
        if getattr(self, "_plan", None):
            lib().fcamd_conv_plan_destroy(self._plan)
            self._plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


#: boundary rules of convolve3d_boundary (FCAMD_BOUNDARY_* in the C header)
BOUNDARY = {"periodic": 0, "zero": 1, "nearest": 2, "reflect": 3, "mirror": 4}


def _boundary_code(boundary):
    if isinstance(boundary, str):
        if boundary not in BOUNDARY:
            raise ValueError(f"unknown boundary {boundary!r}; one of {sorted(BOUNDARY)}")
        return BOUNDARY[boundary]
    return int(boundary)


def padded_extents(image_shape, kernel_shape, boundary):
    """C-order transform shape a boundary-mode convolution runs at."""
    p = (ctypes.c_int * 3)()
    check(lib().fcamd_padded_extents(ints(tuple(image_shape)[::-1]),
                                     ints(tuple(kernel_shape)[::-1]),
                                     _boundary_code(boundary), p))
    return (p[2], p[1], p[0])


def boundary_workspace_bytes(image_shape, kernel_shape, boundary):
    """Device workspace bytes of convolve3d_boundary for C-order shapes."""
    r = lib().fcamd_boundary_workspace_bytes(ints(tuple(image_shape)[::-1]),
                                             ints(tuple(kernel_shape)[::-1]),
                                             _boundary_code(boundary))
    if r < 0:
        check(lib().fcamd_last_error())
    return int(r)


def convolve3d_boundary(image, kernel, boundary="zero", op=CONVOLVE, out=None, workspace=None,
                        stream=None):
    """Linear convolution of a torch CUDA float32 image extended beyond its
    edges by ``boundary`` ('zero', 'nearest', 'reflect', 'mirror'; 'periodic'
    is convolve3d).  Kernel centred at k/2, any size; ``out`` may be
    ``image``.  The padding is virtual: no padded copy of the image is
    made."""
    import torch
    _check_tensor(image, "image", contiguous=False)
    _check_tensor(kernel, "kernel", device=image.device, contiguous=False)
    if image.dim() != 3 or kernel.dim() != 3:
        raise ValueError("image and kernel must be 3D")
    image = image.contiguous()
    kernel = kernel.contiguous()
    if out is None:
        out = torch.empty_like(image)
    _check_tensor(out, "out", image.shape, image.device)
    need = boundary_workspace_bytes(image.shape, kernel.shape, boundary)
    wptr, wbytes, _keep = _workspace(need, workspace, image.device, stream)
    with torch.cuda.device(image.device):
        check(lib().fcamd_convolve3d_boundary_device(
            out.data_ptr(), image.data_ptr(), ints(image.shape[::-1]),
            kernel.data_ptr(), ints(kernel.shape[::-1]), int(op), _boundary_code(boundary),
            wptr, wbytes, _stream_handle(stream)))
    return out


def convolve3d(image, kernel, op=CONVOLVE, out=None, workspace=None, stream=None):
    """Device-resident convolution of torch CUDA float32 tensors.

    image [nz, ny, nx], kernel [kz, ky, kx] (same device); ``out`` may be
    ``image`` (in place).  ``workspace`` is a caller's buffer of
    workspace_bytes() bytes, else one comes from torch's caching allocator.
    Enqueued on ``stream`` (default: torch's current stream) without
    synchronising; returns ``out``."""
    import torch
    _check_tensor(image, "image", contiguous=False)
    _check_tensor(kernel, "kernel", device=image.device, contiguous=False)
    if image.dim() != 3 or kernel.dim() != 3:
        raise ValueError("image and kernel must be 3D")
    image = image.contiguous()
    kernel = kernel.contiguous()
    if out is None:
        out = torch.empty_like(image)
    _check_tensor(out, "out", image.shape, image.device)
    need = workspace_bytes(image.shape, kernel.shape)
    wptr, wbytes, _keep = _workspace(need, workspace, image.device, stream)
    with torch.cuda.device(image.device):
        check(lib().fcamd_convolve3d_device(
            out.data_ptr(), image.data_ptr(), ints(image.shape[::-1]),
            kernel.data_ptr(), ints(kernel.shape[::-1]), int(op),
            wptr, wbytes, _stream_handle(stream)))
    return out
