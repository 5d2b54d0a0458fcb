"""CPU oracle for the 3D FFT convolution hot path (ctypes over liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py``, always as the checker and never as
the thing measured or shipped.  The product path never imports this module.

Arrays are numpy C-order ``[nz, ny, nx]`` volumes (x fastest), matching the
reference's memory layout (``src/convolution3Dfft.cu:394-396``).
Citations are relative to the reference repository root.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

_i3 = ctypes.POINTER(ctypes.c_int)
_fp = ctypes.POINTER(ctypes.c_float)
_dp = ctypes.POINTER(ctypes.c_double)


def build() -> None:
    """Compile liboracle.so (and _ref/ when the reference tree is present)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        lib = ctypes.CDLL(path)
        lib.fco_convolve_host.argtypes = [_fp, _i3, _fp, _i3, _fp, _i3]
        lib.fco_fftshift_pad.argtypes = [_fp, _i3, _dp, _i3]
        lib.fco_fftshift_pad.restype = ctypes.c_int
        lib.fco_circular_conv_direct.argtypes = [_fp, _i3, _fp, _i3, _dp]
        lib.fco_circular_conv_direct.restype = ctypes.c_int
        lib.fco_boundary_conv_direct.argtypes = [_fp, _i3, _fp, _i3, ctypes.c_int,
                                                 ctypes.c_int, _dp]
        lib.fco_boundary_conv_direct.restype = ctypes.c_int
        lib.fco_fft1d.argtypes = [_dp, ctypes.c_int, ctypes.c_long, ctypes.c_int]
        lib.fco_fft3d.argtypes = [_dp, _i3, ctypes.c_int]
        lib.fco_fft_conv.argtypes = [_fp, _i3, _fp, _i3, _dp, ctypes.c_int]
        lib.fco_fft_conv.restype = ctypes.c_int
        lib.fco_conv_grad_kernel.argtypes = [_dp, _fp, _i3, _i3, _dp]
        lib.fco_conv_grad_kernel.restype = ctypes.c_int
        lib.fco_zero_padd_extents.argtypes = [_i3, _i3, ctypes.c_int, _i3, _i3]
        lib.fco_insert_at_offsets.argtypes = [_fp, _i3, _fp, _i3, _i3]
        lib.fco_l2norm_ref.argtypes = [_fp, _fp, ctypes.c_long]
        lib.fco_l2norm_ref.restype = ctypes.c_double
        lib.fco_rl_multiview.argtypes = [_fp, ctypes.c_int, _i3, _fp, _i3,
                                         ctypes.c_int, ctypes.c_double, _dp]
        lib.fco_rl_multiview.restype = ctypes.c_int
        _LIB = lib
    return _LIB


def ref_available() -> bool:
    return os.path.exists(os.path.join(_HERE, "_ref", "libref_convolve.so"))


def _ref():
    global _REF
    if _REF is None:
        lib = ctypes.CDLL(os.path.join(_HERE, "_ref", "libref_convolve.so"))
        lib.ref_convolve.argtypes = [_fp, _i3, _fp, _i3, _fp, _i3]
        _REF = lib
    return _REF


def _ints(v):
    return (ctypes.c_int * len(v))(*[int(x) for x in v])


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _ptr(a, t):
    return a.ctypes.data_as(t)


# --------------------------------------------------------------------------
# restatements
# --------------------------------------------------------------------------
def convolve_host(image, kernel, offsets, result=None):
    """tests/test_algorithms.hpp:9-58 restated in C (bit-identical)."""
    image = _f32(image)
    kernel = _f32(kernel)
    out = np.zeros_like(image) if result is None else result
    _lib().fco_convolve_host(_ptr(image, _fp), _ints(image.shape),
                             _ptr(kernel, _fp), _ints(kernel.shape),
                             _ptr(out, _fp), _ints(offsets))
    return out


def convolve_host_reference(image, kernel, offsets, result=None):
    """The reference's own template, compiled from its source into _ref/."""
    image = _f32(image)
    kernel = _f32(kernel)
    out = np.zeros_like(image) if result is None else result
    _ref().ref_convolve(_ptr(image, _fp), _ints(image.shape),
                        _ptr(kernel, _fp), _ints(kernel.shape),
                        _ptr(out, _fp), _ints(offsets))
    return out


def fftshift_pad(kernel, image_shape):
    """src/convolution3Dfft.cu:128-166 with x-fastest axis roles (FP64 out)."""
    kernel = _f32(kernel)
    out = np.empty(tuple(image_shape), dtype=np.float64)
    rc = _lib().fco_fftshift_pad(_ptr(kernel, _fp), _ints(kernel.shape),
                                 _ptr(out, _dp), _ints(image_shape))
    if rc:
        raise ValueError("kernel larger than image")
    return out


def circular_conv_direct(image, kernel):
    """FP64 direct evaluation of convolution3DfftCUDAInPlace's result."""
    image = _f32(image)
    kernel = _f32(kernel)
    out = np.empty(image.shape, dtype=np.float64)
    rc = _lib().fco_circular_conv_direct(_ptr(image, _fp), _ints(image.shape),
                                         _ptr(kernel, _fp), _ints(kernel.shape),
                                         _ptr(out, _dp))
    if rc:
        raise ValueError("kernel larger than image")
    return out


BOUNDARY = {"periodic": 0, "zero": 1, "nearest": 2, "reflect": 3, "mirror": 4}


def boundary_conv_direct(image, kernel, mode, correlate=False):
    """FP64 direct linear convolution (or correlation) of the image extended
    by boundary rule `mode` (a BOUNDARY name or number); kernel centred at
    k/2, any kernel size."""
    image = _f32(image)
    kernel = _f32(kernel)
    m = BOUNDARY[mode] if isinstance(mode, str) else int(mode)
    out = np.empty(image.shape, dtype=np.float64)
    rc = _lib().fco_boundary_conv_direct(_ptr(image, _fp), _ints(image.shape),
                                         _ptr(kernel, _fp), _ints(kernel.shape), m,
                                         1 if correlate else 0, _ptr(out, _dp))
    if rc:
        raise ValueError("bad boundary mode or shape")
    return out


def fft1d(x, sign=-1):
    """FP64 unnormalised DFT of a 1D complex vector (any length)."""
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    _lib().fco_fft1d(a.view(np.float64).ctypes.data_as(_dp), a.size, 1, sign)
    return a


def fft3d(x, sign=-1):
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    _lib().fco_fft3d(a.view(np.float64).ctypes.data_as(_dp), _ints(a.shape), sign)
    return a


def fft_conv(image, kernel, conj_kernel=False):
    """FP64 FFT circular convolution (own C FFT); conj_kernel -> adjoint."""
    image = _f32(image)
    kernel = _f32(kernel)
    out = np.empty(image.shape, dtype=np.float64)
    rc = _lib().fco_fft_conv(_ptr(image, _fp), _ints(image.shape),
                             _ptr(kernel, _fp), _ints(kernel.shape),
                             _ptr(out, _dp), int(bool(conj_kernel)))
    if rc:
        raise ValueError("kernel larger than image")
    return out


def conv_grad_kernel(g, image, kernel_shape):
    """FP64 direct d(sum g*conv(im,k))/dk on the kernel support."""
    g = np.ascontiguousarray(g, dtype=np.float64)
    image = _f32(image)
    dk = np.empty(tuple(kernel_shape), dtype=np.float64)
    rc = _lib().fco_conv_grad_kernel(_ptr(g, _dp), _ptr(image, _fp),
                                     _ints(image.shape), _ints(kernel_shape),
                                     _ptr(dk, _dp))
    if rc:
        raise ValueError("kernel larger than image")
    return dk


def zero_padd(image, kernel_shape, factor=1):
    """tests/padd_utils.h:100-115,157-171: returns (padded, offsets)."""
    image = _f32(image)
    ext = (ctypes.c_int * 3)()
    off = (ctypes.c_int * 3)()
    _lib().fco_zero_padd_extents(_ints(image.shape), _ints(kernel_shape),
                                 int(factor), ext, off)
    padded = np.empty(tuple(ext), dtype=np.float32)
    _lib().fco_insert_at_offsets(_ptr(image, _fp), _ints(image.shape),
                                 _ptr(padded, _fp), ext, off)
    return padded, tuple(off)


def l2norm_ref(a, b):
    """tests/test_utils.hpp:76-88: sqrt(sum d^2)/N (the reference's metric)."""
    a = _f32(a)
    b = _f32(b)
    return _lib().fco_l2norm_ref(_ptr(a, _fp), _ptr(b, _fp), a.size)


def rl_multiview(views, kernels, iterations, eps=1e-6):
    """FP64 multiview Richardson-Lucy oracle (psi0 = mean of views)."""
    views = _f32(views)
    V = views.shape[0]
    shape = views.shape[1:]
    kshapes = []
    flat = []
    for k in kernels:
        k = _f32(k)
        kshapes += list(k.shape)
        flat.append(k.ravel())
    kflat = _f32(np.concatenate(flat))
    psi = np.empty(shape, dtype=np.float64)
    rc = _lib().fco_rl_multiview(_ptr(views, _fp), V, _ints(shape),
                                 _ptr(kflat, _fp), _ints(kshapes),
                                 int(iterations), float(eps), _ptr(psi, _dp))
    if rc:
        raise ValueError("kernel larger than image")
    return psi


# --------------------------------------------------------------------------
# metrics used by every parity test
# --------------------------------------------------------------------------
def rel_l2(a, ref):
    a = np.asarray(a, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    den = np.linalg.norm(ref.ravel())
    num = np.linalg.norm((a - ref).ravel())
    return float(num / den) if den > 0 else float(num)


def max_abs(a, ref):
    return float(np.max(np.abs(np.asarray(a, np.float64) - np.asarray(ref, np.float64))))


def fft_workers():
    """Threads for the FP64 FFTs below: the host cores this process may use,
    at most 16 (one GPU's share of the host)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def _rfftn(a):
    try:
        import scipy.fft as sf
        return sf.rfftn(a, workers=fft_workers())
    except ImportError:
        return np.fft.rfftn(a)


def _irfftn(A, shape):
    axes = tuple(range(len(shape)))
    try:
        import scipy.fft as sf
        return sf.irfftn(A, s=shape, axes=axes, workers=fft_workers())
    except ImportError:
        return np.fft.irfftn(A, s=shape, axes=axes)


def kernel_spectrum(kernel, image_shape, conj_kernel=False):
    """FP64 half spectrum of the kernel zero-padded and centred at k/2 on
    the image grid (src/convolution3Dfft.cu:128-166, x-fastest roles)."""
    K = _rfftn(fftshift_pad(kernel, image_shape))
    return np.conj(K) if conj_kernel else K


def numpy_fft_conv(image, kernel, conj_kernel=False, K=None):
    """FP64 FFT circular convolution via pocketfft (scipy.fft, threaded;
    numpy.fft when scipy is absent), for shapes where the C oracle's O(n log n)
    recursion is too slow; pinned against fft_conv() and the direct form on
    small shapes by the oracle tests.  ``K`` is a precomputed
    kernel_spectrum()."""
    image = np.asarray(image, dtype=np.float64)
    if K is None:
        K = kernel_spectrum(kernel, image.shape, conj_kernel)
    return _irfftn(_rfftn(image) * K, image.shape)


# numpy.pad modes of the boundary rules (half-sample symmetric is numpy's
# 'symmetric', whole-sample symmetric its 'reflect')
_NP_PAD = {1: "constant", 2: "edge", 3: "symmetric", 4: "reflect"}


def fft_boundary_conv(image, kernel, mode, correlate=False):
    """FP64 boundary-mode convolution by FFT: the image extended by rule
    ``mode`` (a BOUNDARY name or number other than periodic) far enough that
    no interior output wraps, circularly convolved on that grid, cropped.
    The FFT form of boundary_conv_direct() (pinned against it by the oracle
    tests) for sizes where the direct sum is too slow."""
    m = BOUNDARY[mode] if isinstance(mode, str) else int(mode)
    if m not in _NP_PAD:
        raise ValueError("periodic is numpy_fft_conv")
    image = np.asarray(image, dtype=np.float64)
    k = np.asarray(kernel, dtype=np.float64)
    if correlate:
        # correlation is convolution with the kernel flipped about its centre
        # k/2: tap j moves to 2(k/2) - j, which needs an odd extent per axis
        for d, n in enumerate(k.shape):
            if n % 2 == 0:
                k = np.concatenate([k, np.zeros_like(k.take([0], axis=d))], axis=d)
        k = k[::-1, ::-1, ::-1]
    pads = [(n_k - 1 - n_k // 2, n_k // 2) for n_k in k.shape]
    padded = np.pad(image, pads, mode=_NP_PAD[m])
    full = numpy_fft_conv(padded, k.astype(np.float32) if k.dtype != np.float32 else k)
    sl = tuple(slice(lo, lo + n) for (lo, _), n in zip(pads, image.shape))
    return full[sl]


def fft_conv_grad_kernel(g, image, kernel_shape):
    """FP64 gradient of sum(g * conv(image, k)) with respect to the kernel
    taps by FFT: u = IFFT(FFT(g) conj(FFT(image))), dk[j] = u[(j - k/2) mod n];
    pinned against the direct conv_grad_kernel() by the oracle tests."""
    g = np.asarray(g, dtype=np.float64)
    im = np.asarray(image, dtype=np.float64)
    u = _irfftn(_rfftn(g) * np.conj(_rfftn(im)), g.shape)
    idx = [(np.arange(k) - k // 2) % n for k, n in zip(kernel_shape, g.shape)]
    return u[np.ix_(*idx)]


def fft_rl_multiview(views, kernels, iterations, eps=1e-6):
    """FP64 multiview Richardson-Lucy by FFT (psi0 = mean of the views):
    the update of rl_multiview(), with each PSF's spectrum computed once;
    pinned against rl_multiview() by the oracle tests."""
    views = [np.asarray(v, dtype=np.float64) for v in views]
    shape = views[0].shape
    Ks = [kernel_spectrum(k, shape) for k in kernels]
    psi = np.mean(views, axis=0)
    for _ in range(int(iterations)):
        P = _rfftn(psi)
        acc = None
        for v, K in zip(views, Ks):
            blur = _irfftn(P * K, shape)
            R = _rfftn(v / np.maximum(blur, eps)) * np.conj(K)
            acc = R if acc is None else acc + R
        psi = psi * _irfftn(acc, shape) / len(views)
    return psi
