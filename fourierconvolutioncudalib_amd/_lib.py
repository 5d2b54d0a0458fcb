"""Loading of libFourierConvolutionCUDALib.so (the MI355X build) via ctypes.

The library is built in-tree by ``make -C fourierconvolutioncudalib_amd/csrc``
(``__graft_entry__.build()`` does it).  Loading fails loudly when it is
missing: there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libFourierConvolutionCUDALib.so")
CSRC = os.path.join(_HERE, "csrc")

_lib = None

_i = ctypes.c_int
_ip = ctypes.POINTER(ctypes.c_int)
_fp = ctypes.POINTER(ctypes.c_float)
_vp = ctypes.c_void_p
_ll = ctypes.c_longlong

# name -> (restype, argtypes); every symbol include/*.h declares
SIGNATURES = {
    # include/convolution3Dfft.h (reference ABI)
    "convolution3DfftCUDA_test": (_fp, [_fp, _ip, _fp, _i]),
    "convolution3DfftCUDA": (_fp, [_fp, _ip, _fp, _ip, _i]),
    "convolution3DfftCUDAInPlace": (None, [_fp, _ip, _fp, _ip, _i]),
    "selectDeviceWithHighestComputeCapability": (_i, []),
    "getCUDAcomputeCapabilityMinorVersion": (_i, [_i]),
    "getCUDAcomputeCapabilityMajorVersion": (_i, [_i]),
    "getNumDevicesCUDA": (_i, []),
    "getNameDeviceCUDA": (None, [_i, ctypes.c_char_p]),
    "getMemDeviceCUDA": (_ll, [_i]),
    "cuda_version": (_i, []),
    "gpu_mem_needed_mb": (_i, [_ip, _i]),
    # include/fourierconvolution_amd.h (additive ABI)
    "fcamd_last_error": (_i, []),
    "fcamd_last_error_message": (ctypes.c_char_p, []),
    "fcamd_free_result": (None, [_fp]),
    "fcamd_convolve3d_host": (_i, [_fp, _ip, _fp, _ip, _i, _i]),
    "fcamd_workspace_bytes": (_ll, [_ip, _ip]),
    "fcamd_host_device_bytes": (_ll, [_ip, _ip, _i]),
    "fcamd_convolve3d_device": (_i, [_vp, _vp, _ip, _vp, _ip, _i, _vp, _ll, _vp]),
    "fcamd_convolve3d_device_noalloc": (_i, [_vp, _vp, _ip, _vp, _ip, _i, _vp]),
    "fcamd_radix_plan": (_i, [_i, _ip, _i]),
    "fcamd_grad_kernel_workspace_bytes": (_ll, [_ip, _ip]),
    "fcamd_conv_grad_kernel_device": (_i, [_vp, _vp, _vp, _ip, _ip, _vp, _ll, _vp]),
    "fcamd_conv_plan_create": (_vp, [_ip, _ip, _vp, _i, _vp, _ll, _vp]),
    "fcamd_conv_plan_set_kernel": (_i, [_vp, _vp, _vp]),
    "fcamd_conv_plan_execute": (_i, [_vp, _vp, _vp, _vp]),
    "fcamd_conv_plan_destroy": (None, [_vp]),
    "fcamd_boundary_workspace_bytes": (_ll, [_ip, _ip, _i]),
    "fcamd_padded_extents": (_i, [_ip, _ip, _i, _ip]),
    "fcamd_convolve3d_boundary_device": (_i, [_vp, _vp, _ip, _vp, _ip, _i, _i, _vp, _ll, _vp]),
    "fcamd_convolve3d_boundary_host": (_i, [_fp, _ip, _fp, _ip, _i, _i, _i]),
    "fcamd_has_static_kernel": (_i, [_i, _i]),
    "fcamd_static_lengths": (_i, [_ip, _i]),
    "fcamd_use_static_kernels": (None, [_i]),
    "fcamd_static_variant_count": (_i, [_i, _i]),
    "fcamd_static_variant_name": (ctypes.c_char_p, [_i, _i, _i]),
    "fcamd_select_static_variant": (_i, [_i, _i, _i]),
    "fcamd_selected_static_variant": (_i, [_i, _i]),
    "fcamd_batch_workspace_bytes": (_ll, [_i, _ip, _ip]),
    "fcamd_batch_create": (_vp, [_i, _ip, _ip, _vp, _vp, _vp, _i, _vp, _ll]),
    "fcamd_batch_execute": (_i, [_vp, _vp]),
    "fcamd_batch_graph_launch": (_i, [_vp, _vp]),
    "fcamd_batch_destroy": (None, [_vp]),
    "fcamd_rl_workspace_bytes": (_ll, [_i, _ip, _ip, _i]),
    "fcamd_rl_fused_available": (_i, [_i, _ip, _ip]),
    "fcamd_rl_multiview_device": (_i, [_vp, _vp, _i, _ip, _vp, _ip, _i, ctypes.c_float, _i, _i,
                                       _vp, _ll, _vp]),
    "fcamd_autotune": (_i, [_ip, _ip, _i, _vp]),
    "fcamd_autotune_load": (_i, [ctypes.c_char_p]),
    "fcamd_pass_timing": (None, [_i]),
    "fcamd_pass_times": (_i, [ctypes.POINTER(ctypes.c_double), _ip, _i]),
    "fcamd_pass_name": (ctypes.c_char_p, [_i]),
    "fcamd_modulate_device": (_i, [_vp, _vp, _ll, ctypes.c_float, _i, _vp]),
    "fcamd_stream_copy_device": (_i, [_vp, _vp, _ll, _vp]),
}


class FCError(RuntimeError):
    """A library call failed; ``code`` is the fcamd status code."""

    def __init__(self, code: int, message: str):
        super().__init__(f"[fcamd error {code}] {message}")
        self.code = code


def lib():
    """The loaded library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with "
                f"`make -C {CSRC}` (or __graft_entry__.build())")
        # torch ships its own libamdhip64.so.7 (same soname as the system
        # one).  Two HIP/HSA runtimes cannot share a process, and the first
        # one loaded wins the soname, so load torch's before ours when torch
        # is installed; the library then runs on the runtime torch uses.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        handle = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(rc: int) -> None:
    if rc != 0:
        msg = lib().fcamd_last_error_message()
        raise FCError(rc, msg.decode() if msg else "unknown error")


def ints(values):
    values = [int(v) for v in values]
    return (ctypes.c_int * len(values))(*values)
