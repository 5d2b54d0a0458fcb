"""Host-pointer convolutions of volumes several times larger than the device
memory they may use (FCAMD_DEVICE_BUDGET_MB): the z-slab pipeline holds only
a ring of slab windows, the wrap planes and one slab's workspace, so the
volume never sits on the device whole (the reference's open blocking TODO,
src/convolution3Dfft.h:39,54,61).  Results match the FP64 oracle, in place
and through the returning legacy entry point; a budget too small for any
slab is refused with the image untouched."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FP = ctypes.POINTER(ctypes.c_float)
TOL = 1e-5


@pytest.mark.parametrize("shape,kshape,budget_mb", [
    ((800, 128, 128), (9, 7, 5), 10),    # 52 MB through 10 MiB
    ((1000, 96, 160), (15, 5, 5), 16),   # kz/2 = 7 wrap planes
    ((1000, 64, 128), (1, 3, 3), 6),     # kz = 1: no halo
    ((700, 80, 96), (6, 4, 3), 6),       # even kz
])
def test_volume_beyond_budget(fc, oracle, monkeypatch, shape, kshape, budget_mb):
    from fourierconvolutioncudalib_amd import synth
    from fourierconvolutioncudalib_amd._lib import ints
    print("beyond budget", shape, kshape, budget_mb, flush=True)
    monkeypatch.setenv("FCAMD_DEVICE_BUDGET_MB", str(budget_mb))
    lib = fc.lib()
    need = lib.fcamd_host_device_bytes(ints(shape[::-1]), ints(kshape[::-1]), 0)
    vol = 4 * int(np.prod(shape))
    assert 0 < need <= budget_mb << 20 and vol >= 3 * (budget_mb << 20), (need, vol)
    a = synth.uniform(shape, 300 + shape[0])
    k = synth.random_psf(kshape, 301)
    ref = oracle.numpy_fft_conv(a, k)
    im = np.ascontiguousarray(a.copy())
    fc.convolution3DfftCUDAInPlace(im.reshape(-1), list(shape[::-1]), k.reshape(-1),
                                   list(kshape[::-1]), 0)
    err = oracle.rel_l2(im, ref)
    assert err < TOL, err
    legacy = fc.convolution3DfftCUDA(a.reshape(-1), list(shape), k.reshape(-1), list(kshape), 0)
    assert oracle.rel_l2(legacy.reshape(shape), ref) < TOL


def test_budget_too_small_is_refused(fc, monkeypatch):
    from fourierconvolutioncudalib_amd import synth
    shape, kshape = (200, 256, 256), (9, 9, 9)
    monkeypatch.setenv("FCAMD_DEVICE_BUDGET_MB", "2")
    a = synth.uniform(shape, 5)
    k = synth.random_psf(kshape, 6)
    im = a.copy()
    with pytest.raises(fc.FCError) as err:
        fc.convolution3DfftCUDAInPlace(im.reshape(-1), list(shape[::-1]), k.reshape(-1),
                                       list(kshape[::-1]), 0)
    assert err.value.code == 3 and "budget" in str(err.value)
    assert np.array_equal(im, a)
