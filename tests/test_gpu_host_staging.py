"""Host-pointer entry points on volumes large enough for pinned staging and
for the z-slab (overlap-save) pipeline: several slabs, a last slab of a
different length, the tail planes wrapping into the first slab, kz = 1 and
even kz.  Results agree with the device-resident path and the FP64 oracle."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FP = ctypes.POINTER(ctypes.c_float)
TOL = 1e-5


@pytest.mark.parametrize("shape,kshape", [((160, 256, 256), (15, 15, 15)),
                                          ((161, 257, 255), (9, 7, 5)),
                                          ((130, 200, 260), (1, 3, 5)),
                                          ((150, 210, 240), (6, 4, 3)),
                                          ((40, 64, 64), (5, 5, 5))])
def test_inplace_host_entry(fc, oracle, shape, kshape):
    import torch
    from fourierconvolutioncudalib_amd import synth
    from fourierconvolutioncudalib_amd._lib import ints
    print("host entry", shape, kshape, flush=True)
    a = synth.uniform(shape, 41 + shape[0])
    k = synth.random_psf(kshape, 43)
    dev = fc.convolve3d(torch.from_numpy(a).cuda(), torch.from_numpy(k).cuda()).cpu().numpy()
    im = np.ascontiguousarray(a.copy())
    kk = np.ascontiguousarray(k)
    fc.lib().convolution3DfftCUDAInPlace(im.ctypes.data_as(FP), ints(shape[::-1]),
                                         kk.ctypes.data_as(FP), ints(kshape[::-1]), 0)
    assert fc.lib().fcamd_last_error() == 0
    assert oracle.rel_l2(im, dev) < 1e-6
    assert oracle.rel_l2(im, oracle.numpy_fft_conv(a, k)) < TOL


def test_boundary_host_entry_large(fc, oracle):
    import torch
    from fourierconvolutioncudalib_amd import synth
    from fourierconvolutioncudalib_amd._lib import check, ints
    shape, kshape = (150, 200, 180), (11, 9, 13)
    a = synth.uniform(shape, 77)
    k = synth.random_psf(kshape, 78)
    dev = fc.convolve3d_boundary(torch.from_numpy(a).cuda(), torch.from_numpy(k).cuda(),
                                 "reflect").cpu().numpy()
    im = np.ascontiguousarray(a.copy())
    check(fc.lib().fcamd_convolve3d_boundary_host(
        im.ctypes.data_as(FP), ints(shape[::-1]), np.ascontiguousarray(k).ctypes.data_as(FP),
        ints(kshape[::-1]), fc.CONVOLVE, fc.BOUNDARY["reflect"], 0))
    assert np.array_equal(im, dev)


def test_concurrent_host_calls(fc, oracle):
    """Host-pointer calls from several threads on one device (pooled buffers,
    pinned staging and the slab streams are shared per device): every result
    is that of a call made alone."""
    import threading
    from fourierconvolutioncudalib_amd import synth
    from fourierconvolutioncudalib_amd._lib import ints
    cases = [((96, 192, 200), (9, 7, 5)), ((40, 48, 32), (5, 5, 5)), ((160, 128, 256), (15, 3, 7)),
             ((24, 20, 16), (3, 3, 3))]
    inputs = [(synth.uniform(s, 500 + i), synth.random_psf(k, 600 + i)) for i, (s, k) in enumerate(cases)]
    refs = []
    for a, k in inputs:
        im = np.ascontiguousarray(a.copy())
        fc.lib().convolution3DfftCUDAInPlace(im.ctypes.data_as(FP), ints(a.shape[::-1]),
                                             np.ascontiguousarray(k).ctypes.data_as(FP),
                                             ints(k.shape[::-1]), 0)
        refs.append(im)
    outs = [None] * (3 * len(inputs))
    errs = []

    def work(j):
        a, k = inputs[j % len(inputs)]
        im = np.ascontiguousarray(a.copy())
        fc.lib().convolution3DfftCUDAInPlace(im.ctypes.data_as(FP), ints(a.shape[::-1]),
                                             np.ascontiguousarray(k).ctypes.data_as(FP),
                                             ints(k.shape[::-1]), 0)
        if fc.lib().fcamd_last_error() != 0:
            errs.append(j)
        outs[j] = im

    ths = [threading.Thread(target=work, args=(j,)) for j in range(len(outs))]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    assert not errs
    for j, o in enumerate(outs):
        assert np.array_equal(o, refs[j % len(inputs)]), j
