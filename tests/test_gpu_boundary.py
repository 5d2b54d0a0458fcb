"""Boundary-mode convolution (virtually padded transform, cropped result)
against the FP64 direct oracle for every mode, the circular path for
'periodic', the reference's host convolution on the interior, and the
adjoint identity of the zero-boundary correlation."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-5
MODES = ["zero", "nearest", "reflect", "mirror"]

# (image C-order, kernel C-order): compile-time padded lengths, runtime
# (generic) padded lengths, a kernel larger than the image, a 2D block,
# even-sized kernels
SHAPES = [((40, 48, 32), (15, 15, 15)),
          ((24, 20, 16), (5, 7, 3)),
          ((13, 50, 36), (3, 5, 9)),
          ((3, 4, 5), (7, 9, 3)),
          ((1, 40, 48), (1, 9, 7)),
          ((20, 22, 18), (4, 6, 2))]


def _data(shape, kshape, seed):
    from fourierconvolutioncudalib_amd import synth
    return synth.uniform(shape, seed), synth.random_psf(kshape, seed + 1)


@pytest.mark.parametrize("shape,kshape", SHAPES)
@pytest.mark.parametrize("mode", MODES)
def test_boundary_matches_direct_oracle(fc, oracle, shape, kshape, mode):
    import torch
    a, k = _data(shape, kshape, 700 + sum(shape))
    print("boundary", mode, shape, kshape, flush=True)
    got = fc.convolve3d_boundary(torch.from_numpy(a).cuda(), torch.from_numpy(k).cuda(),
                                 mode).cpu().numpy()
    ref = oracle.boundary_conv_direct(a, k, mode)
    assert oracle.rel_l2(got, ref) < TOL, (mode, shape, kshape)


@pytest.mark.parametrize("mode", ["zero", "reflect"])
def test_boundary_correlate(fc, oracle, mode):
    import torch
    a, k = _data((20, 24, 30), (5, 7, 9), 811)
    got = fc.convolve3d_boundary(torch.from_numpy(a).cuda(), torch.from_numpy(k).cuda(), mode,
                                 op=fc.CORRELATE).cpu().numpy()
    ref = oracle.boundary_conv_direct(a, k, mode, correlate=True)
    assert oracle.rel_l2(got, ref) < TOL


def test_zero_boundary_adjoint_identity(fc):
    """<conv_zero(x, k), g> = <x, corr_zero(g, k)> in FP64 of the FP32 results."""
    import torch
    x, k = _data((18, 26, 22), (7, 5, 9), 901)
    g, _ = _data((18, 26, 22), (1, 1, 1), 903)
    tk = torch.from_numpy(k).cuda()
    y = fc.convolve3d_boundary(torch.from_numpy(x).cuda(), tk, "zero").cpu().numpy()
    xt = fc.convolve3d_boundary(torch.from_numpy(g).cuda(), tk, "zero",
                                op=fc.CORRELATE).cpu().numpy()
    lhs = float(np.dot(y.ravel().astype(np.float64), g.ravel().astype(np.float64)))
    rhs = float(np.dot(x.ravel().astype(np.float64), xt.ravel().astype(np.float64)))
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)


def test_periodic_is_the_circular_convolution(fc, oracle):
    import torch
    a, k = _data((24, 20, 16), (5, 7, 3), 950)
    ta, tk = torch.from_numpy(a).cuda(), torch.from_numpy(k).cuda()
    b = fc.convolve3d_boundary(ta, tk, "periodic").cpu().numpy()
    c = fc.convolve3d(ta, tk).cpu().numpy()
    assert np.array_equal(b, c)
    assert fc.padded_extents(a.shape, k.shape, "periodic") == a.shape


def test_zero_boundary_interior_is_the_reference_host_convolution(fc, oracle):
    """tests/test_algorithms.hpp:9-58 computes the interior [k/2, n - k/2)
    without touching a boundary; every mode must reproduce it there."""
    import torch
    a, k = _data((30, 34, 28), (7, 5, 9), 977)
    hk = [d // 2 for d in k.shape]
    ref = oracle.convolve_host(a, k, hk)  # reference restatement, interior only
    sl = tuple(slice(h, n - h) for h, n in zip(hk, a.shape))
    for mode in MODES:
        got = fc.convolve3d_boundary(torch.from_numpy(a).cuda(), torch.from_numpy(k).cuda(),
                                     mode).cpu().numpy()
        assert oracle.rel_l2(got[sl], ref[sl]) < TOL, mode


def test_boundary_host_entry_and_workspace(fc, oracle):
    import torch
    from fourierconvolutioncudalib_amd._lib import check, ints
    a, k = _data((26, 30, 24), (9, 5, 7), 990)
    ref = oracle.boundary_conv_direct(a, k, "mirror")
    im = np.ascontiguousarray(a.copy())
    fp = ctypes.POINTER(ctypes.c_float)
    check(fc.lib().fcamd_convolve3d_boundary_host(
        im.ctypes.data_as(fp), ints(a.shape[::-1]), np.ascontiguousarray(k).ctypes.data_as(fp),
        ints(k.shape[::-1]), fc.CONVOLVE, fc.BOUNDARY["mirror"], 0))
    assert oracle.rel_l2(im, ref) < TOL
    nbytes = fc.boundary_workspace_bytes(a.shape, k.shape, "mirror")
    ws = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    ta = torch.from_numpy(a).cuda()
    got = fc.convolve3d_boundary(ta, torch.from_numpy(k).cuda(), "mirror", out=ta, workspace=ws)
    assert oracle.rel_l2(got.cpu().numpy(), ref) < TOL
    small = torch.empty(nbytes - 256, dtype=torch.uint8, device="cuda")
    tk = torch.from_numpy(k).cuda()
    # the Python wrapper refuses a short workspace before calling the library ...
    with pytest.raises(ValueError, match="workspace"):
        fc.convolve3d_boundary(ta, tk, "mirror", workspace=small)
    # ... and so does the C entry point itself
    rc = fc.lib().fcamd_convolve3d_boundary_device(
        ta.data_ptr(), ta.data_ptr(), ints(a.shape[::-1]), tk.data_ptr(), ints(k.shape[::-1]),
        fc.CONVOLVE, fc.BOUNDARY["mirror"], small.data_ptr(), nbytes - 256, None)
    assert rc == 1 and b"workspace" in fc.lib().fcamd_last_error_message()
