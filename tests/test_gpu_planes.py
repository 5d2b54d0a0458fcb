"""Whole-plane x/y kernels (plane_kernels.hpp): for plane extents with
compile-time plane kernels the forward x-R2C and y-FFT run as one launch per
z-plane (plus one per kernel plane) and the inverse y-IFFT and x-C2R as
one, around the unchanged fused z pass (plane extents up to 256 x 64 and
64 x 256 whose row pairs fit one exchange group, and wide planes up to
256 x 256 whose even or odd half fits).  Volumes with few planes
take the split kernels (two workgroups per plane: even/odd x bins forward,
even/odd rows inverse); the cases cover split and whole-plane kernels in
either direction.  Checked against the FP64 oracle at
rel-L2 <= 1e-5 (the bar of every convolution here) for every instantiated
plane extent (each in the split or wide regime and, where it has them, on
the one-workgroup kernels), convolution and correlation, kernels from one
tap to the full plane, off-centre and even extents, through the drop-in
host entry point and the device entry point; the per-pass event timing confirms that no
separate y pass was launched.  Semantics: src/convolution3Dfft.cu:399-578
(circular convolution, kernel centred at floor(k/2)).
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-5

# (image C-order shape, kernel C-order shape)
CASES = [((64, 128, 128), (15, 15, 15)),   # the 128 x 128 x 64 workload
         ((5, 128, 128), (3, 128, 128)),    # kernel as wide as the plane
         ((33, 64, 64), (1, 1, 1)),         # one tap
         ((17, 96, 96), (4, 6, 9)),         # even kernel extents (centre floor(k/2))
         ((12, 64, 128), (7, 13, 5)),
         ((20, 128, 64), (9, 3, 11)),
         ((1, 128, 128), (1, 7, 5)),        # one plane
         ((12, 120, 80), (5, 9, 11)),       # radix-5 and radix-15 row and column plans
         ((9, 96, 112), (3, 7, 9)),
         ((7, 64, 120), (7, 3, 21)),
         ((30, 112, 64), (13, 11, 3)),
         ((20, 256, 64), (9, 31, 5)),       # larger planes (one row-pair group)
         ((11, 64, 256), (5, 7, 33)),
         ((17, 96, 160), (7, 9, 15)),
         ((13, 112, 144), (13, 5, 9)),
         # 128 x 128 planes under the split rule (256 CUs): forward and
         # inverse one workgroup per plane; forward whole, inverse split
         ((200, 128, 128), (5, 5, 5)),
         ((100, 128, 128), (31, 5, 5)),
         # wide planes (half a half-spectrum plane per workgroup)
         ((24, 256, 256), (9, 21, 7)),
         ((3, 256, 256), (3, 256, 256)),     # kernel as wide as the plane
         ((13, 192, 192), (5, 7, 9)),        # two row-pair groups forward
         ((11, 192, 256), (7, 5, 3)),        # two column groups inverse
         ((9, 160, 160), (1, 1, 1)),
         ((17, 128, 256), (4, 6, 8)),
         ((10, 256, 128), (3, 9, 5)),
         ((7, 144, 144), (5, 5, 5)),
         ((12, 240, 240), (11, 3, 7))]


def _ids(v):
    return "x".join(map(str, v))


@pytest.mark.parametrize("shape,kshape", CASES, ids=_ids)
def test_plane_extents(fc, oracle, shape, kshape):
    from fourierconvolutioncudalib_amd import synth
    print("plane case", shape, kshape, flush=True)
    stack = synth.uniform(shape, 7100 + sum(shape))
    kernel = synth.random_psf(kshape, 7200 + sum(kshape))
    im = np.ascontiguousarray(stack, dtype=np.float32).copy()
    k = np.ascontiguousarray(kernel, dtype=np.float32)
    fc.convolution3DfftCUDAInPlace(im.reshape(-1), list(im.shape[::-1]), k.reshape(-1),
                                   list(k.shape[::-1]), 0)
    ref = oracle.numpy_fft_conv(stack, kernel)
    assert oracle.rel_l2(im, ref) < TOL
    import torch
    ti = torch.from_numpy(stack).cuda()
    tk = torch.from_numpy(kernel).cuda()
    got = fc.convolve3d(ti, tk, op=fc.CORRELATE).cpu().numpy()
    ref_c = oracle.numpy_fft_conv(stack, kernel, conj_kernel=True)
    assert oracle.rel_l2(got, ref_c) < TOL


# Every plane extent with plane kernels (plane_kernels.hip,
# plane_kernels_wide.hip): the 69 with one-workgroup kernels and the 20 wide
# ones, as (nx, ny).
REGULAR_PLANES = ([(nx, ny) for nx in (64, 80, 96, 112, 120, 128) for ny in (64, 80, 96, 112, 120, 128)]
                  + [(64, 144), (64, 160), (64, 192), (64, 224), (64, 240), (64, 256), (80, 144),
                     (80, 160), (80, 192), (80, 224), (96, 144), (96, 160), (112, 144), (112, 160),
                     (120, 144), (128, 144), (144, 64), (144, 80), (144, 96), (144, 112), (144, 120),
                     (144, 128), (160, 64), (160, 80), (160, 96), (160, 112), (192, 64), (192, 80),
                     (192, 96), (224, 64), (224, 80), (240, 64), (256, 64)])
WIDE_PLANES = [(128, 160), (160, 128), (144, 144), (160, 160), (128, 192), (192, 128), (160, 192),
               (192, 160), (192, 192), (224, 224), (240, 240), (256, 256), (128, 256), (256, 128),
               (192, 256), (256, 192), (224, 256), (256, 224), (240, 256), (256, 240)]


def _check_planes(fc, oracle, planes, nz, kshape, op, seed):
    import torch
    from fourierconvolutioncudalib_amd import synth
    worst = 0.0
    for nx, ny in planes:
        shape = (nz, ny, nx)
        stack = synth.uniform(shape, seed + nx + 7 * ny)
        kernel = synth.random_psf(kshape, seed + nx - ny)
        got = fc.convolve3d(torch.from_numpy(stack).cuda(), torch.from_numpy(kernel).cuda(),
                            op=op).cpu().numpy()
        ref = oracle.numpy_fft_conv(stack, kernel, conj_kernel=op == fc.CORRELATE)
        err = oracle.rel_l2(got, ref)
        assert err < TOL, (shape, kshape, op, err)
        worst = max(worst, err)
    return worst


def test_every_plane_extent_few_planes(fc, oracle):
    """Every registered plane extent at 6 planes and a 3 x 5 x 7 kernel: on
    256 CUs the split kernels run in both directions where they exist, the
    wide split kernels for the wide extents, the one-workgroup kernels
    elsewhere; convolution against the FP64 oracle."""
    worst = _check_planes(fc, oracle, REGULAR_PLANES + WIDE_PLANES, 6, (3, 5, 7), fc.CONVOLVE, 500)
    print(f"{len(REGULAR_PLANES) + len(WIDE_PLANES)} extents, worst rel-L2 {worst:.2e}", flush=True)


def test_every_regular_plane_extent_many_planes(fc, oracle):
    """Every extent with one-workgroup kernels at 132 planes and a 1 x 3 x 3
    kernel: on 256 CUs 2 * (nz + kz) and 2 * nz exceed the CU count, so the
    one-workgroup kernels run in both directions; correlation against the
    FP64 oracle."""
    worst = _check_planes(fc, oracle, REGULAR_PLANES, 132, (1, 3, 3), fc.CORRELATE, 900)
    print(f"{len(REGULAR_PLANES)} extents, worst rel-L2 {worst:.2e}", flush=True)


def test_plane_path_runs(fc, oracle):
    """One launch forward and one inverse around the z pass: the per-pass
    timing records no y pass for a plane extent, and records it for a shape
    without plane kernels."""
    import torch
    from fourierconvolutioncudalib_amd import synth
    lib = fc.lib()

    def passes(shape, kshape):
        img = torch.from_numpy(synth.uniform(shape, 1)).cuda()
        k = torch.from_numpy(synth.random_psf(kshape, 2)).cuda()
        fc.convolve3d(img, k)
        torch.cuda.synchronize()
        lib.fcamd_pass_timing(1)
        fc.convolve3d(img, k)
        torch.cuda.synchronize()
        lib.fcamd_pass_timing(0)
        ms = (ctypes.c_double * 16)()
        cnt = (ctypes.c_int * 16)()
        n = lib.fcamd_pass_times(ms, cnt, 16)
        return {lib.fcamd_pass_name(i).decode(): cnt[i] for i in range(n)}

    for shape in ((16, 128, 128), (16, 256, 256)):
        plane = passes(shape, (5, 5, 5))
        print("plane path launches", shape, plane, flush=True)
        assert plane["image_xfwd"] == 1 and plane["image_xinv"] == 1 and plane["zfused"] == 1
        assert plane["image_yfwd"] == 0 and plane["image_yinv"] == 0
    other = passes((16, 320, 320), (5, 5, 5))
    assert other["image_yfwd"] == 1 and other["image_yinv"] == 1
