"""Transform lengths at the single-tile limits of the fused passes (x <= 10239,
y <= 10240, z <= 6826: one row, column or z column of the transform held in
160 KiB of LDS), and past them, where the long-axis path takes over (four-step
FFTs through HBM for an axis beyond one tile, the reference's unfused
structure otherwise: src/convolution3Dfft.cu:519-547), against the FP64
oracle.  A length past one tile with no split into two tiles (a prime above
10240) runs as Bluestein's chirp-z over a 7-smooth length that splits
(20580 = 140 * 147 for the primes here), as cuFFT plans any size.  Only a
length whose chirp-z length would pass two tiles (2L - 1 > 10240^2) is
rejected with the unsupported-size status before any device work, leaving
the image untouched; zero extents are rejected as invalid arguments.

The exact limits 10239 = 3 * 3413 and 6826 = 2 * 3413 would run a direct
3413-point DFT stage in one tile; a direct stage of a prime above 41 is
replaced by the long-axis path's chirp-z (an axis of 4099 points took 32.7 ms
on the direct stage against 0.12 ms at 4096), so they run there.  10206 =
2 * 3^6 * 7 and 10240 = 2^11 * 5 run butterflies only.
Tolerance: rel-L2 <= 1e-5 against the FP64 oracle, as for every other shape.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-5
INVALID_ARGUMENT = 1
UNSUPPORTED_SIZE = 2


def _run(fc, stack, kernel):
    im = np.ascontiguousarray(stack, dtype=np.float32).copy()
    k = np.ascontiguousarray(kernel, dtype=np.float32)
    fc.convolution3DfftCUDAInPlace(im.reshape(-1), list(im.shape[::-1]), k.reshape(-1),
                                   list(k.shape[::-1]), 0)
    return im


# (image C-order shape, kernel C-order shape)
CASES = [((2, 3, 10239), (1, 3, 7)),    # x at its limit (prime factor 3413: chirp-z)
         ((2, 3, 10206), (1, 3, 9)),    # x near its limit, butterflies only
         ((2, 10240, 3), (1, 9, 3)),    # y at its limit
         ((6826, 2, 3), (5, 1, 3)),     # z at its limit (prime factor 3413: chirp-z)
         ((6720, 2, 2), (7, 1, 1))]     # z near its limit, butterflies only


@pytest.mark.parametrize("shape,kshape", CASES, ids=lambda v: "x".join(map(str, v)))
def test_maximum_extents(fc, oracle, shape, kshape):
    from fourierconvolutioncudalib_amd import synth
    print("case", shape, kshape, flush=True)
    stack = synth.uniform(shape, 4100 + sum(shape))
    kernel = synth.random_psf(kshape, 4200 + sum(kshape))
    out = _run(fc, stack, kernel)
    ref = oracle.fft_conv(stack, kernel)
    assert oracle.rel_l2(out, ref) < TOL


# past the single-tile limits: the long-axis path
LONG_CASES = [((2, 3, 10240), (1, 3, 7)),     # x: 80 x 128 four-step
              ((2, 10241, 3), (1, 9, 3)),     # y: 77 x 133 four-step
              ((6827, 2, 3), (5, 1, 3)),      # z: one tile of the unfused path (prime)
              ((2, 3, 20000), (1, 3, 9)),     # x: 125 x 160
              ((8000, 2, 3), (7, 1, 3)),      # z: 80 x 100
              ((3, 12000, 2), (1, 15, 1)),    # y: 100 x 120
              ((7000, 1, 10500), (9, 1, 11)),  # x and z past their limits together
              ((2, 3, 10243), (1, 3, 7)),     # x: prime, chirp-z over 140 x 147
              ((2, 10247, 3), (1, 9, 3)),     # y: prime, chirp-z
              ((10253, 2, 3), (5, 1, 3)),     # z: prime, chirp-z
              ((10253, 2, 1), (5, 1, 1))]     # z: chirp-z over single-bin lines (nx = 1)


@pytest.mark.parametrize("shape,kshape", LONG_CASES, ids=lambda v: "x".join(map(str, v)))
def test_past_single_tile_limits(fc, oracle, shape, kshape):
    import torch
    from fourierconvolutioncudalib_amd import synth
    print("long case", shape, kshape, flush=True)
    stack = synth.uniform(shape, 4300 + sum(shape))
    kernel = synth.random_psf(kshape, 4400 + sum(kshape))
    ref = oracle.numpy_fft_conv(stack, kernel)
    out = _run(fc, stack, kernel)
    assert oracle.rel_l2(out, ref) < TOL
    # device entry point with a caller workspace, and the correlation
    ti = torch.from_numpy(stack).cuda()
    tk = torch.from_numpy(kernel).cuda()
    ws = torch.empty(fc.workspace_bytes(shape, kshape), dtype=torch.uint8, device="cuda")
    got = fc.convolve3d(ti, tk, op=fc.CORRELATE, workspace=ws).cpu().numpy()
    refc = oracle.numpy_fft_conv(stack, kernel, conj_kernel=True)
    assert oracle.rel_l2(got, refc) < TOL


# one-tile lengths whose plan would hold a direct stage of a prime above 41:
# the chirp-z of the long-axis path (M = 8232 in one tile for 4099; M = 20000
# = 125 x 160 for 9973 and 16464 for 8198, over a four-step)
PRIME_FACTOR_CASES = [((8, 64, 4099), (3, 5, 7)),   # x
                      ((8, 4099, 6), (3, 5, 5)),    # y
                      ((4099, 8, 6), (7, 5, 5)),    # z
                      ((3, 5, 9973), (1, 3, 5)),    # x, chirp-z over a four-step
                      ((6, 3, 8198), (3, 3, 5))]    # x = 2 * 4099


@pytest.mark.parametrize("shape,kshape", PRIME_FACTOR_CASES, ids=lambda v: "x".join(map(str, v)))
def test_large_prime_factor_takes_chirp_z(fc, oracle, shape, kshape):
    import torch
    from fourierconvolutioncudalib_amd import synth
    print("prime factor case", shape, kshape, flush=True)
    stack = synth.uniform(shape, 4500 + sum(shape))
    kernel = synth.random_psf(kshape, 4600 + sum(kshape))
    out = _run(fc, stack, kernel)
    assert oracle.rel_l2(out, oracle.numpy_fft_conv(stack, kernel)) < TOL
    ti = torch.from_numpy(stack).cuda()
    tk = torch.from_numpy(kernel).cuda()
    ws = torch.empty(fc.workspace_bytes(shape, kshape), dtype=torch.uint8, device="cuda")
    got = fc.convolve3d(ti, tk, op=fc.CORRELATE, workspace=ws).cpu().numpy()
    assert oracle.rel_l2(got, oracle.numpy_fft_conv(stack, kernel, conj_kernel=True)) < TOL


@pytest.mark.parametrize("shape", [(1, 1, 52428841)], ids=lambda v: "x".join(map(str, v)))
def test_prime_past_two_tiles_rejected(fc, shape):
    im = np.ones(shape, np.float32)
    k = np.ones((1, 1, 1), np.float32)
    with pytest.raises(fc.FCError) as err:
        fc.convolution3DfftCUDAInPlace(im.reshape(-1), list(shape[::-1]), k.reshape(-1),
                                       [1, 1, 1], 0)
    assert err.value.code == UNSUPPORTED_SIZE
    assert "no split" in str(err.value)
    assert np.all(im == 1)


@pytest.mark.parametrize("dims,kdims", [((0, 4, 4), (1, 1, 1)), ((4, 4, 4), (3, 0, 3))],
                         ids=["empty-image", "empty-kernel"])
def test_empty_extents_rejected(fc, dims, kdims):
    """A zero extent is an invalid argument: no device work, and the image
    (when it has voxels) stays untouched."""
    im = np.ones(dims, np.float32)
    k = np.ones(kdims, np.float32)
    with pytest.raises(fc.FCError) as err:
        fc.convolution3DfftCUDAInPlace(im.reshape(-1), list(dims[::-1]), k.reshape(-1),
                                       list(kdims[::-1]), 0)
    assert err.value.code == INVALID_ARGUMENT
    assert np.all(im == 1)
