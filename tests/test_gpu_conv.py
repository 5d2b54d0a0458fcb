"""GPU parity of the drop-in path (convolution3DfftCUDAInPlace and friends)
against the oracle and the reference's golden vectors.

Tolerances: rel-L2 = ||gpu - ref||_2 / ||ref||_2 <= 1e-5 against the FP64
oracle and against the reference's host convolution (FP32 arithmetic
throughout the GPU path); the reference's own metric sqrt(sum d^2)/N is also
restated where its tests use it.
"""
import numpy as np
import pytest

from conftest import golden_cases, load_golden

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _conv_inplace(fc, padded, kernel):
    """Calls the reference ABI exactly as tests/test_gpu_convolve.cpp does:
    C-order stacks, dims passed reversed ({x, y, z})."""
    im = np.ascontiguousarray(padded, dtype=np.float32).copy()
    k = np.ascontiguousarray(kernel, dtype=np.float32)
    fc.convolution3DfftCUDAInPlace(im.reshape(-1), list(im.shape[::-1]), k.reshape(-1),
                                   list(k.shape[::-1]), fc.selectDeviceWithHighestComputeCapability())
    return im


def _crop(a, off, shape):
    return a[tuple(slice(o, o + s) for o, s in zip(off, shape))]


# ---------------------------------------------------------------------------
# device queries (src/standardCUDAfunctions.cu)
# ---------------------------------------------------------------------------
def test_device_queries(fc):
    n = fc.getNumDevicesCUDA()
    assert n >= 1
    dev = fc.selectDeviceWithHighestComputeCapability()
    assert 0 <= dev < n
    assert fc.getCUDAcomputeCapabilityMajorVersion(dev) >= 9
    assert fc.getCUDAcomputeCapabilityMinorVersion(dev) >= 0
    assert fc.getNameDeviceCUDA(dev)
    assert fc.getMemDeviceCUDA(dev) > (1 << 30)


# ---------------------------------------------------------------------------
# golden vectors from the reference's host convolution
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("name", golden_cases())
def test_golden(fc, oracle, name):
    g = load_golden(name)
    padded, off = oracle.zero_padd(g["image"], g["kernel"].shape, int(g["factor"]))
    out = _crop(_conv_inplace(fc, padded, g["kernel"]), off, g["image"].shape)
    if not g["expected"].any():
        assert np.max(np.abs(out)) < 1e-6 * max(1.0, float(np.abs(g["image"]).max()))
        return
    assert oracle.rel_l2(out, g["expected"]) < TOL
    f64 = _crop(oracle.fft_conv(padded, g["kernel"]), off, g["image"].shape)
    assert oracle.rel_l2(out, f64) < TOL
    # tests/test_gpu_convolve.cpp:57-89 check: sums agree to 1e-7 relative
    # (BOOST_CHECK_CLOSE(., ., 1e-5 percent)); FP64 sums of both
    if name.startswith("fixture_"):
        s_ref = float(np.sum(g["expected"], dtype=np.float64))
        assert abs(float(np.sum(out, dtype=np.float64)) - s_ref) <= 1e-6 * abs(s_ref)


# ---------------------------------------------------------------------------
# tests/test_gpu_convolve.cpp:204-467 (asymmetric / prime shapes)
# ---------------------------------------------------------------------------
def _prime_stack():
    return (19, 17, 13)  # z, y, x


@pytest.mark.parametrize("kind", ["identity", "horizontal", "vertical", "diagonal"])
def test_prime_shapes(fc, oracle, kind):
    s = _prime_stack()
    k = np.zeros((3, 3, 3), np.float32)
    if kind == "identity":
        stack = np.arange(np.prod(s), dtype=np.float32).reshape(s)
        k[1, 1, 1] = 1
    else:
        stack = np.ones(s, np.float32)
        if kind == "horizontal":
            k[1, 1, 0:3] = 1
        elif kind == "vertical":
            k[1, 0:3, 1] = 1
        else:
            for i in range(3):
                k[i, i, i] = 1
    padded, off = oracle.zero_padd(stack, k.shape)
    assert padded.shape == (21, 19, 15)
    out = _crop(_conv_inplace(fc, padded, k), off, s)
    ref = _crop(oracle.circular_conv_direct(padded, k), off, s)
    assert oracle.rel_l2(out, ref) < TOL
    if kind == "identity":
        # reference threshold l2norm < 1e-4 (:262) holds a fortiori
        assert oracle.l2norm_ref(stack, out) < 1e-4


# ---------------------------------------------------------------------------
# FFT round trips (tests/test_how_cufft_works.cu): a centred delta kernel
# makes the whole R2C -> C2R pipeline an identity
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("shape", [(13, 17, 19), (17, 17, 17), (32, 16, 8), (9, 27, 3),
                                   (5, 25, 125), (7, 49, 343), (7, 14, 343), (1, 1, 1),
                                   (1, 5, 7), (4, 1, 6), (3, 2, 1)])
def test_round_trip_identity(fc, oracle, shape):
    stack = np.arange(np.prod(shape), dtype=np.float32).reshape(shape)
    k = np.ones((1, 1, 1), np.float32)
    out = _conv_inplace(fc, stack, k)
    assert oracle.rel_l2(out, stack) < 1e-6


# ---------------------------------------------------------------------------
# tests/test_gpu_numerical_stability.cpp
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("case", ["times_two_128", "ramp_tiny_kernel_times_two",
                                  "ramp_tiny_kernel_padd_by_10fold", "ramp_normalized_16",
                                  "ramp_normalized_128", "ramp_normalized_256"])
def test_numerical_stability(fc, oracle, case):
    factor = 1
    if case == "times_two_128":
        stack = np.full((128, 128, 128), 42, np.float32)
        kshape, amp = (91, 31, 31), 2.0
    elif case.startswith("ramp_tiny"):
        stack = np.arange(128 ** 3, dtype=np.float32).reshape(128, 128, 128)
        kshape, amp = (5, 3, 3), 2.0
        factor = 10 if case.endswith("10fold") else 1
    else:
        n = int(case.rsplit("_", 1)[1])
        stack = (np.arange(n ** 3, dtype=np.float32) * np.float32(1.0 / n ** 3)).reshape(n, n, n)
        kshape, amp = (91, 31, 31), 1.0
    k = np.zeros(kshape, np.float32)
    k[kshape[0] // 2, kshape[1] // 2, kshape[2] // 2] = amp
    padded, off = oracle.zero_padd(stack, kshape, factor)
    out = _crop(_conv_inplace(fc, padded, k), off, stack.shape)
    expected = stack.astype(np.float64) * amp  # a scaled centred delta
    assert oracle.rel_l2(out, expected) < TOL
    assert oracle.l2norm_ref(out, expected.astype(np.float32)) < 1e-3


# ---------------------------------------------------------------------------
# workloads
# ---------------------------------------------------------------------------
def test_plain_128x128x64_k15(fc, oracle):
    """128x128x64 uniform image, 15^3 Gaussian PSF: GPU vs FP64 oracle, and
    vs the reference's host convolution on the zero-padded interior."""
    from fourierconvolutioncudalib_amd import synth
    img, psf = synth.workload(synth.CFG1)
    out = _conv_inplace(fc, img, psf)
    assert oracle.rel_l2(out, oracle.fft_conv(img, psf)) < TOL
    padded, off = oracle.zero_padd(img, psf.shape)
    lin = _crop(_conv_inplace(fc, padded, psf), off, img.shape)
    host = _crop(oracle.convolve_host(padded, psf, off), off, img.shape)
    assert oracle.rel_l2(lin, host) < TOL


def test_correlate_op_is_adjoint(fc, oracle):
    rng = np.random.default_rng(3)
    img = rng.random((20, 24, 30), dtype=np.float32)
    k = rng.random((5, 7, 3), dtype=np.float32)
    out = fc.convolve(img, k, op=fc.CORRELATE)
    assert oracle.rel_l2(out, oracle.fft_conv(img, k, conj_kernel=True)) < TOL


def test_legacy_entry_points(fc, oracle):
    """convolution3DfftCUDA (dims z-fastest, returns new buffer) and
    convolution3DfftCUDA_test (image-sized pre-shifted kernel)."""
    rng = np.random.default_rng(4)
    img = rng.random((12, 10, 14), dtype=np.float32)
    k = rng.random((3, 5, 7), dtype=np.float32)
    img0 = img.copy()
    ref = oracle.fft_conv(img, k)
    # the legacy convention names the contiguous axis last: dims = C-order shape
    out = fc.convolution3DfftCUDA(img.reshape(-1), list(img.shape), k.reshape(-1), list(k.shape), 0)
    assert oracle.rel_l2(out.reshape(img.shape), ref) < TOL
    assert np.array_equal(img, img0)
    full = oracle.fftshift_pad(k, img.shape).astype(np.float32)
    out = fc.convolution3DfftCUDA_test(img.reshape(-1), list(img.shape), full.reshape(-1), 0)
    assert oracle.rel_l2(out.reshape(img.shape), ref) < TOL


def test_errors_leave_image_untouched(fc):
    im = np.ones((8, 8, 8), np.float32)
    k = np.ones((3, 3, 9), np.float32)
    with pytest.raises(fc.FCError):
        fc.convolution3DfftCUDAInPlace(im.reshape(-1), [8, 8, 8], k.reshape(-1), [9, 3, 3], 0)
    with pytest.raises(fc.FCError):
        fc.convolution3DfftCUDAInPlace(im.reshape(-1), [8, 8, 8], k.reshape(-1)[:27], [3, 3, 3], 99)
    assert np.all(im == 1)


def test_device_api_matches_host_api(fc, oracle):
    import torch
    from fourierconvolutioncudalib_amd import synth
    img = synth.uniform((24, 40, 36), 5)
    k = synth.gaussian_psf((5, 7, 9), (1.0, 1.5, 2.0))
    host = fc.convolve(img, k)
    ti = torch.from_numpy(img).cuda()
    tk = torch.from_numpy(k).cuda()
    out = fc.convolve3d(ti, tk)
    ws = torch.empty(fc.workspace_bytes(img.shape, k.shape), dtype=torch.uint8, device="cuda")
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        out2 = fc.convolve3d(ti, tk, workspace=ws, stream=s)
    s.synchronize()
    inplace = ti.clone()
    fc.convolve3d(inplace, tk, out=inplace)
    torch.cuda.synchronize()
    for o in (out, out2, inplace):
        assert oracle.rel_l2(o.cpu().numpy(), host) < 1e-6
    assert oracle.rel_l2(host, oracle.fft_conv(img, k)) < TOL


def test_metric_shape_512x512x256_k15(fc, oracle):
    import torch
    from fourierconvolutioncudalib_amd import synth
    img, psf = synth.workload(synth.METRIC)
    out = fc.convolve3d(torch.from_numpy(img).cuda(), torch.from_numpy(psf).cuda()).cpu().numpy()
    ref = oracle.numpy_fft_conv(img, psf)
    assert oracle.rel_l2(out, ref) < TOL
    assert oracle.max_abs(out, ref) < 1e-4


def test_lightsheet_600x500x240_k31x31x61(fc, oracle):
    import torch
    from fourierconvolutioncudalib_amd import synth
    img, psf = synth.workload(synth.CFG2)
    out = fc.convolve3d(torch.from_numpy(img).cuda(), torch.from_numpy(psf).cuda()).cpu().numpy()
    ref = oracle.numpy_fft_conv(img, psf)
    assert oracle.rel_l2(out, ref) < TOL


def test_metric_shape_properties(fc, oracle):
    """Size-independent properties at 512x512x256 (x,y,z) with the 15^3 PSF,
    on the GPU alone: linearity, circular-shift equivariance (a roll of the
    input rolls the output), and a centred delta kernel as the identity."""
    import torch
    from fourierconvolutioncudalib_amd import synth
    img, psf = synth.workload(synth.METRIC)
    a = torch.from_numpy(img).cuda()
    b = torch.from_numpy(synth.uniform(img.shape, 77)).cuda()
    k = torch.from_numpy(psf).cuda()
    ca, cb = fc.convolve3d(a, k), fc.convolve3d(b, k)
    lin = fc.convolve3d(2.0 * a - 0.5 * b, k)
    assert oracle.rel_l2(lin.cpu().numpy(), (2.0 * ca - 0.5 * cb).cpu().numpy()) < 1e-5
    shift = (3, -5, 7)
    rolled = fc.convolve3d(torch.roll(a, shift, dims=(0, 1, 2)), k)
    assert oracle.rel_l2(rolled.cpu().numpy(),
                         torch.roll(ca, shift, dims=(0, 1, 2)).cpu().numpy()) < 1e-5
    delta = torch.zeros_like(k)
    delta[7, 7, 7] = 1.0
    ident = fc.convolve3d(a, delta)
    assert oracle.rel_l2(ident.cpu().numpy(), img) < 2e-6  # FP32 forward + inverse round trip
