"""Pins the CPU oracle (oracle/) before any GPU result is compared with it.

Each check is chosen so that a plausible mistake in the oracle (dropped
term, wrong sign, flipped or transposed kernel, off-by-one centre) fails:
bit-exact agreement with the reference's own compiled template, golden
vectors written by it, brute-force O(n^2) DFTs, closed forms (delta kernels,
shifts, tones), and exact identities (adjoint, Parseval, flux conservation).
"""
import numpy as np
import pytest

from conftest import golden_cases, load_golden


def _rng(seed):
    return np.random.default_rng(seed)


# ---------------------------------------------------------------------------
# host convolution restatement (tests/test_algorithms.hpp:9-58)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("shape,kshape", [((8, 9, 10), (3, 3, 3)), ((12, 7, 16), (5, 3, 7)),
                                          ((10, 10, 10), (1, 1, 1)), ((9, 14, 11), (4, 2, 6))])
def test_convolve_host_bitexact_vs_reference_template(oracle, shape, kshape):
    if not oracle.ref_available():
        pytest.skip("oracle/_ref not built (reference tree absent)")
    img = _rng(1).random(shape, dtype=np.float32)
    k = _rng(2).random(kshape, dtype=np.float32)
    padded, off = oracle.zero_padd(img, kshape)
    a = oracle.convolve_host(padded, k, off)
    b = oracle.convolve_host_reference(padded, k, off)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("name", golden_cases())
def test_convolve_host_matches_golden(oracle, name):
    g = load_golden(name)
    padded, off = oracle.zero_padd(g["image"], g["kernel"].shape, int(g["factor"]))
    assert tuple(off) == tuple(g["offsets"])
    full = oracle.convolve_host(padded, g["kernel"], off)
    sl = tuple(slice(o, o + s) for o, s in zip(off, g["image"].shape))
    assert np.array_equal(full[sl], g["expected"])


def test_golden_fixture_closed_forms():
    """tests/test_fixtures.hpp:145-151: identity -> image, trivial -> 0,
    horizontal taps (1,2,3) along x -> out[x] = img[x+1] + 2 img[x] + 3 img[x-1]."""
    g = load_golden("fixture_identity")
    assert np.array_equal(g["expected"], g["image"])
    assert not load_golden("fixture_trivial")["expected"].any()
    g = load_golden("fixture_horizontal")
    im = np.pad(g["image"].astype(np.float64), 1)
    ref = im[1:-1, 1:-1, 2:] + 2 * im[1:-1, 1:-1, 1:-1] + 3 * im[1:-1, 1:-1, :-2]
    assert np.allclose(g["expected"], ref, rtol=1e-6)
    g = load_golden("fixture_depth")
    im = np.pad(g["image"].astype(np.float64), 1)
    ref = im[2:, 1:-1, 1:-1] + 2 * im[1:-1, 1:-1, 1:-1] + 3 * im[:-2, 1:-1, 1:-1]
    assert np.allclose(g["expected"], ref, rtol=1e-6)


# ---------------------------------------------------------------------------
# FFT (replaces cuFFT, src/convolution3Dfft.cu:519-547)
# ---------------------------------------------------------------------------
def _brute_dft(x, sign=-1):
    n = x.size
    j = np.arange(n)
    W = np.exp(sign * 2j * np.pi * np.outer(j, j) / n)
    return W @ x


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 9, 11, 12, 13, 16, 17, 25, 30, 49, 64, 105,
                               109, 120, 125, 173, 210, 343])
def test_fft1d_vs_brute_force(oracle, n):
    x = _rng(n).random(n) + 1j * _rng(n + 1).random(n)
    for sign in (-1, 1):
        got = oracle.fft1d(x, sign)
        ref = _brute_dft(x, sign)
        assert np.max(np.abs(got - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))


@pytest.mark.parametrize("n", [8, 15, 64, 97, 600])
def test_fft1d_closed_forms(oracle, n):
    delta = np.zeros(n, complex)
    delta[0] = 1
    assert np.allclose(oracle.fft1d(delta), np.ones(n), atol=1e-14)
    const = np.ones(n, complex)
    out = oracle.fft1d(const)
    assert abs(out[0] - n) < 1e-10 and np.max(np.abs(out[1:])) < 1e-10
    m = 3 % n
    tone = np.exp(2j * np.pi * m * np.arange(n) / n)
    out = oracle.fft1d(tone)
    assert abs(out[m] - n) < 1e-9
    out[m] = 0
    assert np.max(np.abs(out)) < 1e-9
    x = _rng(n).random(n) + 1j * _rng(2 * n).random(n)
    X = oracle.fft1d(x)
    assert np.isclose(np.sum(np.abs(X) ** 2), n * np.sum(np.abs(x) ** 2), rtol=1e-12)  # Parseval
    assert np.allclose(oracle.fft1d(oracle.fft1d(x), +1) / n, x, atol=1e-13)  # round trip


def test_fft3d_vs_numpy(oracle):
    x = _rng(5).random((6, 10, 14)) + 1j * _rng(6).random((6, 10, 14))
    assert np.allclose(oracle.fft3d(x), np.fft.fftn(x), atol=1e-11)
    assert np.allclose(oracle.fft3d(x, +1), np.fft.ifftn(x) * x.size, atol=1e-11)


# ---------------------------------------------------------------------------
# convolution semantics (src/convolution3Dfft.cu:41-62,128-166,399-578)
# ---------------------------------------------------------------------------
def test_centred_delta_is_identity_and_offcentre_delta_is_roll(oracle):
    img = _rng(7).random((6, 9, 11), dtype=np.float32)
    k = np.zeros((3, 5, 7), np.float32)
    k[1, 2, 3] = 1.0
    assert np.allclose(oracle.circular_conv_direct(img, k), img, atol=0)
    k[:] = 0
    k[2, 0, 6] = 2.5  # tap offset from the centre by (+1, -2, +3)
    ref = 2.5 * np.roll(img.astype(np.float64), (1, -2, 3), axis=(0, 1, 2))
    assert np.allclose(oracle.circular_conv_direct(img, k), ref, rtol=1e-14)
    assert np.allclose(oracle.fft_conv(img, k), ref, atol=1e-12)


def test_fftshift_pad_places_taps(oracle):
    k = np.arange(1, 1 + 3 * 2 * 5, dtype=np.float32).reshape(3, 2, 5)
    P = oracle.fftshift_pad(k, (7, 4, 9))
    # tap (z, y, x) -> ((z-1) mod 7, (y-1) mod 4, (x-2) mod 9)
    for z in range(3):
        for y in range(2):
            for x in range(5):
                assert P[(z - 1) % 7, (y - 1) % 4, (x - 2) % 9] == k[z, y, x]
    assert np.count_nonzero(P) == k.size


@pytest.mark.parametrize("shape,kshape", [((8, 10, 12), (3, 5, 7)), ((5, 13, 6), (5, 1, 3)),
                                          ((9, 9, 9), (9, 9, 9))])
def test_direct_vs_fft_vs_numpy(oracle, shape, kshape):
    img = _rng(9).random(shape, dtype=np.float32)
    k = _rng(10).random(kshape, dtype=np.float32)
    d = oracle.circular_conv_direct(img, k)
    assert oracle.rel_l2(oracle.fft_conv(img, k), d) < 1e-14
    assert oracle.rel_l2(oracle.numpy_fft_conv(img, k), d) < 1e-14


def test_circular_equals_linear_on_padded_interior(oracle):
    """The reference's linear host convolution equals the circular FFT
    convolution on the interior of the zero-padded image (padd_utils.h)."""
    img = _rng(11).random((10, 12, 9), dtype=np.float32)
    k = _rng(12).random((5, 3, 7), dtype=np.float32)
    padded, off = oracle.zero_padd(img, k.shape)
    lin = oracle.convolve_host(padded, k, off)
    circ = oracle.circular_conv_direct(padded, k)
    sl = tuple(slice(o, o + s) for o, s in zip(off, img.shape))
    assert oracle.rel_l2(lin[sl], circ[sl]) < 1e-6


def test_adjoint_identity(oracle):
    a = _rng(13).random((7, 8, 9), dtype=np.float32)
    b = _rng(14).random((7, 8, 9)).astype(np.float32)
    k = _rng(15).random((3, 5, 3), dtype=np.float32)
    lhs = np.vdot(oracle.fft_conv(a, k), b.astype(np.float64))
    rhs = np.vdot(a.astype(np.float64), oracle.fft_conv(b, k, conj_kernel=True))
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)
    # correlation == convolution with the flipped kernel (odd extents)
    assert oracle.rel_l2(oracle.fft_conv(b, k, conj_kernel=True),
                         oracle.fft_conv(b, np.ascontiguousarray(k[::-1, ::-1, ::-1]))) < 1e-13


def test_grad_kernel_vs_finite_differences(oracle):
    img = _rng(16).random((6, 7, 8), dtype=np.float32)
    g = _rng(17).random((6, 7, 8))
    kshape = (3, 3, 5)
    dk = oracle.conv_grad_kernel(g, img, kshape)
    k0 = (np.round(_rng(18).random(kshape) * 8) / 8).astype(np.float32)  # k0 +- 1 exact
    for idx in [(0, 0, 0), (1, 1, 2), (2, 0, 4), (0, 2, 1)]:
        e = np.zeros(kshape, np.float32)
        e[idx] = 1.0
        # the loss is linear in k: a unit central difference is exact
        fp = np.vdot(g, oracle.fft_conv(img, k0 + e))
        fm = np.vdot(g, oracle.fft_conv(img, k0 - e))
        assert abs((fp - fm) / 2 - dk[idx]) < 1e-9 * max(1.0, abs(dk[idx]))


# ---------------------------------------------------------------------------
# helpers restated from the reference tests
# ---------------------------------------------------------------------------
def test_zero_padd_extents_match_reference_tests(oracle):
    # tests/test_gpu_numerical_stability.cpp: 128^3 with 31x31x91 -> 158x158x218
    p, off = oracle.zero_padd(np.zeros((128, 128, 128), np.float32), (91, 31, 31))
    assert p.shape == (218, 158, 158) and off == (45, 15, 15)
    # padd by 10fold (:176): 128^3 with 3x3x5 -> 148x148x168
    p, off = oracle.zero_padd(np.zeros((128, 128, 128), np.float32), (5, 3, 3), 10)
    assert p.shape == (168, 148, 148) and off == (20, 10, 10)
    img = _rng(19).random((3, 4, 5), dtype=np.float32)
    p, off = oracle.zero_padd(img, (3, 5, 1))
    assert p.shape == (5, 8, 5) and np.array_equal(p[1:4, 2:6, 0:5], img) and np.count_nonzero(p) == np.count_nonzero(img)


def test_l2norm_ref_formula(oracle):
    a = _rng(20).random(1000, dtype=np.float32)
    b = _rng(21).random(1000, dtype=np.float32)
    d = (b - a).astype(np.float64)
    assert oracle.l2norm_ref(a, b) == pytest.approx(np.sqrt(np.sum(d * d)) / 1000, rel=1e-12)


# ---------------------------------------------------------------------------
# Richardson-Lucy oracle
# ---------------------------------------------------------------------------
def _gauss(shape, sig):
    zz, yy, xx = np.meshgrid(*[np.arange(s) - s // 2 for s in shape], indexing="ij")
    g = np.exp(-0.5 * ((zz / sig[0]) ** 2 + (yy / sig[1]) ** 2 + (xx / sig[2]) ** 2))
    return (g / g.sum()).astype(np.float32)


def test_rl_flux_conservation_and_numpy_restatement(oracle):
    truth = _rng(22).random((8, 10, 12)).astype(np.float32) + 0.5
    ks = [_gauss((3, 5, 3), (1, 1.5, 0.8)), _gauss((5, 3, 3), (1.2, 0.7, 1.0))]
    views = np.stack([oracle.fft_conv(truth, k).astype(np.float32) for k in ks])
    psi = oracle.rl_multiview(views, ks, 3)
    # sum-normalised PSFs and circular convolution: every update keeps the
    # mean flux of the views
    assert np.isclose(psi.sum(), views.astype(np.float64).sum() / len(ks), rtol=1e-10)
    # independent numpy restatement of the same update
    p = views.astype(np.float64).mean(0)
    for _ in range(3):
        acc = np.zeros_like(p)
        for v, k in zip(views, ks):
            blur = oracle.numpy_fft_conv(p, k)
            acc += oracle.numpy_fft_conv(v / np.maximum(blur, 1e-6), k, conj_kernel=True)
        p = p * acc / len(ks)
    assert oracle.rel_l2(psi, p) < 1e-12
    # the FFT form used at full size
    assert oracle.rel_l2(oracle.fft_rl_multiview(views, ks, 3), psi) < 1e-12


@pytest.mark.parametrize("shape,kshape", [((6, 7, 8), (3, 5, 3)), ((9, 4, 5), (5, 3, 1)),
                                          ((5, 5, 6), (5, 5, 6))])
def test_fft_grad_kernel_vs_direct(oracle, shape, kshape):
    g = _rng(31).random(shape)
    img = _rng(32).random(shape, dtype=np.float32)
    d = oracle.conv_grad_kernel(g, img, kshape)
    f = oracle.fft_conv_grad_kernel(g, img, kshape)
    assert f.shape == tuple(kshape)
    assert oracle.rel_l2(f, d) < 1e-13


# ---------------------------------------------------------------------------
# boundary-mode oracle (the restatement behind convolve3d_boundary)
# ---------------------------------------------------------------------------
_SCIPY_MODE = {"periodic": "grid-wrap", "zero": "constant", "nearest": "nearest",
               "reflect": "reflect", "mirror": "mirror"}


@pytest.mark.parametrize("shape,kshape", [((5, 6, 7), (3, 3, 3)), ((4, 9, 6), (5, 3, 7)),
                                          ((3, 4, 5), (7, 9, 3)), ((1, 6, 5), (1, 3, 5))])
@pytest.mark.parametrize("mode", sorted(_SCIPY_MODE))
def test_boundary_oracle_matches_scipy_ndimage(oracle, shape, kshape, mode):
    """Independent pin: scipy.ndimage.convolve / correlate in float64 with the
    same extension rule, including kernels larger than the image."""
    ndi = pytest.importorskip("scipy.ndimage")
    rng = np.random.default_rng(sum(shape) * 7 + sum(kshape))
    a = rng.random(shape).astype(np.float32)
    k = rng.random(kshape).astype(np.float32)
    ref = ndi.convolve(a.astype(np.float64), k.astype(np.float64), mode=_SCIPY_MODE[mode], cval=0.0)
    got = oracle.boundary_conv_direct(a, k, mode)
    assert np.abs(got - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())
    refc = ndi.correlate(a.astype(np.float64), k.astype(np.float64), mode=_SCIPY_MODE[mode],
                         cval=0.0)
    gotc = oracle.boundary_conv_direct(a, k, mode, correlate=True)
    assert np.abs(gotc - refc).max() <= 1e-12 * max(1.0, np.abs(refc).max())


@pytest.mark.parametrize("shape,kshape", [((5, 6, 7), (3, 3, 3)), ((4, 9, 6), (5, 3, 7)),
                                          ((3, 4, 5), (7, 9, 3)), ((6, 5, 8), (4, 2, 3))])
@pytest.mark.parametrize("mode", ["zero", "nearest", "reflect", "mirror"])
def test_fft_boundary_form_vs_direct(oracle, shape, kshape, mode):
    """The FFT form used at full size against the direct form pinned above,
    for convolution and correlation, odd and even kernels, kernels larger
    than the image."""
    rng = np.random.default_rng(sum(shape) * 3 + sum(kshape))
    a = rng.random(shape).astype(np.float32)
    k = rng.random(kshape).astype(np.float32)
    for corr in (False, True):
        d = oracle.boundary_conv_direct(a, k, mode, correlate=corr)
        f = oracle.fft_boundary_conv(a, k, mode, correlate=corr)
        assert oracle.rel_l2(f, d) < 1e-13, (mode, corr)


def test_boundary_oracle_closed_forms(oracle):
    """Periodic equals the circular oracle; a constant image stays constant
    under nearest/reflect/mirror with a unit-sum kernel; zero boundary of a
    constant image near an edge sums only the kernel taps that land inside."""
    rng = np.random.default_rng(5)
    a = rng.random((6, 7, 8)).astype(np.float32)
    k = rng.random((3, 5, 3)).astype(np.float32)
    assert np.allclose(oracle.boundary_conv_direct(a, k, "periodic"),
                       oracle.circular_conv_direct(a, k), rtol=0, atol=1e-12)
    c = np.full((5, 6, 7), 2.5, np.float32)
    ku = (k / k.sum()).astype(np.float32)
    for mode in ("nearest", "reflect", "mirror"):
        assert np.allclose(oracle.boundary_conv_direct(c, ku, mode), 2.5, rtol=0, atol=1e-6)
    ones = np.ones((4, 4, 4), np.float32)
    kk = np.ones((3, 3, 3), np.float32)
    z = oracle.boundary_conv_direct(ones, kk, "zero")
    assert z[0, 0, 0] == 8.0 and z[1, 1, 1] == 27.0 and z[0, 1, 1] == 18.0
