"""Every transform length with a compile-time kernel, on every axis, against
the FP64 oracle; compile-time and runtime-length kernels against each other
at the benchmark shape; runtime-length kernels (incl. prime radices) on
lengths without compile-time kernels."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _lengths():
    import fourierconvolutioncudalib_amd as fc
    return fc.static_lengths()


def _check(fc, oracle, shape_zyx, kshape_zyx, seed):
    from fourierconvolutioncudalib_amd import synth
    img = synth.uniform(shape_zyx, seed)
    k = synth.gaussian_psf(kshape_zyx, (1.3, 0.9, 1.7))
    k[0, 0, 0] += 0.25  # break the PSF's symmetry so a flip would show
    out = fc.convolve(img, k)
    ref = oracle.numpy_fft_conv(img, k)
    err = oracle.rel_l2(out, ref)
    assert err < TOL, (shape_zyx, kshape_zyx, err)
    return err


@pytest.mark.parametrize("axis", ["x", "y", "z"])
def test_every_static_length(fc, oracle, axis):
    lengths = _lengths()
    assert {512, 256, 600, 500, 240, 384, 192, 128}.issubset(lengths)
    for i, L in enumerate(lengths):
        if axis == "x":
            shape = (6, 10, L)
        elif axis == "y":
            shape = (6, L, 24)
        else:
            shape = (L, 10, 24)
        kshape = tuple(min(s, k) for s, k in zip(shape, (5, 7, 3)))
        _check(fc, oracle, shape, kshape, 100 + i)


@pytest.mark.parametrize("shape", [(13, 17, 19), (15, 19, 21), (46, 46, 106), (22, 26, 34),
                                   (11, 121, 9), (1, 1, 3000), (2, 1100, 1), (700, 3, 2)])
def test_runtime_length_kernels(fc, oracle, shape):
    """Lengths without compile-time kernels (generic radices 11, 13, 17, 19,
    23, 53 and long axes)."""
    kshape = tuple(min(s, k) for s, k in zip(shape, (3, 5, 7)))
    _check(fc, oracle, shape, kshape, 7)


# lengths whose largest prime factor exceeds 31 run Bluestein's chirp-z
# (a circular convolution of 7-smooth length M >= 2L - 1, M <= 6826);
# past that, 3700 = 37 * 100 keeps its direct 37-point stage (primes up to
# 41) and 3737 = 37 * 101 and 4091 take the chirp-z of the long-axis path
@pytest.mark.parametrize("L", [37, 79, 109, 173, 346, 1009, 2018, 3700, 3737, 4091])
@pytest.mark.parametrize("axis", ["x", "y", "z"])
def test_bluestein_lengths(fc, oracle, axis, L):
    from fourierconvolutioncudalib_amd import synth
    shape = {"x": (3, 5, L), "y": (3, L, 6), "z": (L, 4, 6)}[axis]
    kshape = tuple(min(n, k) for n, k in zip(shape, (3, 5, 7)))
    print("bluestein", axis, L, flush=True)
    img = synth.uniform(shape, 5000 + L)
    k = synth.random_psf(kshape, 5001 + L)
    for conj in (False, True):
        out = fc.convolve(img, k, op=fc.CORRELATE if conj else fc.CONVOLVE)
        ref = oracle.numpy_fft_conv(img, k, conj_kernel=conj)
        assert oracle.rel_l2(out, ref) < TOL, (axis, L, conj)


def test_static_matches_runtime_kernels_at_benchmark_shape(fc, oracle):
    import torch
    from fourierconvolutioncudalib_amd import synth
    img, psf = synth.workload(synth.METRIC)
    ti = torch.from_numpy(img).cuda()
    tk = torch.from_numpy(psf).cuda()
    fast = fc.convolve3d(ti, tk).cpu().numpy()
    fc.use_static_kernels(False)
    try:
        slow = fc.convolve3d(ti, tk).cpu().numpy()
    finally:
        fc.use_static_kernels(True)
    assert oracle.rel_l2(fast, slow) < 2e-6
    ref = oracle.numpy_fft_conv(img, psf)
    assert oracle.rel_l2(fast, ref) < TOL
    assert oracle.rel_l2(slow, ref) < TOL


def test_correlate_static(fc, oracle):
    from fourierconvolutioncudalib_amd import synth
    img = synth.uniform((64, 80, 96), 9)
    k = synth.random_psf((9, 11, 7), 3)
    out = fc.convolve(img, k, op=fc.CORRELATE)
    assert oracle.rel_l2(out, oracle.numpy_fft_conv(img, k, conj_kernel=True)) < TOL


# Kernel extents of the timed workloads along the transformed axis (15, 21,
# 25, 31, 41, 51, 61) and a short one: the fused z pass picks its tap path,
# the general FFT path, the direct DFT or that variant's fallback from the
# kernel's z extent, so every (length, variant) runs at each extent a
# workload can give it.
KERNEL_EXTENTS = (3, 15, 21, 25, 31, 41, 51, 61)
PASS_NAMES = {0: "x forward", 1: "x inverse", 2: "y forward", 3: "fused z", 4: "y inverse"}


def _variant_cases(p, L):
    """(image shape, kernel shape) pairs, C order, that put length L on the
    axis pass p transforms."""
    ext = [k for k in KERNEL_EXTENTS if k <= L]
    if p in (0, 1):
        return [((6, 10, L), (3, 5, k)) for k in (ext[0], ext[-1])]
    if p in (2, 4):
        return [((6, L, 24), (3, k, 5)) for k in (ext[0], ext[-1])]
    return [((L, 6, 18), (k, 5, 3)) for k in ext]


@pytest.mark.parametrize("p", [0, 1, 2, 3, 4])
def test_every_variant(fc, oracle, p):
    """Every compile-time variant of pass p at every registered length,
    selected in turn, against the FP64 oracle: whatever the autotuner picks
    for a timed workload has run here at that workload's length and kernel
    extent."""
    from fourierconvolutioncudalib_amd import synth
    lib = fc.lib()
    checked = 0
    for L in _lengths():
        prev = lib.fcamd_selected_static_variant(p, L)
        try:
            n = lib.fcamd_static_variant_count(p, L)
            assert n >= 1
            for v in range(n):
                assert lib.fcamd_select_static_variant(p, L, v) == 0
                name = lib.fcamd_static_variant_name(p, L, v).decode()
                for shape, kshape in _variant_cases(p, L):
                    img = synth.uniform(shape, 1000 * p + L + v)
                    k = synth.random_psf(kshape, 7 * v + kshape[0])
                    err = oracle.rel_l2(fc.convolve(img, k), oracle.numpy_fft_conv(img, k))
                    assert err < TOL, (PASS_NAMES[p], L, name, shape, kshape, err)
                    checked += 1
        finally:
            lib.fcamd_select_static_variant(p, L, prev)
    print(f"{PASS_NAMES[p]}: {checked} (length, variant, kernel) cases", flush=True)


# z extents around the fused z pass's paths for two-stage plans (L = R0 x R1):
# the tap path up to R1 taps, the kernel column's own transform past that,
# up to the longest kernels
Z_PATH_EXTENTS = {64: (8, 9, 17, 32, 33), 128: (8, 9, 21, 64, 65), 240: (15, 16, 61, 64, 65),
                  256: (16, 17, 40, 64, 65)}


@pytest.mark.parametrize("L", sorted(Z_PATH_EXTENTS))
def test_z_paths_by_kernel_extent(fc, oracle, L):
    """Fused z pass at kernel extents on both sides of each path boundary,
    convolution and correlation, odd and even extents (centre kz // 2),
    against the FP64 oracle."""
    from fourierconvolutioncudalib_amd import synth
    for kz in Z_PATH_EXTENTS[L]:
        print("z path", L, kz, flush=True)
        img = synth.uniform((L, 6, 18), 40 * L + kz)
        k = synth.random_psf((kz, 5, 3), 3 * kz + 1)
        err = oracle.rel_l2(fc.convolve(img, k), oracle.numpy_fft_conv(img, k))
        assert err < TOL, (L, kz, "convolve", err)
        out = fc.convolve(img, k, op=fc.CORRELATE)
        err = oracle.rel_l2(out, oracle.numpy_fft_conv(img, k, conj_kernel=True))
        assert err < TOL, (L, kz, "correlate", err)


@pytest.mark.parametrize("cfg", ["METRIC", "CFG1", "CFG2"])
def test_tuned_workloads_full_size(fc, oracle, cfg):
    """The plain workloads the benchmark times, at full size, with the
    variants the autotuner selects for them (as the benchmark does before
    timing), against the FP64 oracle."""
    import torch
    from fourierconvolutioncudalib_amd import synth
    lib = fc.lib()
    w = getattr(synth, cfg)
    img, psf = synth.workload(w)
    nz, ny, nx = img.shape
    keep = {(q, L): lib.fcamd_selected_static_variant(q, L)
            for q, L in ((0, nx), (1, nx), (2, ny), (3, nz), (4, ny))}
    print(cfg, img.shape, psf.shape, flush=True)
    try:
        ti = torch.from_numpy(img).cuda()
        tk = torch.from_numpy(psf).cuda()
        fc.autotune(ti.shape, tk.shape, reps=2)
        tuned = fc.selected_variants(ti.shape)
        out = fc.convolve3d(ti, tk).cpu().numpy()
    finally:
        for (q, L), v in keep.items():
            if v >= 0:
                lib.fcamd_select_static_variant(q, L, v)
    ref = oracle.numpy_fft_conv(img, psf)
    err = oracle.rel_l2(out, ref)
    print(cfg, tuned, f"rel-L2 {err:.2e} max-abs {oracle.max_abs(out, ref):.2e}", flush=True)
    assert err < TOL, (cfg, tuned, err)


@pytest.mark.parametrize("L", [64, 105, 128, 210, 240, 512])
def test_flat_x_variants_many_workgroups(fc, oracle, L):
    """The flat x variants (workgroup-wide staging through LDS) over many
    workgroups with a ragged last one, with 16-byte-aligned tensors (float4
    path where L % 4 == 0) and with image and output one float off that
    alignment (scalar path)."""
    import torch
    from fourierconvolutioncudalib_amd import synth
    lib = fc.lib()
    shape, kshape = (7, 41, L), (5, 7, 3)
    img = synth.uniform(shape, 31)
    k = synth.random_psf(kshape, 5)
    ref = oracle.numpy_fft_conv(img, k)
    tk = torch.from_numpy(k).cuda()
    n = img.size
    # one row and three rows (a single, ragged workgroup)
    few = [(synth.uniform(s, 33), synth.random_psf((1, min(3, s[1]), 5), 6)) for s in
           ((1, 1, L), (1, 3, L))]
    prev = {p: lib.fcamd_selected_static_variant(p, L) for p in (0, 1)}
    try:
        for name in ("w256-flat", "w128-flat"):
            for p in (0, 1):
                names = [lib.fcamd_static_variant_name(p, L, v).decode()
                         for v in range(lib.fcamd_static_variant_count(p, L))]
                assert lib.fcamd_select_static_variant(p, L, names.index(name)) == 0
            out = fc.convolve(img, k)
            assert oracle.rel_l2(out, ref) < TOL, (name, "aligned")
            for a, kk in few:
                assert oracle.rel_l2(fc.convolve(a, kk), oracle.numpy_fft_conv(a, kk)) < TOL, a.shape
            buf = torch.zeros(n + 1, device="cuda")
            ti = buf[1:].view(shape)
            ti.copy_(torch.from_numpy(img).cuda())
            obuf = torch.zeros(n + 1, device="cuda")
            to = obuf[1:].view(shape)
            assert ti.data_ptr() % 16 != 0 and to.data_ptr() % 16 != 0
            fc.convolve3d(ti, tk, out=to)
            assert oracle.rel_l2(to.cpu().numpy(), ref) < TOL, (name, "unaligned")
            assert obuf[0].item() == 0.0  # nothing written before the output
    finally:
        for p in (0, 1):
            lib.fcamd_select_static_variant(p, L, prev[p])


def test_autotune_keeps_results_and_writes_cache(fc, oracle, tmp_path, monkeypatch):
    """The autotuner times every variant, keeps the fastest, records it in
    the cache file, and the tuned path still matches the oracle."""
    from fourierconvolutioncudalib_amd import synth
    cache = tmp_path / "tune.txt"
    monkeypatch.setenv("FCAMD_TUNE_CACHE", str(cache))
    shape, kshape = (64, 96, 128), (9, 7, 5)
    lib = fc.lib()
    tuned = [(0, 128), (1, 128), (2, 96), (3, 64), (4, 96)]
    prev = {pl: lib.fcamd_selected_static_variant(*pl) for pl in tuned}
    fc.autotune(shape, kshape, reps=2)
    sel = fc.selected_variants(shape)
    assert set(sel) == {"x_forward", "x_inverse", "y_forward", "z", "y_inverse"}
    lines = cache.read_text().splitlines()
    assert len(lines) == 5 and all(len(l.split()) == 4 for l in lines)
    img = synth.uniform(shape, 17)
    k = synth.random_psf(kshape, 18)
    assert oracle.rel_l2(fc.convolve(img, k), oracle.numpy_fft_conv(img, k)) < 1e-5
    # restore the defaults for the tests that follow
    for (p, L), v in prev.items():
        lib.fcamd_select_static_variant(p, L, v)
