"""Compiled callers of the drop-in library, as the reference's own callers
are (C++ executables that include convolution3Dfft.h and link
libFourierConvolutionCUDALib, tests/CMakeLists.txt; call pattern
tests/test_gpu_convolve.cpp:233-234): a C99 program (two translation units
including the header) and a C++ program, each linked against the shared
library and against the static archive (tests/consumer/Makefile).

CPU: the four programs are built, start, see the header's constants and, with
no device, report none.  GPU: each convolves golden fixtures through
convolution3DfftCUDAInPlace and convolution3DfftCUDA and the results match
the fixtures (the reference's host convolution) and the FP64 oracle."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, load_golden

BIN = os.path.join(ROOT, "tests", "consumer", "_bin")
PROGS = ["consumer_c_shared", "consumer_c_static", "consumer_cxx_shared", "consumer_cxx_static"]


@pytest.fixture(scope="module")
def programs():
    if not all(os.path.exists(os.path.join(BIN, p)) for p in PROGS):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "consumer")], check=True)
    return {p: os.path.join(BIN, p) for p in PROGS}


def _run(exe, args):
    r = subprocess.run([exe] + [str(a) for a in args], capture_output=True, text=True,
                       timeout=120)
    kv = dict(l.split("=", 1) for l in r.stdout.splitlines() if "=" in l)
    return r.returncode, kv, r.stderr


def test_static_programs_link_every_kernel_group(programs):
    """The static archive has no load-time registration: every group of
    compile-time kernels must be linked in through the registry's calls."""
    for p in ("consumer_c_static", "consumer_cxx_static"):
        syms = subprocess.run(["nm", programs[p]], capture_output=True, text=True,
                              check=True).stdout
        for g in range(6):
            assert f"static_register_group_{g}" in syms, (p, g)


def test_programs_start_without_a_device(programs, tmp_path):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present (covered by the GPU test)")
    im = tmp_path / "im.raw"
    k = tmp_path / "k.raw"
    np.zeros(8, np.float32).tofile(im)
    np.zeros(1, np.float32).tofile(k)
    for p, exe in programs.items():
        rc, kv, err = _run(exe, [im, 2, 2, 2, k, 1, 1, 1, tmp_path / "a", tmp_path / "b"])
        assert kv.get("header_constants_ok") == "1", (p, kv, err)
        assert kv.get("device") == "-1" and rc == 3, (p, rc, kv, err)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["fixture_vertical", "rand_16x20x24_k3x7x5",
                                  "rand_24x20x16_k5x7x3", "rand_13x17x19_k3_pad10"])
def test_programs_convolve_golden_fixtures(programs, oracle, tmp_path, name):
    g = load_golden(name)
    img, k, off = g["image"], g["kernel"], tuple(int(o) for o in g["offsets"])
    padded, off2 = oracle.zero_padd(img, k.shape, int(g["factor"]))
    assert off2 == off
    sl = tuple(slice(o, o + s) for o, s in zip(off, img.shape))
    ref64 = oracle.numpy_fft_conv(padded, k)
    pz, py, px = padded.shape
    kz, ky, kx = k.shape
    padded.astype(np.float32).tofile(tmp_path / "im.raw")
    np.ascontiguousarray(k, np.float32).tofile(tmp_path / "k.raw")
    for p, exe in programs.items():
        print(p, name, flush=True)
        out_a, out_b = tmp_path / f"{p}_inplace.raw", tmp_path / f"{p}_legacy.raw"
        rc, kv, err = _run(exe, [tmp_path / "im.raw", px, py, pz, tmp_path / "k.raw", kx, ky, kz,
                                 out_a, out_b])
        assert rc == 0 and kv.get("status") == "ok", (p, rc, kv, err)
        assert kv["header_constants_ok"] == "1"
        assert kv["cc"].startswith("9."), kv
        assert kv["name"] and int(kv["device"]) >= 0
        if "mem" in kv:
            assert int(kv["mem"]) > (1 << 30)
        for path in (out_a, out_b):
            got = np.fromfile(path, np.float32).reshape(padded.shape)
            assert oracle.rel_l2(got, ref64) < 1e-5, (p, path)
            assert oracle.rel_l2(got[sl], g["expected"]) < 1e-5, (p, path)
