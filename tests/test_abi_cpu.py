"""C-ABI checks that need no GPU: the library loads, exports every symbol the
public headers declare, and its host-only logic (plans, workspace sizes,
argument validation, device queries without a device) behaves."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT

HEADERS = [os.path.join(ROOT, "include", h)
           for h in ("convolution3Dfft.h", "fourierconvolution_amd.h")]


def _declared_symbols():
    names = []
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"(?:FUNCTION_PREFIX|FCAMD_API)\s+[\w\s\*]+?\b(\w+)\s*\(", text):
            names.append(m.group(1))
    return names


def test_headers_declare_the_reference_abi():
    names = set(_declared_symbols())
    reference = {  # src/convolution3Dfft.h:29-84
        "convolution3DfftCUDA_test", "convolution3DfftCUDA", "convolution3DfftCUDAInPlace",
        "selectDeviceWithHighestComputeCapability", "getCUDAcomputeCapabilityMinorVersion",
        "getCUDAcomputeCapabilityMajorVersion", "getNumDevicesCUDA", "getNameDeviceCUDA",
        "getMemDeviceCUDA", "cuda_version", "gpu_mem_needed_mb"}
    assert reference <= names


def test_library_exports_every_declared_symbol(fc):
    handle = ctypes.CDLL(fc._lib.LIB_PATH)
    missing = [n for n in _declared_symbols() if not hasattr(handle, n)]
    assert not missing, missing
    # and the Python binding knows every one of them
    assert set(_declared_symbols()) <= set(fc._lib.SIGNATURES)


def test_reference_signatures_in_header():
    text = open(HEADERS[0]).read()
    assert re.search(r"void\s+convolution3DfftCUDAInPlace\(imageType\* im, int\* imDim, "
                     r"imageType\* kernel,\s+int\* kernelDim, int devCUDA\)", text)
    assert "typedef float imageType" in text


@pytest.mark.parametrize("n,expect", [(512, [8, 8, 8]), (256, [8, 8, 4]), (600, [8, 3, 5, 5]),
                                      (105, [3, 5, 7]), (1, []), (173, [173]), (346, [2, 173]),
                                      (218, [2, 109])])
def test_radix_plans(fc, n, expect):
    plan = fc.radix_plan(n)
    assert plan == expect
    assert int(np.prod(plan)) == n


def test_workspace_bytes(fc):
    # [nz][ny][nxp] + [kz][ky][nxp] + [kz][ny][nxp] float2, nxp = 257 -> 272,
    # each region rounded up to 256 bytes
    def a256(b):
        return (b + 255) // 256 * 256
    n = fc.workspace_bytes((256, 512, 512), (15, 15, 15))
    assert n == a256(8 * 272 * 512 * 256) + a256(8 * 272 * 15 * 15) + a256(8 * 272 * 15 * 512)
    with pytest.raises(fc.FCError):
        fc.workspace_bytes((8, 8, 8), (9, 3, 3))


def test_long_workspace_bytes_chirp_z(fc):
    """Long-axis workspace (host arithmetic only): spectrum, kernel spectrum,
    kernel volume, packed x rows (nx > 10239), and the chirp-z scratch of an
    axis with no split into two tiles (M = 20580 positions per line for the
    primes 10243, 10247); a prime whose chirp-z length passes two tiles is
    refused."""
    def a256(b):
        return (b + 255) // 256 * 256
    M = 20580
    # y prime: nx = 3 -> 2 bins, row pitch 16; lines = bins x nz
    nz, ny, nx = 2, 10247, 3
    spec = a256(8 * 16 * ny * nz)
    n = fc.workspace_bytes((nz, ny, nx), (1, 9, 3))
    assert n == 2 * spec + a256(4 * nx * ny * nz) + a256(8 * 2 * nz * M)
    # x prime: 5122 bins, pitch 5136; three packed row pairs of nx, lines = pairs
    nz, ny, nx = 2, 3, 10243
    spec = a256(8 * 5136 * ny * nz)
    n = fc.workspace_bytes((nz, ny, nx), (1, 3, 7))
    assert n == 2 * spec + a256(4 * nx * ny * nz) + a256(8 * 3 * nx) + a256(8 * 3 * M)
    # x = 4099 fits one tile, but its plan would hold a direct 4099-point
    # stage: the long-axis layout with packed rows and a one-tile chirp-z
    # (M = 8232); 3700 = 37 * 100 keeps its direct stage and the fused layout
    nz, ny, nx = 8, 64, 4099
    spec = a256(8 * 2064 * ny * nz)
    n = fc.workspace_bytes((nz, ny, nx), (3, 5, 7))
    assert n == (2 * spec + a256(4 * nx * ny * nz) + a256(8 * 256 * nx) + a256(8 * 256 * 8232))
    n = fc.workspace_bytes((8, 64, 3700), (3, 5, 7))
    assert n == a256(8 * 1856 * 64 * 8) + a256(8 * 1856 * 3 * 5) + a256(8 * 1856 * 3 * 64)
    with pytest.raises(fc.FCError) as err:
        fc.workspace_bytes((1, 1, 52428841), (1, 1, 1))
    assert err.value.code == 2 and "no split" in str(err.value)


def test_batch_workspace_lanes(fc, monkeypatch):
    """Batch workspace under each schedule (host arithmetic only): one slot
    per lane, the slot being the lane's largest block; blocks dealt largest
    estimated cost first to the least-loaded lane.  With 0 lanes (length
    groups, one wave) every block has its own region."""
    def a256(b):
        return (b + 255) // 256 * 256
    shapes = [((64, 64, 64), (5, 5, 5)), ((32, 32, 32), (3, 3, 3)), ((128, 64, 64), (5, 5, 5)),
              ((16, 16, 16), (3, 3, 3))]
    ws = [fc.workspace_bytes(s, k) for s, k in shapes]
    n = len(shapes)
    dims = (ctypes.c_int * (3 * n))(*[d for s, _ in shapes for d in s[::-1]])
    kdims = (ctypes.c_int * (3 * n))(*[d for _, k in shapes for d in k[::-1]])
    lib = fc.lib()

    def need(lanes):
        monkeypatch.setenv("FCAMD_BATCH_LANES", str(lanes))
        return lib.fcamd_batch_workspace_bytes(n, dims, kdims)

    assert all(w == a256(w) for w in ws)
    assert need(0) == sum(ws)
    assert need(1) == max(ws)
    assert need(n) == sum(ws) and need(3 * n) == sum(ws)
    # two lanes, costs voxels + 3.5 M (4.02 M, 3.76 M, 3.53 M, 3.50 M for
    # blocks 2, 0, 1, 3): 2 -> lane 0, 0 -> lane 1, 1 -> lane 1, 3 -> lane 0;
    # slots ws[2] and ws[0]
    assert need(2) == ws[2] + ws[0]
    monkeypatch.delenv("FCAMD_BATCH_LANES")
    assert lib.fcamd_batch_workspace_bytes(n, dims, kdims) == need(4)


def test_gpu_mem_needed_mb(fc, monkeypatch):
    """gpu_mem_needed_mb and fcamd_host_device_bytes report what the
    host-pointer path allocates: the single shot (image, workspace, kernel)
    for small volumes, the z-slab ring (within the budget) for large ones.
    With FCAMD_DEVICE_BUDGET_MB set no device is needed to answer."""
    from fourierconvolutioncudalib_amd._lib import ints
    lib = fc.lib()
    monkeypatch.setenv("FCAMD_DEVICE_BUDGET_MB", "100000")
    a256 = lambda b: (b + 255) // 256 * 256  # noqa: E731
    # small volume: single shot
    shape, kshape = (16, 24, 40), (3, 5, 7)
    single = (a256(4 * 16 * 24 * 40) + fc.workspace_bytes(shape, kshape) + a256(4 * 3 * 5 * 7))
    assert lib.fcamd_host_device_bytes(ints(shape[::-1]), ints(kshape[::-1]), 0) == single
    one = a256(4 * 16 * 24 * 40) + fc.workspace_bytes(shape, (1, 1, 1)) + 256
    assert fc.gpu_mem_needed_mb(list(shape)) == -(-one // (1 << 20))
    # 512 x 512 x 256 streams through the slab ring: less than one volume
    # plus its workspace, more than one slab window
    big, k15 = (256, 512, 512), (15, 15, 15)
    ring = lib.fcamd_host_device_bytes(ints(big[::-1]), ints(k15[::-1]), 0)
    plane = 4 * 512 * 512
    assert 16 * plane < ring < 4 * 256 * plane + fc.workspace_bytes(big, k15)
    # a tighter budget gives smaller slabs, within it
    monkeypatch.setenv("FCAMD_DEVICE_BUDGET_MB", "300")
    tight = lib.fcamd_host_device_bytes(ints(big[::-1]), ints(k15[::-1]), 0)
    assert tight <= 300 << 20 and tight < ring
    assert fc.gpu_mem_needed_mb([256, 512, 512]) <= 300
    # below the smallest slab (48-plane transforms need about 215 MiB here)
    # nothing fits: the figure is the single shot's, which the call refuses
    monkeypatch.setenv("FCAMD_DEVICE_BUDGET_MB", "200")
    assert lib.fcamd_host_device_bytes(ints(big[::-1]), ints(k15[::-1]), 0) > 200 << 20
    assert lib.gpu_mem_needed_mb(None, 3) == -1
    assert fc.gpu_mem_needed_mb([0, 4, 4]) == -1


def test_cuda_version_is_hip_version(fc):
    v = fc.cuda_version()
    assert v >= 60000000  # HIP_VERSION of ROCm >= 6


def test_errors_do_not_exit_and_leave_input_untouched(fc):
    im = np.arange(4 * 5 * 6, dtype=np.float32)
    before = im.copy()
    k = np.ones(3 * 3 * 9, dtype=np.float32)
    # kernel longer than the image along x: rejected before any device work
    fc.lib().convolution3DfftCUDAInPlace(im.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                         fc._lib.ints([6, 5, 4]),
                                         k.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                         fc._lib.ints([9, 3, 3]), 0)
    assert fc.lib().fcamd_last_error() == 1
    assert b"exceeds" in fc.lib().fcamd_last_error_message()
    assert np.array_equal(im, before)
    with pytest.raises(fc.FCError):
        fc.convolution3DfftCUDAInPlace(im, [6, 5, 4], k, [9, 3, 3], 0)
    assert not fc.lib().convolution3DfftCUDA(None, None, None, None, 0)
    assert fc.lib().fcamd_last_error() == 1


def test_device_queries_without_device(fc):
    import torch
    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is present")
    assert fc.getNumDevicesCUDA() == 0
    assert fc.selectDeviceWithHighestComputeCapability() == -1
    assert fc.getCUDAcomputeCapabilityMajorVersion(0) == -1
    assert fc.getMemDeviceCUDA(0) == -1
    assert fc.getNameDeviceCUDA(0) == ""
    im = np.ones(8 * 8 * 8, np.float32)
    k = np.ones(27, np.float32)
    assert fc.lib().fcamd_convolve3d_host(im.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                          fc._lib.ints([8, 8, 8]),
                                          k.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                          fc._lib.ints([3, 3, 3]), 0, 0) == 5
    assert np.all(im == 1)


def test_boundary_padding_lengths(fc):
    """Boundary modes pad each axis with a kernel to a fast length >= n + k - 1:
    the smallest 7-smooth compile-time length within 25 % of the smallest
    7-smooth length (host arithmetic only).  526 -> 540 and 270 -> 270 at
    512^2 x 256 with 15^3; 272 -> 288, skipping the registered 286 = 2*11*13
    (padding keeps to 7-smooth lengths)."""
    assert fc.padded_extents((256, 512, 512), (15, 15, 15), "zero") == (270, 540, 540)
    assert fc.padded_extents((1, 1, 260), (1, 1, 13), "zero") == (1, 1, 288)
    assert fc.padded_extents((1, 1, 274), (1, 1, 13), "reflect") == (1, 1, 288)
    assert 286 in fc.static_lengths()
