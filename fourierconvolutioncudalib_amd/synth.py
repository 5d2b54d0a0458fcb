"""Seeded synthetic volumes and PSFs for the benchmark workloads and tests.

Uniform values come from a SplitMix64 stream (top 24 bits -> [0, 1)), so the
same seed gives the same float32 volume on every machine.  Volumes are
C-order [z, y, x]; sigmas are given as (x, y, z) like the workload
descriptions.
"""
from __future__ import annotations

import os

import numpy as np

_G = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """n consecutive SplitMix64 outputs of the stream started at `seed`."""
    with np.errstate(over="ignore"):
        z = (np.uint64(seed) + _G * (np.arange(offset, offset + n, dtype=np.uint64) + np.uint64(1)))
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(shape, seed: int) -> np.ndarray:
    """float32 uniform [0, 1) volume."""
    n = int(np.prod(shape))
    out = np.empty(n, dtype=np.float32)
    chunk = 1 << 24
    for s in range(0, n, chunk):
        m = min(chunk, n - s)
        out[s:s + m] = (splitmix64(seed, m, s) >> np.uint64(40)).astype(np.float32) * np.float32(2.0 ** -24)
    return out.reshape(shape)


def _rotation(rng) -> np.ndarray:
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                     [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                     [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


def gaussian_psf(kshape_zyx, sigma_xyz, rotation=None) -> np.ndarray:
    """Sum-normalised anisotropic Gaussian sampled at integer offsets from the
    centre voxel (k // 2 on each axis); optional 3x3 rotation of its axes."""
    kz, ky, kx = kshape_zyx
    z, y, x = np.meshgrid(np.arange(kz) - kz // 2, np.arange(ky) - ky // 2,
                          np.arange(kx) - kx // 2, indexing="ij")
    d = np.stack([x, y, z], axis=-1).astype(np.float64)
    inv = np.diag(1.0 / np.square(np.asarray(sigma_xyz, dtype=np.float64)))
    if rotation is not None:
        inv = rotation @ inv @ rotation.T
    q = np.einsum("...i,ij,...j->...", d, inv, d)
    g = np.exp(-0.5 * q)
    return (g / g.sum()).astype(np.float32)


def random_psf(kshape_zyx, seed: int, sigma_range=(1.0, 6.0)) -> np.ndarray:
    rng = np.random.default_rng(seed)
    sig = rng.uniform(*sigma_range, size=3)
    return gaussian_psf(kshape_zyx, sig, _rotation(rng))


def blobs(shape_zyx, seed: int, count=2000, sigma=(2.0, 6.0), amplitude=(100.0, 1000.0),
          background=100.0, poisson=True) -> np.ndarray:
    """Gaussian blobs on a constant background with optional Poisson noise
    (the light-sheet-like stack of the 600x500x240 workload)."""
    rng = np.random.default_rng(seed)
    vol = np.full(shape_zyx, background, dtype=np.float32)
    nz, ny, nx = shape_zyx
    for _ in range(count):
        c = rng.uniform(0, 1, 3) * np.array([nz, ny, nx])
        s = rng.uniform(*sigma)
        a = rng.uniform(*amplitude)
        r = int(np.ceil(3 * s))
        lo = np.maximum(np.floor(c).astype(int) - r, 0)
        hi = np.minimum(np.floor(c).astype(int) + r + 1, [nz, ny, nx])
        if np.any(hi <= lo):
            continue
        zz = np.arange(lo[0], hi[0])[:, None, None] - c[0]
        yy = np.arange(lo[1], hi[1])[None, :, None] - c[1]
        xx = np.arange(lo[2], hi[2])[None, None, :] - c[2]
        vol[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] += (
            a * np.exp(-(zz * zz + yy * yy + xx * xx) / (2 * s * s))).astype(np.float32)
    if poisson:
        vol = rng.poisson(vol).astype(np.float32)
    return vol


# ---------------------------------------------------------------------------
# the named workloads (image C-order shape, PSF C-order shape, sigma (x,y,z))
# ---------------------------------------------------------------------------
METRIC = dict(name="plain_512x512x256_k15", image=(256, 512, 512), kernel=(15, 15, 15),
              sigma=(1.5, 1.5, 3.0), seed=1000)
CFG1 = dict(name="plain_128x128x64_k15", image=(64, 128, 128), kernel=(15, 15, 15),
            sigma=(1.5, 1.5, 3.0), seed=1000)
CFG2 = dict(name="lightsheet_600x500x240_k31x31x61", image=(240, 500, 600), kernel=(61, 31, 31),
            sigma=(2.0, 2.0, 8.0), seed=1002)


MENU = (64, 80, 96, 105, 112, 120, 128, 144, 160, 192, 210, 224, 240, 256)


def uneven_block_shapes(seed: int, count: int = 64, menu=MENU, kmin=9, kmax=41):
    """(image C-order shape, kernel C-order shape, psf seed) per block: every
    image extent drawn from `menu`, every kernel extent odd in [kmin, kmax]
    and no larger than the image extent."""
    rng = np.random.default_rng(seed)
    out = []
    for b in range(count):
        shape = tuple(int(rng.choice(menu)) for _ in range(3))
        kshape = []
        for n in shape:
            hi = min(kmax, n if n % 2 else n - 1)
            k = int(rng.integers(kmin // 2, hi // 2 + 1)) * 2 + 1
            kshape.append(min(k, hi))
        out.append((shape, tuple(kshape), seed * 1000 + b))
    return out


def rotation_y(angle_deg: float) -> np.ndarray:
    """Rotation of the PSF axes about the y axis (x, y, z order)."""
    a = np.deg2rad(angle_deg)
    c, s = np.cos(a), np.sin(a)
    return np.array([[c, 0.0, s], [0.0, 1.0, 0.0], [-s, 0.0, c]])


def multiview(shape_zyx, kshape_zyx, nviews: int, seed: int, sigma_xyz=(2.0, 2.0, 6.0),
              blobs_count=600, noise=True):
    """(phantom, views, psfs): a blob phantom seen through `nviews` PSFs whose
    long (z) axis is rotated by 180/nviews degrees per view, each view the
    circular convolution of the phantom with its PSF plus Poisson noise."""
    phantom = blobs(shape_zyx, seed, count=blobs_count, background=10.0, poisson=False)
    psfs = [gaussian_psf(kshape_zyx, sigma_xyz, rotation_y(180.0 * v / nviews))
            for v in range(nviews)]
    try:  # threaded pocketfft when scipy is there (input generation only)
        import scipy.fft as sf
        workers = max(1, min(16, len(os.sched_getaffinity(0))))
        rfftn = lambda a: sf.rfftn(a, workers=workers)  # noqa: E731
        irfftn = lambda a, s: sf.irfftn(a, s=s, axes=(0, 1, 2), workers=workers)  # noqa: E731
    except ImportError:
        rfftn = np.fft.rfftn
        irfftn = lambda a, s: np.fft.irfftn(a, s=s, axes=(0, 1, 2))  # noqa: E731
    fp = rfftn(phantom.astype(np.float64))
    rng = np.random.default_rng(seed + 1)
    views = []
    for k in psfs:
        pk = np.zeros(shape_zyx)
        kz, ky, kx = k.shape
        zi = (np.arange(kz) - kz // 2) % shape_zyx[0]
        yi = (np.arange(ky) - ky // 2) % shape_zyx[1]
        xi = (np.arange(kx) - kx // 2) % shape_zyx[2]
        pk[np.ix_(zi, yi, xi)] = k
        blur = irfftn(fp * rfftn(pk), shape_zyx)
        blur = np.maximum(blur, 0.0)
        views.append((rng.poisson(blur) if noise else blur).astype(np.float32))
    return phantom, views, psfs


RL_CFG = dict(name="rl_4x384x384x192_k25x25x51", image=(192, 384, 384), kernel=(51, 25, 25),
              views=4, iterations=10, seed=1005)


def workload(cfg):
    """(image, psf) float32 arrays for a named uniform or blob workload."""
    if cfg is CFG2:
        img = blobs(cfg["image"], cfg["seed"])
    else:
        img = uniform(cfg["image"], cfg["seed"])
    return img, gaussian_psf(cfg["kernel"], cfg["sigma"])
