"""Writes the golden vectors under tests/golden/ from the reference's own
host convolution template (tests/test_algorithms.hpp:9-58), compiled from the
reference sources into oracle/_ref/ by oracle/Makefile.

Run in the build container (where the reference tree exists):
    python tests/golden/make_golden.py
Each .npz holds float32 `image` (unpadded, C-order [z, y, x]), `kernel`
([kz, ky, kx]), the zero-padding `offsets` and `expected` = the template's
result on the padded interior, i.e. the linear convolution of image and
kernel (only odd kernel extents: the template centres even kernels at
k/2 - 1, the GPU path at k/2).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402


def fixture_ramp(n=8):
    """tests/test_fixtures.hpp:197-204: 8^3 ramp image_[z][y][x] = index."""
    return np.arange(n ** 3, dtype=np.float32).reshape(n, n, n)


def fixture_kernels(k=3):
    """tests/test_fixtures.hpp:125-151: 3^3 kernels of the legacy suite."""
    h = k // 2
    ks = {}
    ks["trivial"] = np.zeros((k, k, k), np.float32)
    ident = np.zeros((k, k, k), np.float32)
    ident.ravel()[k ** 3 // 2] = 1.0
    ks["identity"] = ident
    hor = np.zeros((k, k, k), np.float32)
    ver = np.zeros((k, k, k), np.float32)
    dep = np.zeros((k, k, k), np.float32)
    for i in range(k):
        hor[h, h, i] = i + 1
        ver[h, i, h] = i + 1
        dep[i, h, h] = i + 1
    ks["horizontal"] = hor
    ks["vertical"] = ver
    ks["depth"] = dep
    ks["all1"] = np.ones((k, k, k), np.float32)
    return ks


def seeded(shape, seed):
    return np.random.default_rng(seed).random(shape, dtype=np.float32)


def write(name, image, kernel, factor=1):
    padded, off = O.zero_padd(image, kernel.shape, factor)
    full = O.convolve_host_reference(padded, kernel, off)
    sl = tuple(slice(o, o + s) for o, s in zip(off, image.shape))
    expected = np.ascontiguousarray(full[sl])
    np.savez_compressed(os.path.join(HERE, name + ".npz"), image=image, kernel=kernel,
                        offsets=np.array(off, np.int32), factor=np.int32(factor),
                        expected=expected)
    print(name, image.shape, kernel.shape, "->", padded.shape)


def main():
    if not O.ref_available():
        O.build()
    ramp = fixture_ramp()
    for kname, k in fixture_kernels().items():
        write("fixture_" + kname, ramp, k)
    write("rand_16x20x24_k3x7x5", seeded((16, 20, 24), 11), seeded((3, 7, 5), 12))
    write("rand_24x20x16_k5x7x3", seeded((24, 20, 16), 21), seeded((5, 7, 3), 22))
    write("rand_8x16x16_k15", seeded((8, 16, 16), 31), seeded((15, 15, 15), 32))
    write("rand_13x17x19_k3_pad10", seeded((19, 17, 13), 41), seeded((5, 3, 3), 42), factor=10)


if __name__ == "__main__":
    main()
