"""Multiview Richardson-Lucy: fused and per-view GPU forms against the FP64
oracle (oracle.rl_multiview) and against each other.

Tolerances: the FP32 iteration is compared with the FP64 oracle at
rel-L2 <= 1e-4 after 10 iterations (errors compound over iterations); the
fused and per-view GPU forms agree to 1e-5."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(views, psfs, iters, fused):
    import torch
    from fourierconvolutioncudalib_amd.rl import richardson_lucy
    tv = [torch.from_numpy(v).cuda() for v in views]
    tk = [torch.from_numpy(k).cuda() for k in psfs]
    out = richardson_lucy(tv, tk, iterations=iters, eps=1e-6, fused=fused)
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("shape,kshape", [((32, 48, 64), (9, 5, 7)), ((30, 26, 22), (7, 5, 9))])
def test_rl_vs_fp64_oracle(fc, oracle, shape, kshape):
    import torch
    from fourierconvolutioncudalib_amd import rl, synth
    _, views, psfs = synth.multiview(shape, kshape, 4, 77, sigma_xyz=(1.0, 1.0, 2.5),
                                     blobs_count=40)
    ref = oracle.rl_multiview(np.stack(views), psfs, 10, eps=1e-6)
    tv = [torch.from_numpy(v).cuda() for v in views]
    tk = [torch.from_numpy(k).cuda() for k in psfs]
    fusable = rl.fused_available(tv, tk)
    assert fusable == (shape == (32, 48, 64))
    for fused in ([True, False] if fusable else [False]):
        got = _run(views, psfs, 10, fused)
        err = oracle.rel_l2(got, ref)
        assert err < 1e-4, (fused, err)
    if fusable:
        assert oracle.rel_l2(_run(views, psfs, 10, True), _run(views, psfs, 10, False)) < 1e-5


def test_rl_long_axis_shape(fc, oracle):
    """An x axis past one tile (10240, the long-axis path's four-step) in the
    unfused form, against the FP64 oracle (whose own transform of a large
    prime length would be quadratic; the chirp-z axes are checked by the
    convolution tests)."""
    import torch
    from fourierconvolutioncudalib_amd import rl, synth
    shape, kshape = (3, 6, 10240), (3, 5, 9)
    _, views, psfs = synth.multiview(shape, kshape, 2, 78, sigma_xyz=(1.5, 1.0, 1.0),
                                     blobs_count=40)
    ref = oracle.rl_multiview(np.stack(views), psfs, 10, eps=1e-6)
    tv = [torch.from_numpy(v).cuda() for v in views]
    tk = [torch.from_numpy(k).cuda() for k in psfs]
    assert not rl.fused_available(tv, tk)
    assert oracle.rel_l2(_run(views, psfs, 10, False), ref) < 1e-4


def test_rl_multiview_384x384x192(fc, oracle):
    """4 views of 384x384x192 with 25x25x51 PSFs: fused vs per-view GPU forms
    over 10 iterations, flux conservation, and the first two iterations
    against an FP64 numpy restatement of the update."""
    import torch
    from fourierconvolutioncudalib_amd import synth
    cfg = synth.RL_CFG
    _, views, psfs = synth.multiview(cfg["image"], cfg["kernel"], cfg["views"], cfg["seed"])
    fused = _run(views, psfs, 10, True)
    plain = _run(views, psfs, 10, False)
    assert oracle.rel_l2(fused, plain) < 1e-5
    flux = np.mean([v.astype(np.float64).sum() for v in views])
    assert abs(fused.astype(np.float64).sum() - flux) < 1e-4 * flux
    two = _run(views, psfs, 2, True)
    p = np.mean([v.astype(np.float64) for v in views], axis=0)
    for _ in range(2):
        acc = np.zeros_like(p)
        for v, k in zip(views, psfs):
            blur = oracle.numpy_fft_conv(p, k)
            acc += oracle.numpy_fft_conv(v / np.maximum(blur, 1e-6), k, conj_kernel=True)
        p = p * acc / len(views)
    assert oracle.rel_l2(two, p) < 1e-5
