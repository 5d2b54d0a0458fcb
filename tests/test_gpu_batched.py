"""Batched uneven blocks: every block against the FP64 oracle and against
the plain per-block call, eager and through hipGraph replays (each replay
checked, with eager work and synchronisation between replays), under both
schedules (lanes: per-block pass sequences dealt to a few streams; groups:
one launch per transform length and pass)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _blocks(seed, shapes):
    from fourierconvolutioncudalib_amd import synth
    imgs, ks = [], []
    for i, (shape, kshape) in enumerate(shapes):
        imgs.append(synth.uniform(shape, seed + i))
        ks.append(synth.random_psf(kshape, seed + 100 + i))
    return imgs, ks


SHAPES = [((64, 80, 96), (9, 11, 13)), ((105, 64, 112), (21, 9, 15)),
          ((96, 96, 96), (41, 9, 9)), ((120, 144, 64), (9, 41, 11)),
          ((80, 105, 210), (15, 15, 15)), ((64, 64, 64), (9, 9, 9)),
          ((36, 50, 13), (5, 7, 3)),      # lengths without compile-time kernels
          ((64, 1, 80), (9, 1, 5)),       # a 2D block (ny = 1)
          ((160, 224, 240), (11, 13, 17))]


@pytest.fixture(params=["lanes", "groups"])
def schedule(request, monkeypatch):
    monkeypatch.setenv("FCAMD_BATCH_LANES", "4" if request.param == "lanes" else "0")
    return request.param


def test_batched_matches_oracle_and_plain_calls(fc, oracle, schedule):
    import torch
    from fourierconvolutioncudalib_amd.batched import BatchedConvolution
    imgs, ks = _blocks(300, SHAPES)
    ti = [torch.from_numpy(a).cuda() for a in imgs]
    tk = [torch.from_numpy(k).cuda() for k in ks]
    plan = BatchedConvolution(ti, tk)
    outs = [o.cpu().numpy() for o in plan.run(graph=False)]
    for a, k, o, t1, t2 in zip(imgs, ks, outs, ti, tk):
        plain = fc.convolve3d(t1, t2).cpu().numpy()
        assert oracle.rel_l2(o, plain) < 1e-6, a.shape
        assert oracle.rel_l2(o, oracle.numpy_fft_conv(a, k)) < TOL, a.shape
    plan.close()


def test_batched_graph_replays(fc, oracle, schedule):
    import torch
    from fourierconvolutioncudalib_amd.batched import BatchedConvolution
    shapes = SHAPES[:6]
    imgs, ks = _blocks(400, shapes)
    ti = [torch.from_numpy(a).cuda() for a in imgs]
    tk = [torch.from_numpy(k).cuda() for k in ks]
    plan = BatchedConvolution(ti, tk)
    for rep in range(4):
        # eager work between replays: new image contents in place
        for t in ti:
            t.mul_(1.0 + 0.25 * rep).add_(0.5)
        torch.cuda.synchronize()
        got = [o.clone() for o in plan.run(graph=True)]
        torch.cuda.synchronize()
        eager = [fc.convolve3d(a, k) for a, k in zip(ti, tk)]
        torch.cuda.synchronize()
        for g, e in zip(got, eager):
            assert oracle.rel_l2(g.cpu().numpy(), e.cpu().numpy()) < 1e-6, rep
    plan.close()


def test_uneven_64_block_workload(fc, oracle, schedule):
    """The 64-block menu workload (2/3/5/7-smooth extents, odd 9-41 kernels
    with random sigma and rotation): batched vs 64 plain calls for every
    block, FP64 oracle for the smaller blocks."""
    import torch
    from fourierconvolutioncudalib_amd import synth
    from fourierconvolutioncudalib_amd.batched import BatchedConvolution
    spec = synth.uneven_block_shapes(3003, 64)
    g = torch.Generator(device="cuda").manual_seed(3003)
    ti = [torch.rand(s, device="cuda", generator=g) for s, _, _ in spec]
    tk = [torch.from_numpy(synth.random_psf(k, sd)).cuda() for _, k, sd in spec]
    plan = BatchedConvolution(ti, tk)
    outs = plan.run(graph=True)
    torch.cuda.synchronize()
    checked = 0
    for (shape, kshape, _), a, k, o in zip(spec, ti, tk, outs):
        plain = fc.convolve3d(a, k)
        assert oracle.rel_l2(o.cpu().numpy(), plain.cpu().numpy()) < 1e-6, shape
        if np.prod(shape) <= 128 ** 3 and checked < 8:
            ref = oracle.numpy_fft_conv(a.cpu().numpy(), k.cpu().numpy())
            assert oracle.rel_l2(o.cpu().numpy(), ref) < TOL, shape
            checked += 1
    assert checked >= 4
    plan.close()


@pytest.mark.parametrize("wave_mb", [1, 24])
def test_batched_waves(fc, oracle, monkeypatch, wave_mb):
    """Blocks split into workspace-budget waves sharing one workspace region
    (1 MiB: one block per wave; 24 MiB: a few blocks per wave): equal to the
    plain calls, eager and over repeated graph replays."""
    import ctypes
    import torch
    from fourierconvolutioncudalib_amd.batched import BatchedConvolution
    imgs, ks = _blocks(500, SHAPES)
    ti = [torch.from_numpy(a).cuda() for a in imgs]
    tk = [torch.from_numpy(k).cuda() for k in ks]
    dims = [d for a in imgs for d in a.shape[::-1]]
    kdims = [d for k in ks for d in k.shape[::-1]]
    n = len(imgs)
    lib = fc.lib()
    monkeypatch.setenv("FCAMD_BATCH_LANES", "0")  # waves belong to the group schedule
    one_wave = lib.fcamd_batch_workspace_bytes(n, (ctypes.c_int * (3 * n))(*dims),
                                               (ctypes.c_int * (3 * n))(*kdims))
    monkeypatch.setenv("FCAMD_BATCH_WAVE_MB", str(wave_mb))
    waved = lib.fcamd_batch_workspace_bytes(n, (ctypes.c_int * (3 * n))(*dims),
                                            (ctypes.c_int * (3 * n))(*kdims))
    assert 0 < waved < one_wave
    plan = BatchedConvolution(ti, tk)
    plain = [fc.convolve3d(a, k).cpu().numpy() for a, k in zip(ti, tk)]
    eager = [o.cpu().numpy() for o in plan.run(graph=False)]
    for e, p in zip(eager, plain):
        assert oracle.rel_l2(e, p) < 1e-6
    for rep in range(3):
        got = [o.clone() for o in plan.run(graph=True)]
        torch.cuda.synchronize()
        for g, p in zip(got, plain):
            assert oracle.rel_l2(g.cpu().numpy(), p) < 1e-6, rep
    plan.close()


@pytest.mark.parametrize("lanes", [1, 3, 16])
def test_batched_lanes(fc, oracle, monkeypatch, lanes):
    """The lane schedule with one lane, a few lanes and more lanes than
    blocks: its workspace (one slot per lane) is at most the group
    schedule's, and the results equal the plain calls, eager and over
    repeated graph replays with eager work between them."""
    import ctypes
    import torch
    from fourierconvolutioncudalib_amd.batched import BatchedConvolution
    imgs, ks = _blocks(600, SHAPES)
    ti = [torch.from_numpy(a).cuda() for a in imgs]
    tk = [torch.from_numpy(k).cuda() for k in ks]
    n = len(imgs)
    dims = (ctypes.c_int * (3 * n))(*[d for a in imgs for d in a.shape[::-1]])
    kdims = (ctypes.c_int * (3 * n))(*[d for k in ks for d in k.shape[::-1]])
    lib = fc.lib()
    monkeypatch.setenv("FCAMD_BATCH_LANES", "0")
    grouped = lib.fcamd_batch_workspace_bytes(n, dims, kdims)
    monkeypatch.setenv("FCAMD_BATCH_LANES", str(lanes))
    laned = lib.fcamd_batch_workspace_bytes(n, dims, kdims)
    assert 0 < laned <= grouped
    plan = BatchedConvolution(ti, tk)
    plain = [fc.convolve3d(a, k).cpu().numpy() for a, k in zip(ti, tk)]
    for e, p in zip([o.cpu().numpy() for o in plan.run(graph=False)], plain):
        assert oracle.rel_l2(e, p) < 1e-6
    for rep in range(3):
        for o in plan.outs:
            o.fill_(float("nan"))
        torch.cuda.synchronize()
        got = [o.clone() for o in plan.run(graph=True)]
        torch.cuda.synchronize()
        for g, p in zip(got, plain):
            assert oracle.rel_l2(g.cpu().numpy(), p) < 1e-6, rep
    plan.close()


def test_batched_in_place_mixed_runs(fc, oracle, schedule):
    """Outputs aliasing the images, with graph and eager runs of one plan
    alternating: every run equals the plain call on the images it read."""
    import torch
    from fourierconvolutioncudalib_amd.batched import BatchedConvolution
    imgs, ks = _blocks(700, SHAPES[:6])
    ti = [torch.from_numpy(a).cuda() for a in imgs]
    tk = [torch.from_numpy(k).cuda() for k in ks]
    plan = BatchedConvolution(ti, tk, outs=ti)
    for rep, graph in enumerate((True, False, True, True, False)):
        want = [fc.convolve3d(a, k) for a, k in zip(ti, tk)]
        torch.cuda.synchronize()
        plan.run(graph=graph)
        torch.cuda.synchronize()
        for g, w in zip(ti, want):
            assert oracle.rel_l2(g.cpu().numpy(), w.cpu().numpy()) < 1e-6, (rep, graph)
    plan.close()


LONG_SHAPES = [((2, 3, 20000), (1, 3, 9)),     # x past one tile: four-step
               ((8, 64, 4099), (3, 5, 7)),     # prime x: chirp-z
               ((64, 80, 96), (9, 11, 13)),
               ((36, 50, 13), (5, 7, 3))]


def test_batched_lanes_with_long_axis_blocks(fc, oracle, monkeypatch):
    """Blocks the drop-in call sends to the long-axis path run there in the
    lane schedule, beside one-tile blocks: eager and over graph replays,
    equal to the plain calls and to FP64.  The length-group schedule refuses
    them before any device work."""
    import ctypes
    import torch
    from fourierconvolutioncudalib_amd.batched import BatchedConvolution
    monkeypatch.setenv("FCAMD_BATCH_LANES", "4")
    imgs, ks = _blocks(900, LONG_SHAPES)
    ti = [torch.from_numpy(a).cuda() for a in imgs]
    tk = [torch.from_numpy(k).cuda() for k in ks]
    plan = BatchedConvolution(ti, tk)
    plain = [fc.convolve3d(a, k).cpu().numpy() for a, k in zip(ti, tk)]
    for e, p, a, k in zip([o.cpu().numpy() for o in plan.run(graph=False)], plain, imgs, ks):
        assert oracle.rel_l2(e, p) < 1e-6
        assert oracle.rel_l2(e, oracle.numpy_fft_conv(a, k)) < TOL
    for rep in range(3):
        for o in plan.outs:
            o.fill_(float("nan"))
        torch.cuda.synchronize()
        got = [o.clone() for o in plan.run(graph=True)]
        torch.cuda.synchronize()
        for g, p in zip(got, plain):
            assert oracle.rel_l2(g.cpu().numpy(), p) < 1e-6, rep
    plan.close()
    n = len(imgs)
    dims = (ctypes.c_int * (3 * n))(*[d for a in imgs for d in a.shape[::-1]])
    kdims = (ctypes.c_int * (3 * n))(*[d for k in ks for d in k.shape[::-1]])
    monkeypatch.setenv("FCAMD_BATCH_LANES", "0")
    assert fc.lib().fcamd_batch_workspace_bytes(n, dims, kdims) < 0
    assert fc.lib().fcamd_last_error() == 2
