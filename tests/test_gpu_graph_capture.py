"""Convolutions recorded into a caller's graph (torch.cuda.CUDAGraph on
torch's graph-capture stream): a plan execute (fcamd_conv_plan_execute) and
the device entry point with a caller workspace (fcamd_convolve3d_device).
Each graph is replayed several times with new input data and eager work of
the program between replays; every replay matches the eager result and the
FP64 oracle.  A transform length used for the first time inside a capture
is refused with an error instead of allocating into the capture."""
import pytest

pytestmark = pytest.mark.gpu


def _replays(torch, g, src, data, outs_fn):
    results = []
    for d in data:
        src.copy_(d)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        results.append(outs_fn())
        # eager work of the program between replays, on the same buffers
        src.mul_(0.5)
        torch.cuda.synchronize()
    return results


@pytest.mark.parametrize("shape,kshape", [((64, 96, 128), (9, 7, 5)), ((40, 50, 60), (5, 3, 7)),
                                          ((32, 128, 128), (7, 7, 7)), ((24, 256, 256), (9, 9, 9)),
                                          # long-axis path: chirp-z x, four-step x
                                          ((8, 64, 4099), (3, 5, 7)), ((2, 3, 20000), (1, 3, 9))])
def test_plan_and_device_call_in_cuda_graph(fc, oracle, shape, kshape):
    import torch
    from fourierconvolutioncudalib_amd import synth
    print("graph capture", shape, kshape, flush=True)
    k = torch.from_numpy(synth.random_psf(kshape, 5)).cuda()
    data = [torch.from_numpy(synth.uniform(shape, 70 + i)).cuda() for i in range(3)]
    src = torch.empty(shape, device="cuda")
    out_plan = torch.empty(shape, device="cuda")
    out_dev = torch.empty(shape, device="cuda")
    ws = torch.empty(fc.workspace_bytes(shape, kshape), dtype=torch.uint8, device="cuda")
    plan = fc.ConvolutionPlan(shape, k)
    # one eager call of this shape creates its twiddle tables before capture
    src.copy_(data[0])
    fc.convolve3d(src, k, out=out_dev, workspace=ws)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            plan.run(src, out=out_plan, stream=s)
            fc.convolve3d(src, k, out=out_dev, workspace=ws, stream=s)
    torch.cuda.synchronize()
    got = _replays(torch, g, src, data,
                   lambda: (out_plan.cpu().numpy(), out_dev.cpu().numpy()))
    for d, (a, b) in zip(data, got):
        eager = fc.convolve3d(d, k).cpu().numpy()
        ref = oracle.numpy_fft_conv(d.cpu().numpy(), k.cpu().numpy())
        assert oracle.rel_l2(a, ref) < 1e-5 and oracle.rel_l2(b, ref) < 1e-5
        assert oracle.rel_l2(a, eager) < 1e-6 and oracle.rel_l2(b, eager) < 1e-6
    plan.close()


def test_first_use_of_a_length_inside_capture_is_refused(fc, oracle):
    """The x length 4091 (prime) appears in no other test, so its twiddle
    tables do not exist yet and the first pass of the call inside the capture
    must fail before launching anything; afterwards the same call runs
    eagerly."""
    import torch
    from fourierconvolutioncudalib_amd import synth
    shape, kshape = (29, 23, 4091), (3, 3, 5)
    a = torch.from_numpy(synth.uniform(shape, 3)).cuda()
    k = torch.from_numpy(synth.random_psf(kshape, 4)).cuda()
    out = torch.empty_like(a)
    ws = torch.empty(fc.workspace_bytes(shape, kshape), dtype=torch.uint8, device="cuda")
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    with pytest.raises(fc.FCError, match="first time inside a stream capture"):
        with torch.cuda.stream(s):
            with torch.cuda.graph(g, stream=s):
                fc.convolve3d(a, k, out=out, workspace=ws, stream=s)
    torch.cuda.synchronize()
    fc.convolve3d(a, k, out=out, workspace=ws)
    ref = oracle.numpy_fft_conv(a.cpu().numpy(), k.cpu().numpy())
    assert oracle.rel_l2(out.cpu().numpy(), ref) < 1e-5
