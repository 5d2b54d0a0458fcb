#!/usr/bin/env python3
"""Benchmark of the 3D FFT convolution hot path on one MI355X.

Workload: plain in-place-semantics circular convolution of a 512x512x256
(x, y, z) float32 image (seeded uniform [0,1)) with a 15x15x15 anisotropic
Gaussian PSF (sigma 1.5, 1.5, 3.0), inputs resident in HBM.  One step = one
full convolution (kernel spectrum included), through the library's C ABI
(fcamd_convolve3d_device) on a torch-allocated buffer and stream.

Prints ONE JSON line (rank 0) on stdout; progress goes to stderr.  The line
holds voxels/s, the whole-convolution HBM roofline fraction (bytes_alg =
8N + 64Nc + 4K per convolution), the roofline of the dominant kernel from
HIP events recorded around each pass over a second timed region, with the
kernel's algorithmic bytes of the roofline model (16 Nc for the fused z
pass: the spectrum read and written once, the kernel's compact planes
counted free), and the side workloads (plan, boundary modes, batched blocks,
Richardson-Lucy, differentiation, the 128x128x64 and 600x500x240 plain
workloads).

The CPU leg runs after all GPU timing: the reference's host convolution
(tests/test_algorithms.hpp:9-58, compiled into oracle/_ref) timed on the
host cores (cpu_baseline), then the FP64 oracle checking the host copy of
every timed GPU output (parity: rel-L2 and max-abs per output).
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s
TOL = 1e-5         # rel-L2 bar of every convolution output
TOL_RL = 1e-4      # rel-L2 bar of the 10-iteration Richardson-Lucy estimate


def log(msg):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def pass_bytes(nx, ny, nz, kx, ky, kz):
    """Algorithmic HBM bytes per pass launch and per convolution (roofline
    model: every pass reads and writes its volume once; the kernel is read
    once, 4K, and its transforms act on its compact support, counted free)."""
    N = nx * ny * nz
    nb = nx // 2 + 1
    Nc = nb * ny * nz
    return {
        "kernel_rows_xfwd": 4 * kx * ky * kz,
        "kernel_planes_yfwd": 0,
        "image_xfwd": 4 * N + 8 * Nc,
        "image_yfwd": 16 * Nc,
        "zfused": 16 * Nc,
        "image_yinv": 16 * Nc,
        "image_xinv": 8 * Nc + 4 * N,
    }, 8 * N + 64 * Nc + 4 * kx * ky * kz


def cpu_threads():
    """Host threads of the CPU leg: the cores this process may run on, at
    most the host share of one GPU (OMP_NUM_THREADS, 16 where unset)."""
    avail = len(os.sched_getaffinity(0))
    share = int(os.environ.get("OMP_NUM_THREADS") or 16)
    return max(1, min(avail, share)), avail


# ---------------------------------------------------------------------------
# GPU side.  Each part returns its result and registers host copies of its
# timed outputs in `checks` for the CPU leg.
# ---------------------------------------------------------------------------
class Checks:
    """Host copies of timed GPU outputs; each entry is checked against the
    FP64 oracle in the CPU leg."""

    def __init__(self):
        self.items = []

    def add(self, name, fn, tol=TOL):
        """fn(O) -> (got, ref) or a list of them, O the oracle module."""
        self.items.append((name, fn, tol))


def timed_calls(torch, stream, fn, reps, warm=2):
    for _ in range(warm):
        fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


# variants are tuned per workload unless --tuned applies recorded choices
TUNE = True


def config_timing(fc, torch, cfg, steps, checks, cpu_jobs, cpu_planes, cpu_single):
    """A further workload timed like the metric: device-resident, autotuned
    variants (restored afterwards), HIP events around `steps` calls; its
    whole-convolution roofline fraction.  Its output is checked against the
    FP64 oracle and its CPU baseline timed in the CPU leg."""
    from fourierconvolutioncudalib_amd import synth
    img_np, psf_np = synth.workload(cfg)
    nz, ny, nx = img_np.shape
    kz, ky, kx = psf_np.shape
    img = torch.from_numpy(img_np).cuda()
    psf = torch.from_numpy(psf_np).cuda()
    out = torch.empty_like(img)
    ws = torch.empty(fc.workspace_bytes(img.shape, psf.shape), dtype=torch.uint8, device="cuda")
    stream = torch.cuda.current_stream()
    lib = fc.lib()
    # selections of the lengths the tuner may change (-1: no compile-time kernel)
    keep = {(p, L): lib.fcamd_selected_static_variant(p, L)
            for p, L in ((0, nx), (1, nx), (2, ny), (3, nz), (4, ny))}
    if TUNE:
        fc.autotune(img.shape, psf.shape, reps=3, stream=stream)
    variants = fc.selected_variants(img.shape)
    ms = timed_calls(torch, stream, lambda: fc.convolve3d(img, psf, out=out, workspace=ws,
                                                            stream=stream), steps, warm=3)
    out_h = out.cpu().numpy()
    for (p, L), v in keep.items():
        if v >= 0:
            lib.fcamd_select_static_variant(p, L, v)
    checks.add(cfg["name"], lambda O: (out_h, O.numpy_fft_conv(img_np, psf_np)))
    cpu_jobs.append((cfg["name"], img_np, psf_np, cpu_planes, cpu_single))
    _, conv_bytes = pass_bytes(nx, ny, nz, kx, ky, kz)
    N = nx * ny * nz
    log(f"{cfg['name']}: {ms:.4f} ms per convolution")
    return {
        "workload": cfg["name"], "ms_per_conv": ms, "voxels_per_s": N / (ms * 1e-3),
        "conv_roofline_frac": conv_bytes / (ms * 1e-3) / HBM_PEAK, "variants": variants,
    }


def batched_comparison(fc, torch, checks, reps=5):
    """64 uneven menu blocks: batched plan (eager and hipGraph replay) vs the
    plain per-block loop of fcamd_convolve3d_device calls on one stream."""
    from fourierconvolutioncudalib_amd import synth
    from fourierconvolutioncudalib_amd.batched import BatchedConvolution
    spec = synth.uneven_block_shapes(3003, 64)
    g = torch.Generator(device="cuda").manual_seed(3003)
    imgs = [torch.rand(s, device="cuda", generator=g) for s, _, _ in spec]
    kers_np = [synth.random_psf(k, sd) for _, k, sd in spec]
    kers = [torch.from_numpy(k).cuda() for k in kers_np]
    outs = [torch.empty_like(t) for t in imgs]
    ws = torch.empty(max(fc.workspace_bytes(s, k) for s, k, _ in spec), dtype=torch.uint8,
                     device="cuda")
    voxels = sum(int(t.numel()) for t in imgs)
    # algorithmic bytes of the whole batch: the sum over its blocks
    bytes_alg = sum(pass_bytes(s[2], s[1], s[0], k[2], k[1], k[0])[1] for s, k, _ in spec)
    stream = torch.cuda.current_stream()

    def plain():
        for a, k, o in zip(imgs, kers, outs):
            fc.convolve3d(a, k, out=o, workspace=ws, stream=stream)

    plan = BatchedConvolution(imgs, kers, outs=[torch.empty_like(t) for t in imgs])

    def timed(fn):
        for _ in range(2):
            fn()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps, (time.perf_counter() - t0) * 1e3 / reps

    t_plain, w_plain = timed(plain)
    t_eager, w_eager = timed(lambda: plan.run(graph=False))
    plain_h = [o.cpu().numpy() for o in outs]
    eager_h = [o.cpu().numpy() for o in plan.outs]
    t_graph, w_graph = timed(lambda: plan.run(graph=True))
    graph_h = [o.cpu().numpy() for o in plan.outs]
    imgs_h = [t.cpu().numpy() for t in imgs]
    plan.close()

    def check(O):
        pairs = []
        for a, k, p, e, gr in zip(imgs_h, kers_np, plain_h, eager_h, graph_h):
            ref = O.numpy_fft_conv(a, k)
            pairs += [(gr, ref), (e, ref), (p, ref)]
        return pairs

    checks.add("batched_64_blocks (graph, eager and plain outputs of every block)", check)
    log(f"batched: plain {t_plain:.3f} ms, graph {t_graph:.3f} ms")
    return {
        "workload": "64 uneven blocks, extents from {64..256} 7-smooth menu, odd 9-41 kernels",
        "voxels": voxels,
        "plain_loop_ms": t_plain, "batched_eager_ms": t_eager, "batched_graph_ms": t_graph,
        "plain_loop_wall_ms": w_plain, "batched_graph_wall_ms": w_graph,
        "batched_voxels_per_s": voxels / (t_graph * 1e-3),
        "bytes_alg": bytes_alg,
        "conv_roofline_frac_graph": bytes_alg / (t_graph * 1e-3) / HBM_PEAK,
        "conv_roofline_frac_plain": bytes_alg / (t_plain * 1e-3) / HBM_PEAK,
        "speedup_graph_vs_plain": t_plain / t_graph,
        "speedup_eager_vs_plain": t_plain / t_eager,
    }


def adjoint_timing(fc, torch, checks, reps=50):
    """Differentiation workload, 256x256x128 with a 21^3 PSF: forward
    convolution, image gradient (correlation with the PSF) and PSF gradient
    (correlation of the output gradient with the image on the PSF support),
    device-resident, autotuned variants (restored afterwards), HIP events per
    part."""
    from fourierconvolutioncudalib_amd import synth
    from fourierconvolutioncudalib_amd.autograd import grad_kernel, grad_kernel_workspace_bytes
    shape, kshape = (128, 256, 256), (21, 21, 21)
    img_np = synth.uniform(shape, 1003)
    g_np = synth.uniform(shape, 1013)
    psf_np = synth.gaussian_psf(kshape, (2.0, 2.5, 3.0))
    img = torch.from_numpy(img_np).cuda()
    g = torch.from_numpy(g_np).cuda()
    psf = torch.from_numpy(psf_np).cuda()
    out = torch.empty_like(img)
    gi = torch.empty_like(img)
    ws = torch.empty(fc.workspace_bytes(shape, kshape), dtype=torch.uint8, device="cuda")
    wk = torch.empty(grad_kernel_workspace_bytes(shape, kshape), dtype=torch.uint8, device="cuda")
    stream = torch.cuda.current_stream()
    gk = [None]

    def gradk():
        gk[0] = grad_kernel(g, img, kshape, workspace=wk, stream=stream)

    parts = {
        "forward_ms": lambda: fc.convolve3d(img, psf, out=out, workspace=ws, stream=stream),
        "grad_image_ms": lambda: fc.convolve3d(g, psf, op=fc.CORRELATE, out=gi, workspace=ws,
                                               stream=stream),
        "grad_kernel_ms": gradk,
    }
    lib = fc.lib()
    nz, ny, nx = shape
    keep = {(p, L): lib.fcamd_selected_static_variant(p, L)
            for p, L in ((0, nx), (1, nx), (2, ny), (3, nz), (4, ny))}
    if TUNE:
        fc.autotune(shape, kshape, reps=3, stream=stream)
    res = {"workload": "adjoint_256x256x128_k21", "variants": fc.selected_variants(shape)}
    for name, fn in parts.items():
        res[name] = timed_calls(torch, stream, fn, reps, warm=3)
    for (p, L), v in keep.items():
        if v >= 0:
            lib.fcamd_select_static_variant(p, L, v)
    res["step_ms"] = res["forward_ms"] + res["grad_image_ms"] + res["grad_kernel_ms"]
    out_h, gi_h, gk_h = out.cpu().numpy(), gi.cpu().numpy(), gk[0].cpu().numpy()
    checks.add("adjoint forward", lambda O: (out_h, O.numpy_fft_conv(img_np, psf_np)))
    checks.add("adjoint image gradient",
               lambda O: (gi_h, O.numpy_fft_conv(g_np, psf_np, conj_kernel=True)))
    checks.add("adjoint PSF gradient", lambda O: (gk_h, O.fft_conv_grad_kernel(g_np, img_np,
                                                                              kshape)))
    # the forward and the image gradient each move one convolution's
    # algorithmic bytes
    _, cb = pass_bytes(shape[2], shape[1], shape[0], kshape[2], kshape[1], kshape[0])
    res["bytes_alg_per_conv"] = cb
    res["forward_roofline_frac"] = cb / (res["forward_ms"] * 1e-3) / HBM_PEAK
    res["grad_image_roofline_frac"] = cb / (res["grad_image_ms"] * 1e-3) / HBM_PEAK
    log(f"adjoint: step {res['step_ms']:.4f} ms")
    return res


def rl_comparison(fc, torch, checks):
    """Multiview Richardson-Lucy, 4 views of 384x384x192 with 25x25x51 PSFs,
    10 iterations: fused pass sequence vs per-view convolutions."""
    from fourierconvolutioncudalib_amd import synth
    from fourierconvolutioncudalib_amd.rl import richardson_lucy
    cfg = synth.RL_CFG
    _, views, psfs = synth.multiview(cfg["image"], cfg["kernel"], cfg["views"], cfg["seed"])
    tv = [torch.from_numpy(v).cuda() for v in views]
    tk = [torch.from_numpy(k).cuda() for k in psfs]
    res = {"workload": cfg["name"], "iterations": cfg["iterations"]}
    est = {}
    for fused in (True, False):
        richardson_lucy(tv, tk, iterations=1, fused=fused)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        psi = richardson_lucy(tv, tk, iterations=cfg["iterations"], fused=fused)
        e1.record()
        torch.cuda.synchronize()
        res["fused_ms" if fused else "per_view_ms"] = e0.elapsed_time(e1)
        est[fused] = psi.cpu().numpy()
    res["speedup_fused_vs_per_view"] = res["per_view_ms"] / res["fused_ms"]
    n = 1
    for d in cfg["image"]:
        n *= d
    res["view_voxel_updates_per_s"] = cfg["views"] * n * cfg["iterations"] / (res["fused_ms"] * 1e-3)

    def check(O):
        ref = O.fft_rl_multiview(views, psfs, cfg["iterations"])
        return [(est[True], ref), (est[False], ref)]

    checks.add(f"richardson_lucy {cfg['iterations']} iterations (fused and per-view)", check,
               tol=TOL_RL)
    log(f"richardson_lucy: fused {res['fused_ms']:.2f} ms")
    return res


def plan_timing(fc, torch, img, psf, steps, checks, ref_fn):
    """The same convolution through a plan (kernel rows/planes transformed
    once, fcamd_conv_plan_*): milliseconds per call."""
    plan = fc.ConvolutionPlan(img.shape, psf)
    out = torch.empty_like(img)
    ms = timed_calls(torch, torch.cuda.current_stream(), lambda: plan.run(img, out=out), steps,
                     warm=3)
    plan.close()
    out_h = out.cpu().numpy()
    checks.add("plan (metric workload)", lambda O: (out_h, ref_fn(O)))
    return ms


def boundary_comparison(fc, torch, img_np, psf_np, checks, reps=10):
    """Linear convolution with a boundary rule at the workload shape: the
    virtually padded transform size and its time per call."""
    img = torch.from_numpy(img_np).cuda()
    psf = torch.from_numpy(psf_np).cuda()
    out = {}
    ws = torch.empty(fc.boundary_workspace_bytes(img.shape, psf.shape, "reflect"),
                     dtype=torch.uint8, device="cuda")
    res = torch.empty_like(img)
    for mode in ("zero", "reflect"):
        out[mode + "_ms"] = timed_calls(
            torch, torch.cuda.current_stream(),
            lambda: fc.convolve3d_boundary(img, psf, mode, out=res, workspace=ws), reps)
        got = res.cpu().numpy()
        checks.add(f"boundary {mode} (metric workload)",
                   lambda O, got=got, mode=mode: (got, O.fft_boundary_conv(img_np, psf_np, mode)))
    out["padded_zyx"] = list(fc.padded_extents(img.shape, psf.shape, "reflect"))
    return out


# the padded shapes of the reference's numerical-stability tests
# (tests/test_gpu_numerical_stability.cpp:40-96, :330-382): every extent
# carries a large prime factor (79, 109, 11 and 13, 173)
PRIME_SHAPES = [((218, 158, 158), (91, 31, 31)), ((346, 286, 286), (91, 31, 31))]
# shapes of the long-axis path: a prime x whose one-tile plan would be a
# direct 4099-point stage (chirp-z over M = 8232), and x past one tile
# (four-step 125 x 160)
LONG_AXIS_SHAPES = [((8, 64, 4099), (3, 5, 7)), ((4, 16, 20000), (3, 5, 9))]
LONG_AXIS_PATHS = ["x: chirp-z over M = 8232 in one tile", "x: four-step 125 x 160"]


def prime_shapes_timing(fc, torch, checks, reps=10, shapes=PRIME_SHAPES, label="prime shape"):
    """The reference's prime-padded test shapes (or `shapes`), device-resident,
    seeded uniform images with Gaussian PSFs: time per convolution and
    roofline fraction.  158 = 2 * 79 and the z lengths run Bluestein on the
    runtime-length kernels, 286 = 2 * 11 * 13 the compile-time kernels."""
    from fourierconvolutioncudalib_amd import synth
    out = []
    for shape, kshape in shapes:
        img_np = synth.uniform(shape, 1500 + shape[0])
        psf_np = synth.gaussian_psf(kshape, (2.0, 2.0, 6.0))
        img = torch.from_numpy(img_np).cuda()
        psf = torch.from_numpy(psf_np).cuda()
        res = torch.empty_like(img)
        ws = torch.empty(fc.workspace_bytes(shape, kshape), dtype=torch.uint8, device="cuda")
        stream = torch.cuda.current_stream()
        ms = timed_calls(torch, stream, lambda: fc.convolve3d(img, psf, out=res, workspace=ws,
                                                                stream=stream), reps)
        got = res.cpu().numpy()
        name = "x".join(map(str, shape[::-1])) + "_k" + "x".join(map(str, kshape[::-1]))
        checks.add(f"{label} {name}", lambda O, got=got, a=img_np, k=psf_np:
                   (got, O.numpy_fft_conv(a, k)))
        nz, ny, nx = shape
        kz, ky, kx = kshape
        _, cb = pass_bytes(nx, ny, nz, kx, ky, kz)
        out.append({"workload": name, "radix_plans_xyz": [fc.radix_plan(nx), fc.radix_plan(ny),
                                                          fc.radix_plan(nz)],
                    "ms_per_conv": ms, "voxels_per_s": nx * ny * nz / (ms * 1e-3),
                    "conv_roofline_frac": cb / (ms * 1e-3) / HBM_PEAK})
        log(f"{label} {name}: {ms:.3f} ms")
    return out


def long_axis_timing(fc, torch, checks, reps=10):
    """The long-axis path (chirp-z and four-step), timed and checked like
    the prime shapes."""
    out = prime_shapes_timing(fc, torch, checks, reps, LONG_AXIS_SHAPES, "long-axis shape")
    for r, path in zip(out, LONG_AXIS_PATHS):
        del r["radix_plans_xyz"]  # the one-tile plans, which this path does not run
        r["path"] = path
    return out


def load_traffic(workload):
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


# ---------------------------------------------------------------------------
# CPU leg: the reference's host convolution timed, then the FP64 oracle
# checking every registered GPU output.  The only part of the benchmark that
# loads oracle/.
# ---------------------------------------------------------------------------
def cpu_baseline(img, psf, max_planes=None, single=True):
    """Reference host convolution (oracle/_ref, else its C restatement) over
    the workload's output z-planes (all of them, or the first `max_planes`),
    split into one contiguous slab per host thread; with `single`, also one
    thread on one output plane."""
    import numpy as np
    from oracle import oracle as O
    kind = "reference" if O.ref_available() else "port"
    fn = O.convolve_host_reference if kind == "reference" else O.convolve_host
    threads, avail = cpu_threads()
    kz, ky, kx = psf.shape
    hz, hy, hx = kz // 2, ky // 2, kx // 2
    nz_all, ny, nx = img.shape
    nz = nz_all if max_planes is None else min(nz_all, max_planes)
    per = -(-nz // threads)
    padded = np.pad(img, ((hz, hz), (hy, hy), (hx, hx)))
    slabs = []
    for z0 in range(0, nz, per):
        n = min(per, nz - z0)
        slabs.append(np.ascontiguousarray(padded[z0:z0 + n + 2 * hz]))
    outs = [None] * len(slabs)

    def work(i):
        outs[i] = fn(slabs[i], psf, (hz, hy, hx))

    fn(np.ascontiguousarray(padded[:2 * hz + 1, :2 * hy + 8, :2 * hx + 8]), psf, (hz, hy, hx))
    t0 = time.perf_counter()
    ths = [threading.Thread(target=work, args=(i,)) for i in range(len(slabs))]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    dt = time.perf_counter() - t0
    what = (f"the whole {nx}x{ny}x{nz_all} k{kx}x{ky}x{kz} workload (all {nz} output z-planes"
            if nz == nz_all else
            f"{nz} of the {nz_all} output z-planes of the {nx}x{ny}x{nz_all} k{kx}x{ky}x{kz} "
            f"workload (extrapolated per voxel")
    res = {
        "value": nz * ny * nx / dt,
        "unit": "voxels/s",
        "cores": len(slabs),
        "host_cores_available": avail,
        "host_cpu_count": os.cpu_count(),
        "kind": kind,
        "sample": (f"{what}; zero-padded direct convolution, tests/test_algorithms.hpp:9-58), "
                   f"{per} planes per thread on {len(slabs)} threads, {dt:.1f} s"),
    }
    if single:
        t1 = time.perf_counter()
        fn(np.ascontiguousarray(padded[:1 + 2 * hz]), psf, (hz, hy, hx))
        dt1 = time.perf_counter() - t1
        res["single_thread"] = {"value": ny * nx / dt1, "unit": "voxels/s",
                                "sample": f"1 output z-plane on 1 thread, {dt1:.2f} s"}
    return res


def parity(checks):
    """rel-L2 = ||got - ref|| / ||ref|| and max-abs of every registered GPU
    output against its FP64 oracle result (the worst pair for entries with
    several outputs)."""
    from oracle import oracle as O
    out = {}
    ok = True
    for name, fn, tol in checks.items:
        t0 = time.perf_counter()
        pairs = fn(O)
        if isinstance(pairs, tuple):
            pairs = [pairs]
        rel = max(O.rel_l2(a, r) for a, r in pairs)
        mab = max(O.max_abs(a, r) for a, r in pairs)
        good = rel <= tol
        ok = ok and good
        out[name] = {"rel_l2": rel, "max_abs": mab, "tol": tol, "outputs": len(pairs),
                     "ok": good}
        log(f"parity {name}: rel-L2 {rel:.2e} max-abs {mab:.2e} ({time.perf_counter() - t0:.1f} s)")
    out["all_within_tol"] = ok
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true",
                    help="skip the CPU leg (CPU baseline and the FP64 parity checks)")
    ap.add_argument("--no-batched", action="store_true", help="time the metric workload only")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--tuned", metavar="FILE",
                    help="no tuning: apply the variant choices recorded in FILE by an earlier "
                         "run with FCAMD_TUNE_CACHE=FILE (so that a profiled run uses the "
                         "kernels an unprofiled one chose)")
    ap.add_argument("--only", choices=["batched", "rl", "adjoint", "configs", "primes", "longaxis"],
                    help="run one side measurement alone and print its JSON (for profiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(every convolution here runs on a single device)", file=sys.stderr)
        sys.exit(2)
    if args.tuned:
        import fourierconvolutioncudalib_amd as fc
        global TUNE
        TUNE = False
        fc.check(fc.lib().fcamd_autotune_load(args.tuned.encode()))
    if args.only:
        import torch
        import fourierconvolutioncudalib_amd as fc
        from fourierconvolutioncudalib_amd import synth
        ck, jobs = Checks(), []
        sides = {
            "batched": lambda: batched_comparison(fc, torch, ck),
            "rl": lambda: rl_comparison(fc, torch, ck),
            "adjoint": lambda: adjoint_timing(fc, torch, ck),
            "configs": lambda: [config_timing(fc, torch, synth.CFG1, 50, ck, jobs, None, True),
                                config_timing(fc, torch, synth.CFG2, 20, ck, jobs, 16, False)],
            "primes": lambda: prime_shapes_timing(fc, torch, ck),
            "longaxis": lambda: long_axis_timing(fc, torch, ck),
        }
        print(json.dumps({args.only: sides[args.only]()}))
        return

    import numpy as np
    import torch
    import fourierconvolutioncudalib_amd as fc
    from fourierconvolutioncudalib_amd import synth

    cfg = synth.METRIC
    img_np, psf_np = synth.workload(cfg)
    nz, ny, nx = img_np.shape
    kz, ky, kx = psf_np.shape
    dev = torch.device("cuda:0")
    img = torch.from_numpy(img_np).to(dev)
    psf = torch.from_numpy(psf_np).to(dev)
    out = torch.empty_like(img)
    ws = torch.empty(fc.workspace_bytes(img.shape, psf.shape), dtype=torch.uint8, device=dev)
    stream = torch.cuda.current_stream(dev)
    lib = fc.lib()
    checks = Checks()
    cpu_jobs = []

    def step():
        fc.convolve3d(img, psf, out=out, workspace=ws, stream=stream)

    # plan-time tuning of the kernel variants for this shape (not timed)
    if TUNE and not args.no_autotune:
        fc.autotune(img.shape, psf.shape, reps=5, stream=stream)
    variants = fc.selected_variants(img.shape)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    # timed region 1: plain steps
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        step()
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = e0.elapsed_time(e1) / args.steps
    log(f"metric: {ms_step:.4f} ms per step")
    out_h = out.cpu().numpy()
    ref_cache = {}

    def metric_ref(O):
        if "ref" not in ref_cache:
            ref_cache["ref"] = O.numpy_fft_conv(img_np, psf_np)
        return ref_cache["ref"]

    checks.add(cfg["name"], lambda O: (out_h, metric_ref(O)))

    # per-call spread: at least 20 further calls, each between its own pair of
    # events; the median is reported beside the mean of the timed region
    n_med = max(20, args.steps)
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(n_med)]
    for a, b in ev:
        a.record(stream)
        step()
        b.record(stream)
    torch.cuda.synchronize()
    per_call = sorted(a.elapsed_time(b) for a, b in ev)
    ms_median = per_call[n_med // 2] if n_med % 2 else 0.5 * (per_call[n_med // 2 - 1] +
                                                                per_call[n_med // 2])

    # timed region 2: the same steps with HIP events around every pass, on
    # the stream the passes are launched on
    lib.fcamd_pass_timing(1)
    torch.cuda.synchronize()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    lib.fcamd_pass_timing(0)
    ms = (ctypes.c_double * 16)()
    cnt = (ctypes.c_int * 16)()
    npass = lib.fcamd_pass_times(ms, cnt, 16)
    per_pass = {}
    for i in range(npass):
        if cnt[i]:
            per_pass[lib.fcamd_pass_name(i).decode()] = ms[i] / cnt[i]

    pbytes, conv_bytes = pass_bytes(nx, ny, nz, kx, ky, kz)
    dominant = max(per_pass, key=per_pass.get)
    t_dom = per_pass[dominant] * 1e-3
    achieved = pbytes[dominant] / t_dom
    traffic = load_traffic(cfg["name"])
    N = nx * ny * nz
    Nc = (nx // 2 + 1) * ny * nz

    # achievable HBM bandwidth: plain 16-byte streaming copy, 1 GiB each way
    nbytes = 1 << 30
    src = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    dst = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    for _ in range(2):
        fc._lib.check(lib.fcamd_stream_copy_device(dst.data_ptr(), src.data_ptr(), nbytes,
                                                   ctypes.c_void_p(stream.cuda_stream)))
    e0.record(stream)
    for _ in range(10):
        lib.fcamd_stream_copy_device(dst.data_ptr(), src.data_ptr(), nbytes,
                                     ctypes.c_void_p(stream.cuda_stream))
    e1.record(stream)
    torch.cuda.synchronize()
    copy_GBps = 2 * nbytes * 10 / (e0.elapsed_time(e1) * 1e-3) / 1e9
    del src, dst

    # drop-in host-pointer call (PCIe-inclusive), reported beside the value
    im_host = img_np.copy()
    t_host = []
    for _ in range(3):
        im_host[:] = img_np
        t1 = time.perf_counter()
        fc.convolution3DfftCUDAInPlace(im_host.reshape(-1), [nx, ny, nz], psf_np.reshape(-1),
                                       [kx, ky, kz], 0)
        t_host.append(time.perf_counter() - t1)
    checks.add("host-pointer convolution3DfftCUDAInPlace (metric workload)",
               lambda O: (im_host, metric_ref(O)))
    log(f"host-pointer call: {min(t_host) * 1e3:.2f} ms")

    result = {
        "metric": "conv voxels/s + %HBM roofline at 512^2x256 fp32",
        "value": N / (ms_step * 1e-3),
        "unit": "voxels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "ms_per_call_median": ms_median,
        "median_calls": n_med,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (SplitMix64 uniform [0,1) image, Gaussian PSF)",
        "config": {"workload": cfg["name"], "image_xyz": [nx, ny, nz], "kernel_xyz": [kx, ky, kz],
                   "op": "convolve", "parallelism": "single-gpu"},
        "roofline": {
            "bound": "hbm",
            "kernel": dominant,
            "achieved": achieved / 1e9,
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK,
            "traffic": traffic.get(dominant) if traffic else None,
            "bytes_per_launch": pbytes[dominant],
            "launch_ms": per_pass[dominant],
            "bytes_model": "16*Nc for the fused z pass (Nc = (nx/2+1)*ny*nz complex bins: the "
                           "half spectrum read and written once); its reads of the kernel's "
                           "x,y-transformed planes, 8*(nx/2+1)*ny*kz, are not credited",
            "frac_with_kernel_planes": (16 * Nc + 8 * (nx // 2 + 1) * ny * kz) / t_dom / HBM_PEAK
            if dominant == "zfused" else None,
        },
        "conv_roofline": {
            "bytes_alg": conv_bytes,
            "achieved_GBps": conv_bytes / (ms_step * 1e-3) / 1e9,
            "frac": conv_bytes / (ms_step * 1e-3) / HBM_PEAK,
        },
        "hbm_copy_GBps": copy_GBps,
        "variants": variants,
        "pass_ms": {k: round(v, 5) for k, v in per_pass.items()},
        "pass_GBps": {k: round(pbytes[k] / (v * 1e-3) / 1e9, 1) for k, v in per_pass.items()
                      if pbytes[k]},
        "wall_s_timed_region": wall,
        "host_api_ms_pcie_inclusive": min(t_host) * 1e3,
    }
    result["plan_ms_per_step"] = plan_timing(fc, torch, img, psf, args.steps, checks, metric_ref)
    if not args.no_batched:
        result["boundary_modes"] = boundary_comparison(fc, torch, img_np, psf_np, checks)
        result["batched"] = batched_comparison(fc, torch, checks)
        result["richardson_lucy"] = rl_comparison(fc, torch, checks)
        result["adjoint"] = adjoint_timing(fc, torch, checks)
        result["configs"] = [
            config_timing(fc, torch, synth.CFG1, 50, checks, cpu_jobs, None, True),
            config_timing(fc, torch, synth.CFG2, 20, checks, cpu_jobs, 16, False)]
        result["reference_prime_shapes"] = prime_shapes_timing(fc, torch, checks)
        result["long_axis_shapes"] = long_axis_timing(fc, torch, checks)
    del img, psf, out, ws
    torch.cuda.empty_cache()

    # CPU leg (after every GPU timing)
    if not args.no_cpu_baseline:
        log("CPU leg: reference host convolution")
        result["cpu_baseline"] = cpu_baseline(img_np, psf_np)
        for c, (name, a, k, planes, single) in zip(result.get("configs", []), cpu_jobs):
            c["cpu_baseline"] = cpu_baseline(a, k, planes, single)
        log("CPU leg: FP64 oracle checks")
        result["parity"] = parity(checks)
    else:
        result["cpu_baseline"] = None
        result["parity"] = None
    print(json.dumps(result))


if __name__ == "__main__":
    main()
