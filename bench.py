"""Benchmark: INSTA per-sample hot path on one MI355X.

Workload (the full-size FP32 configuration): a 5,023-vertex / 9,976-triangle
head mesh with 16 posed frames (device-built BVHs), hash grid L=16 F=2
T=2^19 (N_min 16 -> N_max 2048), density MLP 32->64->16, colour MLP
32->64->64->3, 128^3 occupancy bitfield per frame, 512x512 pinhole render
with at most 1,024 steps per ray.  One step = one 512x512 frame rendered
end to end (ray init, march, BVH deformation, encoding + MLPs, compositing),
frames cycling through the 16 posed frames.  Inputs (meshes, parameters,
occupancy) are resident in HBM before timing starts.

Prints one JSON line: deformed samples/s (samples composited per second) as
`value`, frames/s beside it, the dominant kernel's roofline position
(measured with HIP events on the render stream), and the CPU oracle's
throughput on a bounded pixel subset of the same frame.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
FP32_PEAK_TFLOPS = 157.3       # MI355X FP32 matrix/vector peak
METRIC = "deformed samples/s and 512^2 frames/s, FP16-MFMA vs FP32; serving p50/p99 ms"
# kernel class -> kernels of that class in the rocprofv3 summaries
PROFILE_KERNELS = {"deform": "deform_", "network": "network_f32_kernel", "march": "march_kernel",
                   "composite": "composite_kernel"}
PROFILE_NAMES = {"deform": ["deform_hits_kernel", "deform_apply_kernel"], "march": ["march_kernel"],
                 "composite": ["composite_kernel"]}


def per_sample_costs(cfg, lay):
    """Algorithmic work per sample (unpadded) for each kernel class."""
    L, F = cfg.n_levels, cfg.n_features_per_level
    w = cfg.n_neurons
    flops = 2 * (lay.density_in * w + (cfg.density_hidden_layers - 1) * w * w + w * cfg.density_out)
    flops += 2 * (lay.rgb_in * w + (cfg.rgb_hidden_layers - 1) * w * w + w * 3)
    return {
        # corner gathers L*8*F*4 B + canonical pos/dir in (32 B) + 4 outputs (16 B)
        "network": dict(bytes=L * 8 * F * 4 + 32 + 16, flops=flops),
        # posed pos + ray id in (16 B), ray direction (16 B), canonical pos/dir out (32 B)
        "deform": dict(bytes=64, flops=0),
        # network output + position per sample read (32 B)
        "composite": dict(bytes=32, flops=0),
        # emitted sample written (16 B)
        "march": dict(bytes=16, flops=0),
    }


def setup(args, precision=0, av=None, n_extra=0):
    import torch
    from insta_amd import synth
    from insta_amd.capi import Context, full_config, make_camera
    kw = {} if args.cell_split is None else dict(cell_split=args.cell_split)
    if args.cell_list_entries:
        kw["cell_list_entries"] = args.cell_list_entries
    cfg = full_config(max_frames=16, precision=precision, n_extra_dims=n_extra, **kw)
    if av is None:
        av = synth.head_avatar(seed=1, n_frames=16)
    ctx = Context(cfg)
    ctx.set_canonical_mesh(av.canonical, av.faces, av.masks)
    frames = torch.from_numpy(np.stack(av.frames)).cuda()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ctx.set_frames(0, frames)   # 16 device BVH builds (with FP64 frames), one workgroup each
    torch.cuda.synchronize()
    ctx.bvh_ms_per_frame = (time.perf_counter() - t0) * 1e3 / 16
    params = synth.network_params(7, ctx.layout, cfg)
    ctx.set_params(torch.from_numpy(params).cuda())
    if n_extra:  # per-frame expression values (seeded N(0, 1))
        vals = np.random.default_rng(77).standard_normal((16, n_extra)).astype(np.float32)
        ctx.set_frame_extra_dims(0, torch.from_numpy(vals).cuda())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for f in range(16):
        ctx.build_occupancy(f, 0.015)
    torch.cuda.synchronize()
    setup_ms = (time.perf_counter() - t0) * 1e3 / 16
    rng = np.random.default_rng(2024)
    cams = [make_camera(synth.random_camera(rng, 0.7), synth.focal_from_fov(args.res, 30.0)) for _ in range(16)]
    torch.cuda.synchronize()
    ctx.occupancy_ms_per_frame = setup_ms
    return ctx, cfg, av, params, cams


def phase(name):
    """One stderr line per finished phase (the JSON result goes to stdout)."""
    print(f"bench: {name} done", file=sys.stderr, flush=True)


def timed_renders(ctx, cams, res, steps, warmup, out):
    """Render `steps` frames (after `warmup`), cycling the 16 posed frames.
    The timed pass runs with profiling off and reads nothing back between
    frames (each insta_render still waits for its own frame's counters, as
    the API returns a finished image).  A second, untimed pass over the same
    frames (renders are deterministic: same camera, frame and sample index)
    collects the sample counts and the per-kernel HIP-event breakdown.
    Returns (seconds, samples composited, samples emitted, kernel profile)."""
    import torch
    for s in range(warmup):
        ctx.render(cams[s % 16], res, res, s % 16, s % 16, 0, out)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(steps):
        f = s % 16
        ctx.render(cams[f], res, res, f, f, 0, out)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ctx.set_profiling(True)
    samples = emitted = rounds = rays0 = 0
    for s in range(steps):
        f = s % 16
        ctx.render(cams[f], res, res, f, f, 0, out)
        st = ctx.render_stats()
        samples += st.n_samples
        emitted += st.n_samples_emitted
        rounds += st.n_iterations
        rays0 += st.n_rays_alive0
    torch.cuda.synchronize()
    prof = ctx.profile()
    prof["_rounds_per_frame"] = rounds / steps
    prof["_live_rays_per_frame"] = rays0 / steps
    ctx.set_profiling(False)
    return elapsed, samples, emitted, prof


def accuracy_vs_fp64(ctxs, cfg, av, params, cam, res, n_pix):
    """Per-pixel error of each precision's render against the FP64 oracle on
    a fixed random subset of pixels of frame 0."""
    from oracle.oracle import Oracle, render_fp64
    orc = Oracle(cfg)
    orc.set_canonical_mesh(av.canonical, av.faces, av.masks)
    orc.set_frame(0, av.frames[0])
    orc.set_params(params)
    orc.build_occupancy(0, 0.015)
    pix = np.random.default_rng(11).choice(res * res, n_pix, replace=False).astype(np.uint32)
    ref = render_fp64(orc, cam, res, res, 0, 0, 0, pix)
    out = {}
    for name, ctx in ctxs.items():
        rgba, _, _ = ctx.render(cam, res, res, 0, 0, 0)
        got = rgba.cpu().numpy().astype(np.float64)[pix]
        err = got[:, :3] - ref[:, :3]
        mse = float(np.mean(err ** 2))
        out[name] = dict(max_abs=float(np.abs(err).max()), rms=math.sqrt(mse),
                         psnr_db=(float("inf") if mse == 0 else 10 * math.log10(1.0 / mse)))
    out["pixels"] = n_pix
    return out


def cpu_baseline(cfg, av, params, cam, res, budget_s, runs=5):
    """The oracle on the host cores (OpenMP over rays): one warm-up run, then
    the median of `runs` timed runs of the same work, on a bounded pixel
    subset of the 512x512 frame of posed frame 0 (a fixed random subset of
    pixels, sized from the warm-up so the timed runs take about budget_s)."""
    from oracle.oracle import Oracle
    orc = Oracle(cfg)
    orc.set_canonical_mesh(av.canonical, av.faces, av.masks)
    orc.set_frame(0, av.frames[0])
    orc.set_params(params)
    orc.build_occupancy(0, 0.015)
    order = np.random.default_rng(5).permutation(res * res).astype(np.uint32)
    n_pix = min(res * res, 16384)
    t0 = time.perf_counter()
    orc.render(cam, res, res, 0, 0, 0, pixels=order[:n_pix])       # warm-up
    warm = time.perf_counter() - t0
    n_pix = int(min(res * res, max(1024, n_pix * (budget_s / runs) / max(warm, 1e-3))))
    pix = order[:n_pix]
    times, samples = [], 0
    for _ in range(runs):
        t0 = time.perf_counter()
        samples = int(orc.render(cam, res, res, 0, 0, 0, pixels=pix)[3][0])
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return dict(value=samples / med, unit="samples/s", cores=threads, kind="port",
                sample=f"{n_pix} random pixels of a {res}x{res} frame of posed frame 0 ({samples} samples); "
                       f"1 warm-up, median of {runs} runs ({', '.join(f'{t:.2f}' for t in times)} s)",
                frames_per_s=(samples / med) / (samples * res * res / n_pix))


def small_scalar_bench(runs=5):
    """The small CPU-scalar workload (2,048-triangle sphere, one posed frame,
    L=8 F=2 T=2^14, 64^3 grid at distance 0.05, a 64x64 render from a camera
    at radius 1.0 with a 30 degree field of view, <= 256 steps of 1/256):
    the oracle on one host thread and on all of them, and the library on the
    GPU (FP32), each 1 warm-up and the median of `runs`; the GPU image
    against the oracle's (largest per-pixel RGBA difference)."""
    import torch
    from insta_amd import synth
    from insta_amd.capi import Context, make_camera, small_config
    from oracle.oracle import Oracle
    cfg = small_config(max_frames=1)
    av = synth.small_sphere_avatar(seed=0, n_frames=1)
    res = 64
    cam = make_camera(synth.random_camera(np.random.default_rng(5), 1.0), synth.focal_from_fov(res, 30.0))
    ctx = Context(cfg)
    ctx.set_canonical_mesh(av.canonical, av.faces, av.masks)
    ctx.set_frames(0, torch.from_numpy(av.frames[0][None].astype(np.float32)).cuda())
    params = synth.network_params(7, ctx.layout, cfg)
    ctx.set_params(torch.from_numpy(params).cuda())
    ctx.build_occupancy(0, 0.05)
    orc = Oracle(cfg)
    orc.set_canonical_mesh(av.canonical, av.faces, av.masks)
    orc.set_frame(0, av.frames[0])
    orc.set_params(params)
    orc.build_occupancy(0, 0.05)

    def median_time(fn):
        fn()
        ts = []
        for _ in range(runs):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts))

    out = {}
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    ref = None
    for name, nt in (("cpu_1_thread", 1), ("cpu_all_threads", threads)):
        orc.set_threads(nt)
        ref = orc.render(cam, res, res, 0, 0)
        samples = int(ref[3][0])
        t = median_time(lambda: orc.render(cam, res, res, 0, 0))
        out[name] = dict(threads=nt, samples_per_s=samples / t, ms_per_frame=t * 1e3)
    rgba = None

    def gpu():
        nonlocal rgba
        rgba = ctx.render(cam, res, res, 0, 0)[0]
        torch.cuda.synchronize()
    t = median_time(gpu)
    out["gpu_fp32"] = dict(samples_per_s=int(ref[3][0]) / t, ms_per_frame=t * 1e3)
    out["max_abs_rgba_vs_oracle"] = float(np.abs(rgba.cpu().numpy() - ref[0]).max())
    out["samples"] = int(ref[3][0])
    out["workload"] = ("small CPU-scalar workload: 2,048-tri sphere, L8 T2^14, 64^3 grid, 64x64 render, "
                       "<= 256 steps of 1/256; 1 warm-up, median of %d" % runs)
    return out


def precision_bench(args, n_views=8, res=64, launch=1 << 20, reps=20):
    """Precision form at the small shapes (2,048-triangle sphere, L=8 F=2
    T=2^14, 64^3 grid, 64x64 images, <= 256 steps of 1/256) with
    trained-scale hash tables (U(-0.5, 0.5); Xavier MLPs), so that the colour
    depends on the encoding and the arithmetic shows: per-pixel RGB error of
    each network arithmetic against the FP64 oracle render over `n_views`
    random cameras (max, RMS, PSNR), and the network alone on batches of
    2^20 samples per launch (HIP events on the library's stream): time,
    achieved TFLOP/s on the algorithmic FLOP and its fraction of the peak of
    the units it runs on (MFMA FP32 157.3 TFLOP/s, FP16/BF16 2.5 PFLOP/s;
    the per-sample FP32 path runs on the vector ALUs, same FP32 peak)."""
    import torch
    from insta_amd import synth
    from insta_amd.capi import PRECISION_NAMES, Context, make_camera, small_config
    from oracle.oracle import Oracle, render_fp64
    av = synth.small_sphere_avatar(seed=0, n_frames=1)
    base = small_config(max_frames=1)
    orc = Oracle(base)
    lay = orc.layout
    params = synth.network_params(7, lay, base)
    params[lay.n_mlp_params:] = np.random.default_rng(31).uniform(-0.5, 0.5, lay.n_grid_params).astype(np.float32)
    orc.set_canonical_mesh(av.canonical, av.faces, av.masks)
    orc.set_frame(0, av.frames[0])
    orc.set_params(params)
    orc.build_occupancy(0, 0.05)
    rng = np.random.default_rng(32)
    cams = [make_camera(synth.random_camera(rng, 1.0, front_bias=False), synth.focal_from_fov(res, 30.0))
            for _ in range(n_views)]
    refs = [render_fp64(orc, cam, res, res, 0, 0, 0)[:, :3] for cam in cams]
    # 2^20 network inputs: canonical positions near the surface, random directions
    tri = av.canonical[av.faces].astype(np.float64)
    k = rng.integers(0, len(av.faces), launch)
    pos = np.einsum("nk,nkd->nd", rng.dirichlet([1, 1, 1], launch), tri[k]) + 0.03 * rng.standard_normal((launch, 3))
    d = rng.standard_normal((launch, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    coords = np.zeros((launch, 11), np.float32)
    coords[:, 0:3] = np.clip(pos, 0.0, 1.0)
    coords[:, 8:11] = (d + 1.0) * 0.5
    coords_dev = torch.from_numpy(coords).cuda()
    flops = per_sample_costs(base, lay)["network"]["flops"]
    out = {"workload": f"small shapes, {n_views} random {res}x{res} views, hash tables U(-0.5, 0.5), "
                       f"network batches of {launch} samples"}
    for prec in (0, 3, 1, 2):
        name = PRECISION_NAMES[prec]
        cfg = small_config(max_frames=1, precision=prec)
        ctx = Context(cfg)
        ctx.set_canonical_mesh(av.canonical, av.faces, av.masks)
        ctx.set_frames(0, torch.from_numpy(av.frames[0][None]).cuda())
        ctx.set_params(torch.from_numpy(params).cuda())
        ctx.build_occupancy(0, 0.05)
        errs, imgs = [], []
        for cam, ref in zip(cams, refs):
            rgba, _, _ = ctx.render(cam, res, res, 0, 0, 0)
            img = rgba.cpu().numpy()[:, :3].astype(np.float64)
            imgs.append(img)
            errs.append(img - ref)
        e = np.concatenate(errs)
        mse = float(np.mean(e ** 2))
        if prec == 0:
            fp32_imgs = imgs
        e32 = np.concatenate([a - b for a, b in zip(imgs, fp32_imgs)])
        mse32 = float(np.mean(e32 ** 2))
        stream = torch.cuda.current_stream()
        for _ in range(3):
            ctx.network_inference(coords_dev)
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        for _ in range(reps):
            ctx.network_inference(coords_dev)
        ev1.record(stream)
        torch.cuda.synchronize()
        ms = ev0.elapsed_time(ev1) / reps
        peak = 2500.0 if prec in (1, 2) else FP32_PEAK_TFLOPS
        tflops = flops * launch / (ms * 1e-3) / 1e12
        out[name] = dict(max_abs=float(np.abs(e).max()), rms=math.sqrt(mse),
                         psnr_db=(float("inf") if mse == 0 else 10 * math.log10(1.0 / mse)),
                         vs_fp32_render=dict(max_abs=float(np.abs(e32).max()), rms=math.sqrt(mse32),
                                             psnr_db=(float("inf") if mse32 == 0 else 10 * math.log10(1.0 / mse32))),
                         network_ms_per_launch=ms, samples_per_s=launch / (ms * 1e-3), tflops=tflops,
                         peak_tflops=peak, utilisation=tflops / peak,
                         units="matrix cores" if prec != 3 else "vector ALUs")
        del ctx
    return out


def training_bench(args, full=False):
    """Training step at the small shapes (L=8, T=2^14, 64^3 grid, 2,048-triangle
    sphere), or with full=True at the full-size shapes (L=16, T=2^19, 128^3
    grid, 9,976-triangle head, <= 1,024 steps): rays uniform over 64 posed
    frames seen by 64 random 512x512 cameras with smooth-noise targets, a
    2^18-sample training batch per step (adaptive ray count; marched samples
    are evaluated, the composited ones compacted into the batch), Adam.  FP32
    and FP16 grid-gradient accumulation (loss scale 128); gradient error of one
    batch against the FP64 oracle (4,096 rays; 1,024 at full size)."""
    import torch
    from insta_amd import synth
    from insta_amd.capi import Context, default_train_config, full_config, make_camera, small_config
    n_frames, res = 64, 512
    if full:
        cfg = full_config(max_frames=n_frames)
        av = synth.head_avatar(seed=0, n_frames=n_frames)
        radius, thr, n_check = 0.7, 0.015, 1024
    else:
        cfg = small_config(max_frames=n_frames)
        av = synth.small_sphere_avatar(seed=0, n_frames=n_frames)
        radius, thr, n_check = 1.0, 0.05, 4096
    rng = np.random.default_rng(21)
    cams = [make_camera(synth.random_camera(rng, radius, front_bias=False), synth.focal_from_fov(res, 30.0))
            for _ in range(n_frames)]
    images = synth.smooth_noise_images(5, n_frames, res, res)
    images_dev = torch.from_numpy(images).cuda()
    frames_dev = torch.from_numpy(np.stack(av.frames)).cuda()

    def make_ctx():
        ctx = Context(cfg)
        ctx.set_canonical_mesh(av.canonical, av.faces, av.masks)
        ctx.set_frames(0, frames_dev)
        params = synth.network_params(7, ctx.layout, cfg)
        ctx.set_params(torch.from_numpy(params).cuda())
        for f in range(n_frames):
            ctx.build_occupancy(f, thr)
        ctx.set_training_set(images_dev, cams, list(range(n_frames)))
        return ctx, params

    out = {}
    target = 1 << 18
    for name, kw in (("fp32_grid_grads", {}), ("fp16_grid_grads", dict(grid_grad_fp16=1, loss_scale=128.0))):
        ctx, _ = make_ctx()
        # marched samples per step are not capped below what 2^18 composited
        # samples need (rays stop within a few samples once trained)
        tc = default_train_config(seed=3, max_samples=1 << 24, target_samples=target, **kw)
        first = None
        # warm up until the adaptive ray count holds the batch at the target
        # (three steps in a row within 2 %), at least train_warmup steps
        held, warm = 0, 0
        while warm < args.train_warmup or (held < 3 and warm < args.train_warmup_max):
            st = ctx.train_step(tc)
            first = st.loss if first is None else first
            held = held + 1 if min(st.n_samples_used, target) >= 0.98 * target else 0
            warm += 1
        torch.cuda.synchronize()
        samples = rays = batch = evaluated = 0
        t0 = time.perf_counter()
        for _ in range(args.train_steps):
            st = ctx.train_step(tc)
            samples += st.n_samples
            evaluated += st.n_samples_evaluated
            rays += st.n_rays
            batch += min(st.n_samples_used, tc.target_samples)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        k = args.train_steps
        out[name] = dict(steps_per_s=k / dt, ms_per_step=dt / k * 1e3, batch_samples_per_step=batch / k,
                         marched_samples_per_step=samples / k, evaluated_samples_per_step=evaluated / k,
                         rays_per_step=rays / k,
                         batch_samples_per_s=batch / dt, loss_first=first, loss_last=st.loss, warmup_steps=warm)
        del ctx

    # gradient error of one batch against the FP64 oracle (same rays and samples;
    # a fresh context per variant so both see the batch of step 0)
    from oracle import oracle as O
    n_rays = n_check
    grads = {}
    for name, kw in (("fp32", {}), ("fp16_grid", dict(grid_grad_fp16=1, loss_scale=128.0))):
        ctx, params = make_ctx()
        tc = default_train_config(n_rays=n_rays, max_samples=n_rays * cfg.max_steps,
                                  target_samples=n_rays * cfg.max_steps, apply_update=0, seed=17, **kw)
        ctx.train_step(tc)
        grads[name] = ctx.gradients().cpu().numpy().astype(np.float64)
        del ctx
    orc = O.Oracle(cfg)
    orc.set_canonical_mesh(av.canonical, av.faces, av.masks)
    for f in range(n_frames):
        orc.set_frame(f, av.frames[f])
        orc.build_occupancy(f, thr)
    orc.set_params(params)
    O.set_training_set(orc, images, cams, list(range(n_frames)))
    g64, _, counts = O.train_grads(orc, n_rays, 17)
    lay = orc.layout
    b = np.cumsum([0] + [o * i for (o, i) in lay.mlp_shapes(cfg)] + [lay.n_grid_params])
    names = ["density_W0", "density_W1", "colour_W0", "colour_W1", "colour_W2", "hash_tables"]
    rel = lambda a, r: float(np.linalg.norm(a - r) / (np.linalg.norm(r) + 1e-30))
    out["grad_rel_l2_vs_fp64"] = {"batch": f"{n_rays} rays, {counts[1]} samples"}
    for name, g in grads.items():
        out["grad_rel_l2_vs_fp64"][name] = {n: rel(g[b0:b1], g64[b0:b1]) for n, b0, b1 in zip(names, b[:-1], b[1:])}
    shapes = ("full-size shapes (L16 T2^19, 128^3 grid, 9,976-tri head, <=1024 steps)" if full
              else "small shapes (L8 T2^14, 64^3 grid, 2,048-tri sphere)")
    out["workload"] = (f"{shapes}, 64 posed frames x 512x512 random cameras, a batch of up to 2^18 composited "
                       "samples per step (adaptive ray count, at most 2^18 rays; timed after 60 warm-up steps, "
                       "once it holds the batch; the rays per step keep growing as training terminates rays "
                       "earlier), "
                       "Adam(0.9, 0.99, lr 1e-2)")
    return out


def serving_bench(args, full=False):
    """Serving form: 8 resident avatars at the small shapes, or with full=True
    at the full-size shapes (9,976-triangle heads, L16 T2^19 tables, 128^3
    grids; different seeds for mesh, bases, MLP and tables), FP16 MFMA network, Poisson arrivals at several
    rates; each request = avatar, 50 expression coefficients (device BVH
    rebuild + occupancy), random camera, resolution 128/256/512/1024^2 with
    weights 0.3/0.3/0.3/0.1.  Against sequential plain-FP32 rendering: the
    same requests' service times on the FP32 per-sample (scalar) path, queued
    one at a time under the same arrivals."""
    import torch
    from insta_amd import synth
    from insta_amd.capi import (PRECISION_FP16, PRECISION_FP32_SCALAR, Context, Server, default_server_config,
                                full_config, make_camera, small_config)
    n_av, n_slots = 8, args.serve_slots
    radius, thr = (0.7, 0.015) if full else (1.0, 0.05)
    resolutions, weights = np.array([128, 256, 512, 1024]), np.array([0.3, 0.3, 0.3, 0.1])

    kw = dict(cell_split=args.serve_cell_split)

    def avatars(precision):
        out = []
        for a in range(n_av):
            if full:
                av = synth.head_avatar(seed=300 + a, n_frames=0, n_bases=50)
                cfg = full_config(max_frames=max(n_slots, 1), precision=precision, **kw)
            else:
                av = synth.small_sphere_avatar(seed=300 + a, n_frames=0, n_bases=50)
                cfg = small_config(max_frames=max(n_slots, 1), precision=precision, **kw)
            ctx = Context(cfg)
            ctx.set_canonical_mesh(av.canonical, av.faces, av.masks)
            ctx.set_params(torch.from_numpy(synth.network_params(400 + a, ctx.layout, cfg)).cuda())
            out.append((ctx, torch.from_numpy(av.bases).cuda()))
        torch.cuda.synchronize()
        return out

    def requests(seed, n):
        rng = np.random.default_rng(seed)
        reqs = []
        for _ in range(n):
            res = int(rng.choice(resolutions, p=weights))
            cam = make_camera(synth.random_camera(rng, radius, front_bias=False), synth.focal_from_fov(res, 30.0))
            reqs.append((int(rng.integers(n_av)), rng.normal(0.0, 1.0, 50).astype(np.float32), cam, res))
        return reqs

    def pct(x, q):
        return float(np.percentile(np.asarray(x) * 1e3, q)) if len(x) else None

    fp16 = avatars(PRECISION_FP16)
    server = Server(default_server_config(n_slots=n_slots, hbm_budget_bytes=args.serve_budget_gib << 30,
                                          occupancy_threshold=thr))
    ids = [server.add_avatar(ctx, bases) for ctx, bases in fp16]
    outs = {r: [torch.empty((r * r, 4), dtype=torch.float32, device="cuda") for _ in range(n_slots * 4)]
            for r in resolutions.tolist()}
    # warm-up: graphs captured, workspaces touched
    for k, (a, c, cam, res) in enumerate(requests(1, 2 * n_slots)):
        server.submit(ids[a], c, cam, res, res, outs[res][k % len(outs[res])])
    while server.poll(wait=True):
        pass
    out = {"avatars": n_av, "slots": n_slots, "budget_gib": args.serve_budget_gib, "rates": {}}
    for rate in args.serve_rates:
        reqs = requests(int(rate) + 7, args.serve_requests)
        arrivals = np.cumsum(np.random.default_rng(int(rate)).exponential(1.0 / rate, len(reqs)))
        t0 = server.clock() + 0.05
        results, k = [], 0
        while len(results) < len(reqs):
            now = server.clock()
            while k < len(reqs) and now >= t0 + arrivals[k]:
                a, c, cam, res = reqs[k]
                server.submit(ids[a], c, cam, res, res, outs[res][k % len(outs[res])], tag=k)
                k += 1
            done = server.poll(wait=(k == len(reqs)))
            results += done
            if not done and k < len(reqs):
                time.sleep(min(0.0005, max(0.0, t0 + arrivals[k] - server.clock())))
        lat = {r.user_tag: r.t_done - r.t_submit for r in results}
        by_res = {int(res): [lat[i] for i, q in enumerate(reqs) if q[3] == res] for res in resolutions.tolist()}
        span = max(r.t_done for r in results) - min(r.t_submit for r in results)
        out["rates"][str(rate)] = dict(
            requests=len(reqs), frames_per_s=len(reqs) / span, p50_ms=pct(list(lat.values()), 50),
            p99_ms=pct(list(lat.values()), 99),
            p99_ms_by_res={str(r): pct(v, 99) for r, v in by_res.items()},
            samples_per_s=sum(r.n_samples for r in results) / span)
    server.close()

    # sequential plain FP32: service time of each request alone on the FP32 per-sample path
    fp32 = avatars(PRECISION_FP32_SCALAR)
    frame = 0
    svc = {}
    for rate in args.serve_rates:
        reqs = requests(int(rate) + 7, args.serve_requests)
        times = []
        for (a, c, cam, res) in reqs:
            ctx, bases = fp32[a]
            torch.cuda.synchronize()
            t = time.perf_counter()
            ctx.set_frame_expression(frame, bases, torch.from_numpy(c).cuda())
            ctx.build_occupancy(frame, thr)
            ctx.render(cam, res, res, frame, frame, 0, (outs[res][0], None, None))
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t)
        arrivals = np.cumsum(np.random.default_rng(int(rate)).exponential(1.0 / rate, len(reqs)))
        finish, lat = 0.0, []
        for t_arr, s in zip(arrivals, times):
            finish = max(finish, t_arr) + s
            lat.append(finish - t_arr)
        svc[str(rate)] = dict(frames_per_s=len(reqs) / (finish - arrivals[0]), p50_ms=pct(lat, 50),
                              p99_ms=pct(lat, 99), mean_service_ms=float(np.mean(times) * 1e3))
    out["sequential_fp32"] = svc
    out["workload"] = ("8 avatars at the " + ("full-size shapes (9,976-tri heads, L16 T2^19, 128^3 grid)" if full
                       else "small shapes (2,048-tri spheres, L8 T2^14, 64^3 grid)") +
                       ", resolutions 128/256/512/1024^2 (0.3/0.3/0.3/0.1), Poisson arrivals, candidate lists "
                       f"split {args.serve_cell_split}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=32)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--res", type=int, default=512)
    ap.add_argument("--cpu-budget", type=float, default=10.0, help="seconds of CPU oracle timing (0: skip)")
    ap.add_argument("--skip-forms", action="store_true", help="only the FP32 path")
    ap.add_argument("--accuracy-pixels", type=int, default=4096, help="pixels compared with the FP64 oracle")
    ap.add_argument("--skip-train", action="store_true", help="skip the training-step measurement")
    ap.add_argument("--skip-serve", action="store_true", help="skip the serving measurement")
    ap.add_argument("--serve-rates", type=float, nargs="+", default=[20.0, 50.0, 100.0, 200.0])
    ap.add_argument("--serve-requests", type=int, default=120)
    ap.add_argument("--serve-slots", type=int, default=6)
    ap.add_argument("--serve-budget-gib", type=int, default=24)
    ap.add_argument("--serve-sizes", choices=["both", "small", "full"], default="both")
    ap.add_argument("--serve-cell-split", type=int, default=2,
                    help="candidate-list split of the served avatars' contexts (each request builds its "
                         "frame's lists for one render; 2 measured best for both avatar sizes)")
    ap.add_argument("--train-steps", type=int, default=40)
    ap.add_argument("--train-sizes", choices=["both", "small", "full"], default="both")
    ap.add_argument("--train-warmup", type=int, default=60,
                    help="training steps before the timed ones (the adaptive ray count then holds the batch "
                         "at 2^18 samples)")
    ap.add_argument("--train-warmup-max", type=int, default=120,
                    help="warm-up steps allowed for the batch to reach 2^18 samples")
    ap.add_argument("--cell-split", type=int, default=None,
                    help="candidate-list split of the render contexts (default: the library's)")
    ap.add_argument("--cell-list-entries", type=int, default=0, help="list entries per frame (0: the library's)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run one 256 MiB device copy first (reference kernel for rocprof byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (the library runs on one "
              "GPU; the path would shard by image tiles for rendering and by avatars for serving)", file=sys.stderr)
        sys.exit(2)

    import torch
    from insta_amd.capi import PRECISION_NAMES
    if args.calibrate_copy:
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    res = args.res
    out = (torch.empty((res * res, 4), dtype=torch.float32, device="cuda"),
           torch.empty(res * res, dtype=torch.float32, device="cuda"),
           torch.empty(res * res, dtype=torch.int32, device="cuda"))

    # headline: the FP32 path (FP32 MFMA network)
    ctx, cfg, av, params, cams = setup(args, 0)
    elapsed, samples, emitted, prof = timed_renders(ctx, cams, res, args.steps, args.warmup, out)
    rounds, live_rays = prof.pop("_rounds_per_frame"), prof.pop("_live_rays_per_frame")
    value = samples / elapsed
    costs = per_sample_costs(cfg, ctx.layout)
    dom = max((k for k in prof if prof[k]["launches"]), key=lambda k: prof[k]["ms"])
    p = prof[dom]
    avg_s = p["ms"] / p["launches"] / 1e3
    per_launch_items = p["items"] / p["launches"]
    c = costs.get(dom, dict(bytes=0, flops=0))
    # HBM bytes per launch of the same kernels from the committed rocprofv3
    # FETCH_SIZE / WRITE_SIZE summary of this workload (scaled by a 256 MiB copy)
    traffic = None
    prof_json = ROOT / "profiles" / "render_full_512.json"
    if prof_json.exists():
        kn = json.loads(prof_json.read_text())["kernels"]
        # (the "<true>" build is the traversal-counting one of an untimed frame)
        names = [k for k in kn if k.startswith(PROFILE_KERNELS.get(dom, "-")) and "<true>" not in k]
        if names:
            traffic = sum(kn[k]["fetch_bytes"] + kn[k]["write_bytes"] for k in names)
    if dom == "network":
        achieved = c["flops"] * per_launch_items / avg_s / 1e12
        roof = dict(bound="mfma", achieved=achieved, peak=FP32_PEAK_TFLOPS, unit="TFLOP/s",
                    frac=achieved / FP32_PEAK_TFLOPS, traffic=traffic, kernel="network_f32_kernel<16>",
                    per_unit=f"{c['flops']} FLOP/sample (unpadded), {c['bytes']} B/sample gathered+streamed")
    else:
        achieved = c["bytes"] * per_launch_items / avg_s / 1e9
        roof = dict(bound="hbm", achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s", frac=achieved / HBM_PEAK_GBS,
                    traffic=traffic, kernel=f"{dom} (" + " + ".join(PROFILE_NAMES.get(dom, [dom])) + ")",
                    per_unit=f"{c['bytes']} B/sample streamed", traffic_source="profiles/render_full_512.json")
    kernels = {k: dict(ms_per_frame=v["ms"] / args.steps, launches=v["launches"], items=v["items"])
               for k, v in prof.items() if v["launches"]}
    # nearest-triangle work per deformed sample: one more (untimed) frame with
    # the counting build of the deformation kernel
    ctx.set_profiling(True, count_traversal=True)
    ctx.render(cams[0], res, res, 0, 0, 0, out)
    tc = ctx.traversal_counts()
    ctx.set_profiling(False)
    q = max(tc["queries"], 1)
    traversal = dict(queries=tc["queries"], list_queries=tc["list_queries"], cell_split=cfg.cell_split,
                     list_entries_per_sample=tc["list_entries"] / q,
                     inner_nodes_per_sample=tc["inner"] / q,
                     child_boxes_per_sample=4 * tc["inner"] / q, leaves_per_sample=tc["leaves"] / q,
                     triangles_per_sample=tc["triangles"] / q,
                     nodes_per_sample=(tc["inner"] + tc["leaves"]) / q,
                     occupancy_and_lists_ms_per_frame=ctx.occupancy_ms_per_frame,
                     bvh_ms_per_frame=ctx.bvh_ms_per_frame,
                     list_info_frame0=ctx.cell_list_info(0))

    # precision forms: same workload, other network arithmetic
    forms = {"fp32": dict(frames_per_s=args.steps / elapsed, samples_per_s=value,
                          network_ms_per_frame=prof["network"]["ms"] / args.steps)}
    ctxs = {"fp32": ctx}
    if not args.skip_forms:
        for prec in (3, 1, 2):
            c2, _, _, _, cams2 = setup(args, prec, av)
            e2, s2, _, p2 = timed_renders(c2, cams2, res, args.steps, args.warmup, out)
            p2.pop("_rounds_per_frame"), p2.pop("_live_rays_per_frame")
            forms[PRECISION_NAMES[prec]] = dict(frames_per_s=args.steps / e2, samples_per_s=s2 / e2,
                                                network_ms_per_frame=p2["network"]["ms"] / args.steps)
            ctxs[PRECISION_NAMES[prec]] = c2
        # the reference's main network shape: 16 expression values on the density input
        c3, _, _, _, cams3 = setup(args, 1, av, n_extra=16)
        e3, s3, _, p3 = timed_renders(c3, cams3, res, args.steps, args.warmup, out)
        forms["fp16_extra_dims"] = dict(frames_per_s=args.steps / e3, samples_per_s=s3 / e3,
                                        network_ms_per_frame=p3["network"]["ms"] / args.steps,
                                        note="FP16 with n_extra_dims 16 (per-frame values, 1.0 off the lips)")
        del c3
        forms["fp16_vs_fp32_frames"] = forms["fp16"]["frames_per_s"] / forms["fp32"]["frames_per_s"]
        forms["fp16_vs_fp32_scalar_frames"] = forms["fp16"]["frames_per_s"] / forms["fp32_scalar"]["frames_per_s"]
        if args.accuracy_pixels > 0:
            forms["accuracy_vs_fp64"] = accuracy_vs_fp64(ctxs, cfg, av, params, cams[0], res, args.accuracy_pixels)

    if not args.skip_forms:
        forms["small_shapes"] = precision_bench(args)
    phase("render and precision forms")
    train = None if args.skip_train or args.train_sizes == "full" else training_bench(args)
    phase("training (small shapes)")
    train_full = None if args.skip_train or args.train_sizes == "small" else training_bench(args, full=True)
    phase("training (full size)")
    serve = None if args.skip_serve or args.serve_sizes == "full" else serving_bench(args)
    phase("serving (small shapes)")
    serve_full = None if args.skip_serve or args.serve_sizes == "small" else serving_bench(args, full=True)
    phase("serving (full size)")

    cpu = None
    small = None
    if args.cpu_budget > 0:
        cpu = cpu_baseline(cfg, av, params, cams[0], res, args.cpu_budget)
        small = small_scalar_bench()

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded ellipsoid head, Gaussian expression bases, random-init hash grid + MLPs)",
        "config": {"workload": "full-size FP32 render: 9,976-tri head, 16 frames, L16 F2 T2^19, 128^3 grid, "
                               f"{res}x{res}, <=1024 steps", "resolution": res, "frames": 16,
                   "samples_per_frame": samples / args.steps, "emitted_per_frame": emitted / args.steps,
                   "live_rays_per_frame": live_rays, "march_rounds_per_frame": rounds},
        "frames_per_s": args.steps / elapsed,
        "precision_forms": forms,
        "roofline": roof,
        "kernels": kernels,
        "deform_traversal": traversal,
        # a posed frame's set-up (device BVH + occupancy grid + candidate
        # lists) is outside the timed renders, which reuse 16 resident frames;
        # the rate when every frame is a new pose rendered once includes it
        "posed_frame_setup": dict(
            bvh_ms=ctx.bvh_ms_per_frame, occupancy_and_lists_ms=ctx.occupancy_ms_per_frame,
            frames_per_s_one_render_per_pose=1e3 / (elapsed / args.steps * 1e3 + ctx.bvh_ms_per_frame
                                                    + ctx.occupancy_ms_per_frame)),
        "cpu_baseline": cpu,
        "small_scalar": small,
        "training": train,
        "training_full_size": train_full,
        "serving": serve,
        "serving_full_size": serve_full,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
