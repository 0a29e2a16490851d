"""Benchmark of the rasterizer hot path on one MI355X.

One step = one training iteration of the drop-in API on the synthetic
"heavy" workload (1M Gaussians, 64-cluster mixture with 2% large splats,
SH degree 3, one 1280x840 camera): GaussianRasterizer forward, L2 loss to a
fixed random target, backward (rasterizer adjoint, preprocess adjoint and
the densification statistics that train_faster.py requests). Inputs are
resident in HBM before timing. Prints one JSON line.

    python bench.py [--gpus 1] [--steps 20] [--warmup 5] [--schedule fast|plain]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
COMPOSITE_FLOPS = 21  # per (pixel, splat) evaluation
# untimed steps before a frame shape's first timed run, over which the
# prefix-list divisor of the drop-in API settles (it doubles after every 4
# frames whose prefixes sufficed)
PREFIX_SETTLE_STEPS = 40


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="heavy")
    ap.add_argument("--schedule", default="fast", choices=["plain", "fast"])
    ap.add_argument("--bucket", type=int, default=32)
    ap.add_argument("--list-prefix-div", type=int, default=16,
                    help="build the lists of the nearest P / div Gaussians first (0: whole lists)")
    ap.add_argument("--cpu-baseline", type=int, default=1)
    ap.add_argument("--compare-plain", type=int, default=1)
    ap.add_argument("--small", type=int, default=1,
                    help="also time the small workload (100k, 800x600, bucket 32) on both "
                         "schedules, with its CPU baseline")
    ap.add_argument("--render-loop", type=int, default=1,
                    help="also run the graph-captured 120-frame orbit of the dense workload")
    ap.add_argument("--multiview", type=int, default=1,
                    help="also time the 16-view batched training step (autotuned)")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="only run a few training steps and a 256 MiB copy (for rocprofv3 --pmc "
                         "FETCH_SIZE / WRITE_SIZE runs); prints no JSON")
    ap.add_argument("--sched-eval", type=int, default=1,
                    help="also train the tile cost model online and score it against the heuristic")
    ap.add_argument("--ply", default=None,
                    help="also replay a trained 3DGS model file (point_cloud.ply) ...")
    ap.add_argument("--cameras", default=None,
                    help="... over the views of its cameras.json (forward FPS and fwd+bwd ms)")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(the image-space sharding design is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    import torch

    from fastergs_amd import scenes
    from fastergs_amd._lib import SCHEDULE_FAST, SCHEDULE_PLAIN
    from fastergs_amd.rasterizer import (GaussianRasterizationSettings, GaussianRasterizer,
                                         RasterOptions, get_default_options,
                                         set_default_options)

    dev = torch.device("cuda:0")
    scene, cam = scenes.workload(args.workload, seed=0)
    P = scene["means3D"].shape[0]
    params = {k: torch.from_numpy(v).to(dev).requires_grad_(True) for k, v in scene.items()}
    settings = GaussianRasterizationSettings(
        image_height=cam.height, image_width=cam.width, tanfovx=cam.tanfovx,
        tanfovy=cam.tanfovy, bg=torch.zeros(3, device=dev), scale_modifier=1.0,
        viewmatrix=torch.from_numpy(cam.viewmatrix).to(dev),
        projmatrix=torch.from_numpy(cam.projmatrix).to(dev), sh_degree=3,
        campos=torch.from_numpy(cam.campos).to(dev), mult=1.0, prefiltered=False, debug=False)
    target = torch.from_numpy(scenes.target_image(cam.width, cam.height)).to(dev)
    accum = torch.zeros(P, 1, device=dev)
    max_radii = torch.zeros(P, device=dev)
    denom = torch.zeros(P, 1, device=dev)
    rast = GaussianRasterizer(settings)

    def step():
        for t in params.values():
            t.grad = None
        img, radii = rast(means3D=params["means3D"], opacities=params["opacities"],
                          dc=params["dc"], shs=params["shs"], scales=params["scales"],
                          rotations=params["rotations"], xyz_gradient_accum=accum,
                          max_radii=max_radii, denom=denom)
        loss = torch.nn.functional.mse_loss(img, target)  # L2 loss, one fused reduction
        loss.backward()
        return img

    if args.traffic_probe:
        set_default_options(schedule=SCHEDULE_FAST if args.schedule == "fast" else SCHEDULE_PLAIN,
                            bucket_size=args.bucket)
        for _ in range(3):
            step()
        # calibration: a copy of known size (256 MiB read + 256 MiB written)
        src = torch.ones(64 << 20, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print("traffic probe done", flush=True)
        return

    settled = set()

    def settle_steps(schedule):
        o = get_default_options()
        key = (schedule, o.bucket_size, o.tile_w, o.list_prefix_div)
        if schedule != SCHEDULE_FAST or o.list_prefix_div < 2 or key in settled:
            return 0
        settled.add(key)
        return PREFIX_SETTLE_STEPS

    def timed(schedule: int, steps: int, warmup: int, fwd_only=False, bucket=None, tile_w=16,
              queue=False, prefix_div=None):
        set_default_options(schedule=schedule, bucket_size=args.bucket if bucket is None else bucket,
                            tile_w=tile_w, composite_queue=queue,
                            list_prefix_div=args.list_prefix_div if prefix_div is None
                            else prefix_div)
        # untimed: the warm-up steps, and before them the steps over which the
        # prefix-list policy of this frame shape settles (once per shape)
        for _ in range(warmup + settle_steps(schedule)):
            step()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        e0.record()
        for _ in range(steps):
            if fwd_only:
                with torch.no_grad():
                    rast(means3D=params["means3D"], opacities=params["opacities"],
                         dc=params["dc"], shs=params["shs"], scales=params["scales"],
                         rotations=params["rotations"])
            else:
                step()
        e1.record()
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) * 1e3 / steps
        return e0.elapsed_time(e1) / steps, wall

    sched = SCHEDULE_FAST if args.schedule == "fast" else SCHEDULE_PLAIN
    ms, wall = timed(sched, args.steps, args.warmup)
    prefix = None
    if sched == SCHEDULE_FAST and args.list_prefix_div > 1:
        from fastergs_amd import rasterizer as rz
        key = (cam.width, cam.height, args.bucket, 16, dev, args.list_prefix_div)
        pol = rz._prefix_policies.get(key)
        if pol is not None:
            prefix = {"start_div": args.list_prefix_div, "div": pol.good,
                      "records": -(-P // pol.good) if pol.good > 1 else P,
                      "settle_steps": PREFIX_SETTLE_STEPS}
    fwd_ms, _ = timed(sched, args.steps, 2, fwd_only=True)
    stages, roofline = kernel_roofline(step, params, settings, sched, args, P)
    plain_ms = None
    if args.compare_plain and sched != SCHEDULE_PLAIN:
        plain_ms, _ = timed(SCHEDULE_PLAIN, max(5, args.steps // 2), 2)
    # the drop-in default: the reference's 128-splat buckets (its semantics)
    dropin = None
    if sched == SCHEDULE_FAST and args.bucket != 128:
        d_ms, _ = timed(SCHEDULE_FAST, args.steps, 3, bucket=128)
        d_fwd, _ = timed(SCHEDULE_FAST, args.steps, 2, fwd_only=True, bucket=128)
        dropin = {"bucket_size": 128, "ms_per_step": round(d_ms, 4), "fwd_ms": round(d_fwd, 4),
                  "render_fps": round(1000.0 / d_fwd, 2)}
    # the tile shape (semantic like the bucket size) and the compositing
    # launch form (work-queue grid) at the bench's bucket size
    variants = None
    if sched == SCHEDULE_FAST:
        variants = {f"tile_{tw}x{256 // tw}": round(timed(sched, args.steps, 3, tile_w=tw)[0], 4)
                    for tw in (8, 16, 32)}
        variants["composite_queue"] = round(timed(sched, args.steps, 3, queue=True)[0], 4)
        # whole lists from the start (no prefix lists of the nearest Gaussians)
        variants["whole_lists"] = round(timed(sched, args.steps, 3, prefix_div=0)[0], 4)
    set_default_options(schedule=sched, bucket_size=args.bucket, tile_w=16, composite_queue=False,
                        list_prefix_div=args.list_prefix_div)

    graph_fps = graph_render_fps(settings, params, sched, args, prefix["div"] if prefix else 0)
    step_graph_ms = None
    if sched == SCHEDULE_FAST:
        for t in params.values():
            t.grad = None
        # the prefix size the eager frames settled on, fixed in the graph
        div = prefix["div"] if prefix else 0
        step_graph_ms = captured_step_ms([settings], params, [target],
                                         RasterOptions(schedule=sched, bucket_size=args.bucket,
                                                       list_prefix_div=div),
                                         steps=args.steps, stats=True)
    fma_tflops = fma_peak()
    adam = sparse_adam_step(rast, params)
    ssim = fused_ssim_step(cam)
    lists = list_passes_step(settings, params, args)
    small = small_config(args) if args.small else None
    loop = render_loop(args) if args.render_loop else None
    torch.cuda.empty_cache()
    mv = multiview_step(args) if args.multiview else None
    torch.cuda.empty_cache()
    sched_eval = learned_schedule_eval(args) if args.sched_eval else None
    torch.cuda.empty_cache()
    replay = scene_replay(args) if args.ply else None

    cpu = None
    if args.cpu_baseline:
        cpu = cpu_baseline(scene, cam)

    line = {
        "metric": "fwd+bwd ms/iter, 1M Gaussians @1280x840",
        "value": round(ms, 4),
        "unit": "ms/iter",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded 64-cluster scene, 2% large splats, SH degree 3)",
        "config": {"workload": f"{args.workload}: {P} Gaussians, {cam.width}x{cam.height}, "
                               f"SH3, bucket {args.bucket}, schedule {args.schedule}",
                   "parallelism": "single GPU"},
        "render_fps": round(1000.0 / fwd_ms, 2),
        "fwd_ms": round(fwd_ms, 4),
        "render_fps_graph": graph_fps,
        "wall_ms_per_step": round(wall, 4),
        "ms_per_step_graph": None if step_graph_ms is None else round(step_graph_ms, 4),
        "plain_ms_per_step": None if plain_ms is None else round(plain_ms, 4),
        "speedup_vs_plain": None if plain_ms is None else round(plain_ms / ms, 3),
        "dropin_default": dropin,
        "list_prefix": prefix,
        "variants_ms_per_step": variants,
        "small_config": small,
        "roofline": roofline,
        "fp32_fma_attainable_tflops": fma_tflops,
        "stages": stages,
        "sparse_adam": adam,
        "fused_ssim": ssim,
        "list_passes": lists,
        "render_loop": loop,
        "multiview": mv,
        "learned_scheduler": sched_eval,
        "scene_replay": replay,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


# Algorithmic work per launch of each timed kernel (bytes for HBM-bound
# kernels, flops for VALU-bound ones). P Gaussians, M rest SH coefficients,
# V visible Gaussians, R pairs, T tiles, E active (pixel, splat) pairs of
# the backward (sum of last contributors).
def kernel_units(name, P, M, V, R, T, E, fast):
    sh = 12 + 12 * M  # dc + rest coefficients
    table = {
        # geometry only: in means, opacity, scales, rotations; out geom 32,
        # splat 48, depth key 4, depth-sort key 8 + index 4, large flag 4, count 4, radius 4
        "preprocess_fast": ("hbm", P * (12 + 4 + 12 + 16) + P * (32 + 48 + 4 + 8 + 4 + 4 + 4 + 4)),
        # colour pass: in dc + rest SH, means, depth key; out rgb + clamp 16 and
        # (frames that will be differentiated) the 3x3 SH Jacobian 36
        "sh_color": ("hbm", P * (sh + 12 + 4) + V * 16 + P * 36),
        "preprocess_plain": ("hbm", P * (12 + 4 + 12 + 16 + sh) + P * (32 + 48 + 4 + 4 + 4)),
        # in: means, scales, rotations, opacity, rest SH, raw sums 48, radius, clamp;
        # out: mean2D 16, conic 16, opacity 4, colour 12, mean3D 12, cov3D 24, dc 12, SH, scale 12, rot 16
        "gauss_bwd": ("hbm", P * (12 + 12 + 16 + 4 + 12 * M + 48 + 4 + 4)
                      + P * (16 + 16 + 4 + 12 + 12 + 24 + 12 + 12 * M + 12 + 16)),
        # split adjoint, SH pass: in means, colour sums, clamp, radius, rest SH;
        # out dSH, ddc, mean term 16
        "sh_bwd": ("hbm", P * (12 + 12 + 4 + 4 + 12 * M) + P * (12 * M + 12 + 16)),
        # split adjoint, geometric pass: in raw sums 48, opacity, radius, means,
        # scales, rotations, clamp, SH mean term 16; out mean2D 16, opacity 4,
        # mean3D 12, scale 12, rotation 16 (the autograd path asks for no
        # conic, colour or cov3D gradients)
        "gauss_bwd_geom": ("hbm", P * (48 + 4 + 4 + 12 + 12 + 16 + 4 + 16)
                           + P * (16 + 4 + 12 + 12 + 16)),
        # adjoint with the forward's SH Jacobians: in raw sums 48, opacity,
        # radius, means, scales, rotations, clamp, Jacobian 36; out mean2D 16,
        # opacity 4, mean3D 12, scale 12, rotation 16, dc 12, rest SH
        "gauss_bwd_jac": ("hbm", P * (48 + 4 + 4 + 12 + 12 + 16 + 4 + 36)
                          + P * (16 + 4 + 12 + 12 + 16 + 12 + 12 * M)),
        # key-only pass of the tile sort: read + write 8 B per pair; the
        # last pass writes the 4 B list entry and the 4 B tile id instead
        "onesweep_k": ("hbm", 16 * R),
        "onesweep_k_last": ("hbm", 16 * R),
        "radix_scatter_k": ("hbm", 16 * R),
        "radix_scatter_k_last": ("hbm", 16 * R),
        # tile sort: 16-bit tile ids with 32-bit records; the histogram
        # reads the 2 B ids, each pass reads and writes 6 B per pair
        "tile_hist": ("hbm", 2 * R),
        "tile_scatter": ("hbm", 12 * R),
        # depth sort of the records: 8 B key (depth | record) + 4 B pair
        # count, read + write per pass; the histogram reads the keys
        "depth_onesweep": ("hbm", 24 * P),
        "depth_hist_all": ("hbm", 8 * P),
        # sorted packed keys -> lists (4 B) and ranges
        "unpack_ranges": ("hbm", 12 * R + 8 * T),
        # ~70 flops per active (pixel, splat) in the reverse step (fgs_math.h pix_bwd)
        "backward_fast": ("valu", 70 * E),
        "backward_plain": ("valu", 70 * E),
        # (pixel, splat) evaluations of compositing, 21 flops each (alpha
        # from the log2-domain conic: 3 fma + exp2 + min/compare; blend: 3 fma
        # + mul + sub): the plain pass evaluates each pixel up to its last
        # contributor (E); the fast pass's executed count comes from its
        # lists (composite_evaluations)
        "composite_plain": ("valu", COMPOSITE_FLOPS * E),
    }
    # 64-bit key + 32-bit value, read + write: the pair sort (plain) or the
    # Gaussian depth sort (fast); histograms read the 8 B keys
    table["onesweep_kv"] = ("hbm", 24 * (P if fast else R))
    table["sort_hist_kv"] = ("hbm", 8 * (P if fast else R))
    table["sort_hist_k"] = ("hbm", 8 * R)
    table["radix_scatter_kv"] = ("hbm", 24 * (P if fast else R))
    table["radix_hist_kv"] = ("hbm", 8 * (P if fast else R))
    table["radix_hist_k"] = ("hbm", 8 * R)
    return table.get(name)


def composite_evaluations(fr, bucket):
    """(pixel, splat) evaluations k_composite_fast executes on whole tiles:
    rounds of 1, 2, then 4 buckets, each for all 256 pixels, up to the round
    holding the bucket of the tile's furthest contributor."""
    v = fr.view_arrays()
    rng = v["ranges"].cpu().numpy().view(np.uint32).reshape(-1, 2).astype(np.int64)
    maxc = v["max_contrib"].cpu().numpy().view(np.uint32).astype(np.int64)
    length = rng[:, 1] - rng[:, 0]
    nb = (length + bucket - 1) // bucket
    last = np.maximum((maxc + bucket - 1) // bucket - 1, 0)  # bucket of the furthest contributor
    # buckets after the rounds run: 1, 3, 7, 11, ... up to the first above `last`
    ramp = np.where(last < 1, 1, np.where(last < 3, 3, 3 + 4 * ((last - 3) // 4 + 1)))
    done = np.minimum(nb, ramp)
    return int(256 * np.minimum(length, done * bucket).sum())


def kernel_roofline(step, params, settings, sched, args, P):
    """Times every kernel with HIP events on its launch stream over a few
    steps and reports the dominant one against its roofline."""
    import torch
    from fastergs_amd import _lib
    from fastergs_amd.rasterizer import Frame, RasterOptions
    fr = Frame(settings, params["means3D"].detach(), params["opacities"].detach(),
               params["scales"].detach(), params["rotations"].detach(),
               dc=params["dc"].detach(), shs=params["shs"].detach(),
               options=RasterOptions(schedule=sched, bucket_size=args.bucket)).forward()
    R, T = fr.R, ((fr.W + 15) // 16) * ((fr.H + 15) // 16)
    # the tile sort and the ranges handle the prefix lists' pairs when they
    # sufficed (the step's frames use the same prefix size)
    lc = fr.list_counts()
    R_sorted = lc["prefix_pairs"] if lc["prefix_pairs"] and not lc["redo"] else R
    V = int((fr.radii > 0).sum().item())
    E = int(fr.last_contributor.sum().item())
    M = int(params["shs"].shape[1])
    executed = composite_evaluations(fr, args.bucket) if sched == 1 else None
    del fr
    steps = 5
    torch.cuda.synchronize()
    _lib.profile_enable(True)
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    prof = _lib.profile_read()
    _lib.profile_enable(False)
    stages = {}
    best = None
    for name, (tot_ms, launches) in prof.items():
        avg_s = tot_ms / launches / 1e3
        ent = {"ms_per_iter": round(tot_ms / steps, 4), "launches_per_iter": launches / steps,
               "avg_us": round(avg_s * 1e6, 2)}
        units = kernel_units(name, P, M, V, R_sorted if name.startswith("tile_") else R, T, E,
                             sched == 1)
        if name == "composite_fast" and executed is not None:
            # the reference evaluates every pair for all 256 pixels of its
            # tile (E_f = 256 R); this pass stops at the round in which the
            # tile's last pixel stopped, so it executes fewer. The roofline
            # counts the executed evaluations; the reference-equivalent rate
            # (E_f per launch time) is reported beside it.
            ent.update({"evals_reference": 256 * R, "evals_executed": executed,
                        "reference_equiv_tflops":
                            round(COMPOSITE_FLOPS * 256 * R / avg_s / 1e12, 2)})
            units = ("valu", COMPOSITE_FLOPS * executed)
        if units is not None and launches % steps == 0:
            bound, per_launch = units
            ach = per_launch / avg_s / (1e9 if bound == "hbm" else 1e12)
            peak = HBM_PEAK_GBS if bound == "hbm" else FP32_PEAK_TFLOPS
            ent.update({"bound": bound, "achieved": round(ach, 2), "frac": round(ach / peak, 4)})
            if best is None or tot_ms > best[1]:
                best = (name, tot_ms, ach, peak, bound, per_launch)
        stages[name] = ent
    roof = None
    if best is not None:
        name, _, ach, peak, bound, per_launch = best
        roof = {"kernel": name, "bound": bound, "achieved": round(ach, 2), "peak": peak,
                "unit": "GB/s" if bound == "hbm" else "TFLOP/s", "frac": round(ach / peak, 4),
                "traffic": measured_traffic(name, args), "algorithmic_per_launch": per_launch}
    return stages, roof


def measured_traffic(kernel, args):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC
    summary (FETCH_SIZE and WRITE_SIZE runs of `bench.py --traffic-probe`,
    scaled by a copy of known size), for the configuration it was measured
    on; None otherwise."""
    path = os.path.join(ROOT, "profiles", "pmc_heavy_1M_1280x840_fast.json")
    if args.workload != "heavy" or args.schedule != "fast" or args.bucket != 32:
        return None
    if not os.path.exists(path):
        return None
    with open(path) as f:
        k = json.load(f)["kernels"].get(kernel)
    return None if k is None else k["traffic_bytes"]


def fma_peak(blocks=4096, iters=4096, reps=5):
    """Attainable non-packed FP32 FMA rate of this GPU (fgs_fma_probe: 8
    independent FMA chains per thread), TFLOP/s, beside the published peak."""
    import ctypes
    import torch
    from fastergs_amd._lib import check, lib
    out = torch.empty(blocks * 256, device="cuda")
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    p = ctypes.c_void_p(out.data_ptr())
    check(lib().fgs_fma_probe(p, blocks, iters, s), "fgs_fma_probe")
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        check(lib().fgs_fma_probe(p, blocks, iters, s), "fgs_fma_probe")
    e1.record()
    torch.cuda.synchronize()
    sec = e0.elapsed_time(e1) / reps / 1e3
    return round(16.0 * iters * blocks * 256 / sec / 1e12, 2)


def fused_ssim_step(cam, reps=20):
    """Fused SSIM loss forward + backward on one 3 x H x W image of the bench
    camera's size (the loss term of the training loops). Minimum traffic per
    pass: forward reads 2 images and writes 4 maps, backward reads 6 maps and
    writes 1 (4 B per value)."""
    import torch
    from fastergs_amd.ssim import fused_ssim
    g = torch.Generator(device="cuda").manual_seed(0)
    a = torch.rand(1, 3, cam.height, cam.width, device="cuda", generator=g).requires_grad_(True)
    b = torch.rand(1, 3, cam.height, cam.width, device="cuda", generator=g)
    fused_ssim(a, b).backward()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        a.grad = None
        fused_ssim(a, b).backward()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    n = a.numel()
    gbs = (n * 4 * (2 + 4 + 6 + 1)) / (ms * 1e-3) / 1e9
    return {"image": [3, cam.height, cam.width], "ms_fwd_bwd": round(ms, 4),
            "min_traffic_gbs": round(gbs, 1)}


def list_passes_step(settings, params, args, reps=10):
    """render_simp and render_depth over the bench frame's lists (the
    importance-score and depth-reinitialisation passes of the Mini-Splatting
    training loops), ms per call on the fast schedule's frame."""
    import torch
    from fastergs_amd._lib import SCHEDULE_FAST
    from fastergs_amd.rasterizer import Frame, RasterOptions
    with torch.no_grad():
        fr = Frame(settings, params["means3D"].detach(), params["opacities"].detach(),
                   params["scales"].detach(), params["rotations"].detach(),
                   dc=params["dc"].detach(), shs=params["shs"].detach(),
                   options=RasterOptions(schedule=SCHEDULE_FAST,
                                         bucket_size=args.bucket)).forward()
        out = {"pairs": fr.R}
        for name, fn in (("render_simp", fr.render_simp), ("render_depth", fr.render_depth)):
            fn()
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                fn()
            e1.record()
            torch.cuda.synchronize()
            out[name + "_ms"] = round(e0.elapsed_time(e1) / reps, 4)
    return out


def sparse_adam_step(rast, params, reps=20):
    """SparseGaussianAdam.step over the six parameter tensors of the heavy
    model (59 floats per Gaussian) with the visibility of the bench camera;
    HBM-bound: 28 B per updated element (read p, g, m, v; write p, m, v)."""
    import torch
    from fastergs_amd.optim import SparseGaussianAdam
    P = params["means3D"].shape[0]
    with torch.no_grad():
        _, radii = rast(means3D=params["means3D"], opacities=params["opacities"],
                        dc=params["dc"], shs=params["shs"], scales=params["scales"],
                        rotations=params["rotations"])
    vis = radii > 0
    ps = {k: v.detach().clone().requires_grad_(True) for k, v in params.items()}
    for k, p in ps.items():
        p.grad = torch.randn_like(p) * 1e-3
    opt = SparseGaussianAdam([{"params": [p], "lr": 1e-3} for p in ps.values()], lr=0.0, eps=1e-15)
    opt.step(vis, P)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        opt.step(vis, P)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    per = sum(p.numel() for p in ps.values()) // P
    n_vis = int(vis.sum().item())
    gbs = (n_vis * per * 28 + P) / (ms * 1e-3) / 1e9
    return {"ms_per_step": round(ms, 4), "gaussians": P, "visible": n_vis,
            "floats_per_gaussian": per, "achieved_gbs": round(gbs, 1),
            "frac_hbm": round(gbs / HBM_PEAK_GBS, 4)}


def _time_replays(rend, camt, reps):
    import torch
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for k in range(reps):
        rend.cambuf.copy_(camt[k % camt.shape[0]])
        rend.graph.replay()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1)


def captured_step_ms(settings_list, params, targets, opts, steps=10, stats=False):
    """ms per replay of a training step captured in one HIP graph
    (CapturedTrainingStep: forward, L2 loss gradient, backward, and with
    stats the densification statistics) for the given views and targets."""
    import torch
    from fastergs_amd.graph import CapturedTrainingStep
    p = {k: v.detach() for k, v in params.items()}
    P = p["means3D"].shape[0]
    kw = {}
    if stats:
        dev = p["means3D"].device
        kw = dict(xyz_gradient_accum=torch.zeros(P, 1, device=dev),
                  max_radii=torch.zeros(P, device=dev), denom=torch.zeros(P, 1, device=dev))
    step = CapturedTrainingStep(settings_list, p["means3D"], p["opacities"], p["scales"],
                                p["rotations"], dc=p["dc"], shs=p["shs"], options=opts,
                                headroom=1.1, **kw)
    for t, src in zip(step.targets(), targets):
        t.copy_(src.reshape(t.shape))
    for _ in range(3):
        step.step()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        step.step()
    e1.record()
    torch.cuda.synchronize()
    step.check()
    ms = e0.elapsed_time(e1) / steps
    del step
    torch.cuda.empty_cache()
    return ms


def graph_render_fps(settings, params, sched, args, div=16):
    """Forward-only FPS of the heavy workload with the frame captured in one
    HIP graph (prefix lists of P / div Gaussians)."""
    import torch
    from fastergs_amd.graph import CapturedRenderer
    from fastergs_amd.rasterizer import RasterOptions
    p = {k: v.detach() for k, v in params.items()}
    rend = CapturedRenderer(settings, p["means3D"], p["opacities"], p["scales"], p["rotations"],
                            dc=p["dc"], shs=p["shs"],
                            options=RasterOptions(schedule=sched, bucket_size=args.bucket,
                                                  list_prefix_div=div))
    camt = rend.cambuf.clone().reshape(1, 35)
    _time_replays(rend, camt, 5)
    ms = _time_replays(rend, camt, 50)
    rend.check()
    return round(50 * 1000.0 / ms, 2)


def render_loop(args, frames=120):
    """Render-only loop of the dense workload (3M Gaussians, 1920x1080, 120-frame
    orbit at radius 3.5): each frame is one replay of a captured HIP graph
    after copying the frame's camera into the buffers the graph reads.
    Reports FPS over the orbit and the tile sort's HBM bandwidth (algorithmic
    bytes: per 8-bit pass a histogram read of 8 B and a scatter read + write
    of 8 + 8 B per pair) from per-kernel HIP events of eager frames."""
    import torch
    from fastergs_amd import _lib, scenes
    from fastergs_amd._lib import SCHEDULE_FAST
    from fastergs_amd.graph import CapturedRenderer
    from fastergs_amd.rasterizer import Frame, GaussianRasterizationSettings, RasterOptions
    dev = torch.device("cuda:0")
    scene, cam0 = scenes.workload("dense", seed=0)
    cams = scenes.orbit_cameras(frames, cam0.width, cam0.height, radius=3.5)
    P = scene["means3D"].shape[0]
    p = {k: torch.from_numpy(v).to(dev) for k, v in scene.items()}
    del scene

    def settings_of(c):
        return GaussianRasterizationSettings(
            image_height=c.height, image_width=c.width, tanfovx=c.tanfovx, tanfovy=c.tanfovy,
            bg=torch.zeros(3, device=dev), scale_modifier=1.0,
            viewmatrix=torch.from_numpy(c.viewmatrix).to(dev),
            projmatrix=torch.from_numpy(c.projmatrix).to(dev), sh_degree=3,
            campos=torch.from_numpy(c.campos).to(dev), mult=1.0, prefiltered=False, debug=False)

    camt = torch.stack([torch.from_numpy(np.concatenate(
        [c.viewmatrix.reshape(16), c.projmatrix.reshape(16), c.campos.reshape(3)]))
        for c in cams]).to(dev)
    # the captured frame's prefix size: the largest divisor (shortest
    # prefix) for which no frame of the orbit needs the whole lists, found
    # on an untimed pass over the orbit per divisor
    best = None
    for div in (16, 32, 64, 128):
        opts = RasterOptions(schedule=SCHEDULE_FAST, bucket_size=args.bucket, list_prefix_div=div)
        rend = CapturedRenderer(settings_of(cams[0]), p["means3D"], p["opacities"],
                                p["scales"], p["rotations"], dc=p["dc"], shs=p["shs"],
                                options=opts)
        for _ in range(4):  # size the capacity for the whole orbit
            _time_replays(rend, camt, frames)
            if rend.check(grow=True):
                break
        ok = not rend.prefix_redone() and rend.copts.list_prefix > 0
        del rend
        torch.cuda.empty_cache()
        if not ok:
            break
        best = div
    opts = RasterOptions(schedule=SCHEDULE_FAST, bucket_size=args.bucket,
                         list_prefix_div=best or 0)
    rend = CapturedRenderer(settings_of(cams[0]), p["means3D"], p["opacities"], p["scales"],
                            p["rotations"], dc=p["dc"], shs=p["shs"], options=opts)
    for _ in range(4):  # size the capacity for the whole orbit
        _time_replays(rend, camt, frames)
        if rend.check(grow=True):
            break
    rend.prefix_redone()
    ms = _time_replays(rend, camt, frames)
    rend.check()
    redone = rend.prefix_redone()  # some timed replay needed the whole lists
    R_cap, NB_cap = rend.capacity()
    del rend
    torch.cuda.empty_cache()
    # sort bandwidth and pair counts from eager frames at four orbit positions
    sort_ms, sort_bytes, Rs, sorted_pairs = 0.0, 0, [], []

    def sorted_count(fr):
        # the tile sort's keys: the prefix lists' pairs when they sufficed
        lc = fr.list_counts()
        return lc["prefix_pairs"] if lc["prefix_pairs"] and not lc["redo"] else fr.R
    for k in range(0, frames, frames // 4):
        _lib.profile_enable(True)
        fr = Frame(settings_of(cams[k]), p["means3D"], p["opacities"], p["scales"],
                   p["rotations"], dc=p["dc"], shs=p["shs"], options=opts).forward()
        prof = _lib.profile_read()
        _lib.profile_enable(False)
        tile_sort = ("tile_hist", "tile_scatter", "tile_hist_all", "tile_onesweep")
        if any(n in prof for n in tile_sort):
            # 16-bit tile ids + 32-bit records: 2 B histogram read, 12 B per pass
            passes = sum(prof[n][1] for n in ("tile_scatter", "tile_onesweep") if n in prof)
            hists = sum(prof[n][1] for n in ("tile_hist", "tile_hist_all") if n in prof)
            sort_ms += sum(prof[n][0] for n in tile_sort if n in prof)
            sort_bytes += (2 * hists + 12 * passes) * sorted_count(fr)
        else:
            radix = ("sort_hist_k", "onesweep_k", "onesweep_k_last", "radix_hist_k",
                     "radix_scatter_k", "radix_scatter_k_last")
            passes = sum(prof[n][1] for n in radix[1:3] + radix[4:] if n in prof)
            sort_ms += sum(prof[n][0] for n in radix if n in prof)
            sort_bytes += (8 + passes * 16) * fr.R
        Rs.append(fr.R)
        sorted_pairs.append(sorted_count(fr))
        del fr
    torch.cuda.empty_cache()
    return {"workload": f"dense: {P} Gaussians, {cams[0].width}x{cams[0].height}, SH3, "
                        f"{frames}-frame orbit r=3.5, one HIP graph per frame",
            "fps": round(frames * 1000.0 / ms, 2), "ms_per_frame": round(ms / frames, 4),
            "pairs_sampled": Rs, "pairs_sorted": sorted_pairs, "pair_capacity": R_cap,
            "bucket_capacity": NB_cap, "prefix_div": best, "prefix_redone": redone,
            "sort_gbs": round(sort_bytes / (sort_ms / 1e3) / 1e9, 1),
            "sort_frac_hbm": round(sort_bytes / (sort_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4)}


def scene_replay(args, max_views=200):
    """Replay of a trained model file over its cameras: per view one eager
    forward (timed) and one forward + L2 loss + backward (timed), HIP events
    on the current stream; means over the views."""
    import torch
    from fastergs_amd import plyio
    from fastergs_amd.rasterizer import GaussianRasterizationSettings, GaussianRasterizer
    dev = torch.device("cuda:0")
    scene = plyio.load_ply(args.ply)
    cams = (plyio.load_cameras_json(args.cameras) if args.cameras
            else scenes_orbit(scene))[:max_views]
    deg = scene.pop("sh_degree")
    p = {k: torch.from_numpy(v).to(dev).requires_grad_(True) for k, v in scene.items()}
    fwd, step = [], []
    for c in cams:
        st = GaussianRasterizationSettings(
            image_height=c.height, image_width=c.width, tanfovx=c.tanfovx, tanfovy=c.tanfovy,
            bg=torch.zeros(3, device=dev), scale_modifier=1.0,
            viewmatrix=torch.from_numpy(c.viewmatrix).to(dev),
            projmatrix=torch.from_numpy(c.projmatrix).to(dev), sh_degree=deg,
            campos=torch.from_numpy(c.campos).to(dev), mult=1.0, prefiltered=False, debug=False)
        rast = GaussianRasterizer(st)
        kw = dict(means3D=p["means3D"], opacities=p["opacities"], dc=p["dc"], shs=p["shs"],
                  scales=p["scales"], rotations=p["rotations"])
        for timed_list, train in ((fwd, False), (step, True), (fwd, False), (step, True)):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            if train:
                img, _ = rast(**kw)
                torch.nn.functional.mse_loss(img, torch.zeros_like(img)).backward()
            else:
                with torch.no_grad():
                    rast(**kw)
            e1.record()
            torch.cuda.synchronize()
            timed_list.append(e0.elapsed_time(e1))
    fwd_ms = float(np.mean(fwd[1::2]))  # second of each pair (first one warms up the view)
    step_ms = float(np.mean(step[1::2]))
    return {"ply": os.path.basename(args.ply), "gaussians": int(p["means3D"].shape[0]),
            "sh_degree": deg, "views": len(cams), "fwd_ms": round(fwd_ms, 4),
            "fps": round(1000.0 / fwd_ms, 2), "fwd_bwd_ms": round(step_ms, 4)}


def scenes_orbit(scene, frames=60):
    """Orbit cameras around a scene's centre when no cameras.json is given."""
    from fastergs_amd import scenes
    xyz = scene["means3D"]
    centre = np.median(xyz, axis=0)
    radius = 2.0 * float(np.percentile(np.linalg.norm(xyz - centre, axis=1), 90))
    cams = []
    for k in range(frames):
        a = 2.0 * np.pi * k / frames
        pos = centre + radius * np.array([np.sin(a), 0.2, -np.cos(a)])
        cams.append(scenes.look_at_camera(pos, 1280, 840, target=tuple(centre)))
    return cams


def multiview_step(args, views=16, steps=5):
    """Batched multi-view training step: 2M Gaussians (heavy-tailed cluster
    distribution), 16 views on a seeded sphere r=3-5 with sizes drawn from
    {640x480, 960x720, 1280x840, 1920x1080}; forward + backward of all views
    with one launch per stage (rasterize_views), L2 loss per view. The
    autotuner picks bucket size and split threshold; the same step is also
    timed as 16 consecutive single-view steps with the chosen options."""
    import torch
    from fastergs_amd import scenes
    from fastergs_amd._lib import SCHEDULE_FAST
    from fastergs_amd.autotune import Autotuner
    from fastergs_amd.multiview import rasterize_views
    from fastergs_amd.rasterizer import (GaussianRasterizationSettings, RasterOptions,
                                         rasterize_gaussians)
    dev = torch.device("cuda:0")
    scene = scenes.cluster_scene(2_000_000, seed=0)
    P = scene["means3D"].shape[0]
    cams = scenes.sphere_cameras(views, seed=0)
    params = {k: torch.from_numpy(v).to(dev).requires_grad_(True) for k, v in scene.items()}
    del scene
    settings = [GaussianRasterizationSettings(
        image_height=c.height, image_width=c.width, tanfovx=c.tanfovx, tanfovy=c.tanfovy,
        bg=torch.zeros(3, device=dev), scale_modifier=1.0,
        viewmatrix=torch.from_numpy(c.viewmatrix).to(dev),
        projmatrix=torch.from_numpy(c.projmatrix).to(dev), sh_degree=3,
        campos=torch.from_numpy(c.campos).to(dev), mult=1.0, prefiltered=False, debug=False)
        for c in cams]
    targets = [torch.from_numpy(scenes.target_image(c.width, c.height, seed=v)).to(dev)
               for v, c in enumerate(cams)]

    def batched(opts):
        for t in params.values():
            t.grad = None
        imgs, _ = rasterize_views(settings, params["means3D"], params["opacities"],
                                  dc=params["dc"], shs=params["shs"], scales=params["scales"],
                                  rotations=params["rotations"], options=opts)
        sum(torch.nn.functional.mse_loss(img, t) for img, t in zip(imgs, targets)).backward()

    def sequential(opts):
        for t in params.values():
            t.grad = None
        from fastergs_amd.rasterizer import set_default_options
        set_default_options(schedule=SCHEDULE_FAST, bucket_size=opts.bucket_size,
                            split_tiles=opts.split_tiles, tile_sched=opts.tile_sched,
                            composite_queue=opts.composite_queue, tile_w=opts.tile_w)
        for st, t in zip(settings, targets):
            img, _ = rasterize_gaussians(params["means3D"], params["dc"], params["shs"], None,
                                         params["opacities"], params["scales"],
                                         params["rotations"], None, st, None, None, None, None,
                                         None)
            torch.nn.functional.mse_loss(img, t).backward()

    # bucket size (semantic, opted in), split threshold and compositing work
    # shape (whole 16x16 tiles or 16x8 bands)
    tuner = Autotuner(buckets=(16, 32, 64), splits=(32, 64, 128), work_splits=(None, 1), reps=2,
                      tile_ws=(8, 16, 32))
    best, table = tuner.tune(batched)
    times = {}
    for name, fn in (("batched", batched), ("sequential", sequential)):
        fn(best)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            fn(best)
        e1.record()
        torch.cuda.synchronize()
        times[name] = e0.elapsed_time(e1) / steps
    for t in params.values():
        t.grad = None
    torch.cuda.empty_cache()
    times["captured"] = captured_step_ms(settings, params, targets, best, steps=steps)
    del params
    return {"workload": f"multiview: {P} Gaussians, {views} views "
                        f"({sum(c.width * c.height for c in cams)} pixels, sizes "
                        f"{sorted(set((c.width, c.height) for c in cams))}), fwd+bwd per step",
            "ms_per_step_batched": round(times["batched"], 3),
            "ms_per_step_sequential": round(times["sequential"], 3),
            "ms_per_step_batched_graph": round(times["captured"], 3),
            "views_per_s": round(views * 1000.0 / times["batched"], 1),
            "speedup_vs_sequential": round(times["sequential"] / times["batched"], 3),
            "autotuned": {"bucket_size": best.bucket_size, "tile_w": best.tile_w,
                          "split_tiles": best.split_tiles,
                          "work_shape": "16x16 tiles" if best.tile_sched is None
                          else f"16x8 bands from {best.tile_sched.split_len} entries",
                          "composite_queue": bool(best.composite_queue)},
            "autotune_table_ms": table}


SPLIT_CANDIDATES = (1 << 24, 16384, 8192, 4096, 2048, 1024, 512, 256)  # 1 << 24: never split


def learned_schedule_eval(args, views=200, held_out=40, n=None, check_outputs=True):
    """Online training of the tile cost model on a seeded random camera
    trajectory of the large workload (6M Gaussians, 1237x822, 200 views),
    then compositing time on its last `held_out` views, which it was not
    trained on: the composite kernel and the schedule's own kernels (feature
    and cost-model inference, work list), HIP events, best of 3, for raster
    order, the judge and the learned schedule: split flags in tile order,
    longest-predicted first, and longest-predicted first from the previous
    view's prediction (ahead: the cost model runs on a side stream beside
    compositing, off the critical path). Each schedule walks the held-out
    views in trajectory order three times (after the last training view),
    best of the three per view; `forward_us` is the whole forward (HIP events
    on the caller's stream). The judge is the fixed-threshold heuristic with
    the threshold that was fastest on every 8th training view, among
    SPLIT_CANDIDATES (never splitting included). Regret = learned - judge
    per view on compositing's critical path, the composite kernel plus the
    schedule's kernels before it (negative: the model gains); the whole
    forward varies by tens of us between runs of the same work and is
    reported beside it. With
    check_outputs, every schedule's image of a held-out view must equal
    raster order's bit for bit."""
    import torch
    from fastergs_amd import _lib, scenes
    from fastergs_amd._lib import SCHEDULE_FAST
    from fastergs_amd.rasterizer import Frame, GaussianRasterizationSettings, RasterOptions
    from fastergs_amd.scheduler import CostModel, OnlineTrainer, TileSchedule
    dev = torch.device("cuda:0")
    scene, cam0 = scenes.workload("large", seed=0, n=n)
    P = scene["means3D"].shape[0]
    p = {k: torch.from_numpy(v).to(dev) for k, v in scene.items()}
    del scene
    cams = scenes.trajectory_cameras(views, cam0.width, cam0.height, seed=4)
    train = cams[:views - held_out]

    def frame(c, sched):
        st = GaussianRasterizationSettings(
            image_height=c.height, image_width=c.width, tanfovx=c.tanfovx, tanfovy=c.tanfovy,
            bg=torch.zeros(3, device=dev), scale_modifier=1.0,
            viewmatrix=torch.from_numpy(c.viewmatrix).to(dev),
            projmatrix=torch.from_numpy(c.projmatrix).to(dev), sh_degree=3,
            campos=torch.from_numpy(c.campos).to(dev), mult=1.0, prefiltered=False, debug=False)
        # forward-only frames (no gradient workspace for compositing to zero)
        # with whole lists for every variant: the tile schedules work on them
        # (a frame with a tile schedule never builds prefix lists), so the
        # raster order must too
        return Frame(st, p["means3D"], p["opacities"], p["scales"], p["rotations"], dc=p["dc"],
                     shs=p["shs"], forward_only=True,
                     options=RasterOptions(schedule=SCHEDULE_FAST, bucket_size=args.bucket,
                                           tile_sched=sched, list_prefix_div=0)).forward()

    sched_kernels = ("sched_items", "tile_features", "mlp_infer", "tile_predict")

    def render(c, sc, keep=False):
        """(forward us, composite us, schedule-kernel us on the critical path, image)"""
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        _lib.profile_enable(True)
        e0.record()
        fr = frame(c, sc)
        e1.record()
        torch.cuda.synchronize()
        prof = _lib.profile_read()
        _lib.profile_enable(False)
        comp = 1e3 * prof.get("composite_fast", (0.0, 0))[0]
        sch = 0.0 if sc.ahead else 1e3 * sum(prof[nm][0] for nm in sched_kernels if nm in prof)
        img = fr.out_color.clone() if keep else None
        del fr
        return 1e3 * e0.elapsed_time(e1), comp, sch, img

    def timed(c, sc, reps=3):
        best = None
        for _ in range(reps):
            t = render(c, sc)[:3]
            best = t if best is None or t[1] + t[2] < best[1] + best[2] else best
        return best

    model = CostModel(seed=0, lr=2e-3)
    trainer = OnlineTrainer(model, batch=2048)
    log = TileSchedule.raster(features=True, log_cycles=True)
    t0 = time.perf_counter()
    losses = []
    for c in train:
        fr = frame(c, log)
        torch.cuda.synchronize()
        trainer.observe(fr)
        del fr
        losses.append(trainer.train(20))
    train_s = time.perf_counter() - t0
    # the judge: the fixed threshold fastest on the training views
    table = {}
    for th in SPLIT_CANDIDATES:
        sc = TileSchedule.heuristic(th)
        table[th] = float(np.mean([sum(timed(c, sc, reps=2)[1:]) for c in train[::8]]))
    judge_th = min(table, key=table.get)
    scheds = {"raster": TileSchedule.raster(), "judge": TileSchedule.heuristic(judge_th),
              "learned": TileSchedule.learned(model),
              "learned_ordered": TileSchedule.learned(model, order=True),
              "learned_ahead": TileSchedule.learned(model, order=True, ahead=True)}
    evals = cams[views - held_out:]
    res = {k: [None] * held_out for k in scheds}
    refs = [None] * held_out
    identical = True
    for k, sc in scheds.items():
        for _ in range(3):
            render(train[-1], sc)  # the first held-out view follows its predecessor
            for i, c in enumerate(evals):
                fw, comp, sch, img = render(c, sc, keep=check_outputs)
                if check_outputs:
                    if refs[i] is None:
                        refs[i] = img
                    else:
                        identical = identical and torch.equal(img, refs[i])
                del img
                if res[k][i] is None or comp + sch < res[k][i][1] + res[k][i][2]:
                    res[k][i] = (fw, comp, sch)
    del refs
    torch.cuda.empty_cache()
    fwd = {k: np.array([t[0] for t in v]) for k, v in res.items()}
    # compositing's critical path: the composite kernel plus the schedule's
    # kernels that precede it (none for the ahead schedule, whose kernels run
    # beside the preprocess and beside compositing)
    cp = {k: np.array([t[1] + t[2] for t in v]) for k, v in res.items()}
    lv = ("learned", "learned_ordered", "learned_ahead")
    best_l = min(lv, key=lambda k: cp[k].sum())
    l, j, r = cp[best_l], cp["judge"], cp["raster"]
    return {"workload": f"large: {P} Gaussians, {cam0.width}x{cam0.height}, {views}-view seeded "
                        f"random trajectory r=3-5, {views - held_out} training + {held_out} "
                        "held-out views",
            "train_seconds": round(train_s, 2), "loss_first": round(losses[0], 4),
            "loss_last": round(losses[-1], 4),
            "judge_split_len": None if judge_th == SPLIT_CANDIDATES[0] else judge_th,
            "judge_table_us": {("never" if th == SPLIT_CANDIDATES[0] else str(th)): round(v, 2)
                               for th, v in table.items()},
            "compositing_path_us_mean": {k: round(float(v.mean()), 2) for k, v in cp.items()},
            "composite_us_mean": {k: round(float(np.mean([t[1] for t in v])), 2)
                                  for k, v in res.items()},
            "schedule_kernels_us_mean": {k: round(float(np.mean([t[2] for t in v])), 2)
                                         for k, v in res.items()},
            "forward_us_mean": {k: round(float(v.mean()), 2) for k, v in fwd.items()},
            "learned_variant": best_l,
            "regret_us_mean": round(float((l - j).mean()), 2),
            "regret_vs_raster_us_mean": round(float((l - r).mean()), 2),
            "learned_wins_frac": round(float((l < j).mean()), 3),
            "gain_vs_judge": round(float(j.sum() / l.sum()), 4),

            "outputs_identical": bool(identical) if check_outputs else None}


def small_config(args, steps=20, warmup=5):
    """The small workload (100k Gaussians uniform in a ball, 800x600, bucket
    32): fwd+bwd ms/iter through the drop-in API on both schedules, and the
    CPU oracle's fwd+bwd on the same inputs."""
    import torch
    from fastergs_amd import scenes
    from fastergs_amd._lib import SCHEDULE_FAST, SCHEDULE_PLAIN
    from fastergs_amd.rasterizer import (GaussianRasterizationSettings, GaussianRasterizer,
                                         set_default_options)
    dev = torch.device("cuda:0")
    scene, cam = scenes.workload("small", seed=0)
    P = scene["means3D"].shape[0]
    p = {k: torch.from_numpy(v).to(dev).requires_grad_(True) for k, v in scene.items()}
    st = GaussianRasterizationSettings(
        image_height=cam.height, image_width=cam.width, tanfovx=cam.tanfovx,
        tanfovy=cam.tanfovy, bg=torch.zeros(3, device=dev), scale_modifier=1.0,
        viewmatrix=torch.from_numpy(cam.viewmatrix).to(dev),
        projmatrix=torch.from_numpy(cam.projmatrix).to(dev), sh_degree=3,
        campos=torch.from_numpy(cam.campos).to(dev), mult=1.0, prefiltered=False, debug=False)
    target = torch.from_numpy(scenes.target_image(cam.width, cam.height)).to(dev)
    rast = GaussianRasterizer(st)

    def step():
        for t in p.values():
            t.grad = None
        img, _ = rast(means3D=p["means3D"], opacities=p["opacities"], dc=p["dc"], shs=p["shs"],
                      scales=p["scales"], rotations=p["rotations"])
        torch.nn.functional.mse_loss(img, target).backward()

    res = {"workload": f"small: {P} Gaussians uniform in a ball, {cam.width}x{cam.height}, SH3, "
                       "bucket 32"}
    for name, sch in (("fast", SCHEDULE_FAST), ("plain", SCHEDULE_PLAIN)):
        set_default_options(schedule=sch, bucket_size=32, list_prefix_div=args.list_prefix_div)
        # warm-up, and the steps over which the prefix-list policy settles
        for _ in range(warmup + (PREFIX_SETTLE_STEPS if sch == SCHEDULE_FAST else 0)):
            step()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            step()
        e1.record()
        torch.cuda.synchronize()
        res[f"{name}_ms_per_step"] = round(e0.elapsed_time(e1) / steps, 4)
    res["speedup_vs_plain"] = round(res["plain_ms_per_step"] / res["fast_ms_per_step"], 3)
    set_default_options(schedule=SCHEDULE_FAST, bucket_size=args.bucket)
    # the same step replayed from one HIP graph: GPU time without host overhead
    from fastergs_amd.rasterizer import RasterOptions
    for t in p.values():
        t.grad = None
    res["fast_ms_per_step_graph"] = round(
        captured_step_ms([st], p, [target], RasterOptions(schedule=SCHEDULE_FAST, bucket_size=32),
                         steps=steps), 4)
    del p
    if args.cpu_baseline:
        res["cpu_baseline"] = cpu_baseline(scene, cam, iters=5)
    return res


def cpu_baseline(scene, cam, iters=3):
    """CPU oracle (C++, OpenMP) fwd+bwd on the same workload, a few iterations."""
    from oracle.pyoracle import Oracle, num_threads
    from fastergs_amd import scenes
    tgt = scenes.target_image(cam.width, cam.height)
    times = []
    t_end = time.perf_counter() + 20.0
    while len(times) < iters and (not times or time.perf_counter() < t_end):
        t0 = time.perf_counter()
        o = Oracle(scene, cam, bucket=32, mode=0).forward()
        img = o.image()
        o.backward((2.0 * (img - tgt) / img.size).astype(np.float32))
        times.append((time.perf_counter() - t0) * 1e3)
        del o
    return {"value": round(float(np.median(times)), 1), "unit": "ms/iter",
            "cores": num_threads(), "kind": "port", "cpu": cpu_model(),
            "sample": f"{len(times)} full fwd+bwd iterations of the same workload (median)"}


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


if __name__ == "__main__":
    main()
