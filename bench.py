#!/usr/bin/env python3
"""Benchmark of the splat path on one MI355X.

Workload (default): plain forward frame of 6M splats at 1600x900 -- 256
anisotropic clusters in an 80 x 20 x 80 m box plus 20% ground-plane points,
orbit camera at r = 60 m, phi = theta = 45 deg, fovy 60 deg.  One step is one
frame: cull + project + compact, depth sort, tile binning, tile sort, ranges,
raster; replayed from a hipGraph with the splat storage resident in HBM.

Prints one JSON line with FPS (`value`), the per-stage split, the roofline of
the dominant kernel measured with HIP events over the timed region, and the
CPU oracle timed on the host cores as `cpu_baseline`.
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import math
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)

METRIC = "FPS @6M splats 1600x900; batched vs plain fwd+bwd views/s; % HBM peak"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
# Without packed FP32 instructions (this build disables them) a wave64 FP32 instruction issues once per
# 4 cycles per SIMD: 256 CUs x 4 SIMDs x 16 lanes x 2 (FMA) x 2.4 GHz
FP32_ISSUE_TFLOPS = 78.6
PAIRS_ALWAYS, PAIRS_NEVER = 4, 8  # vkgs_raster_config.flags
# FP32 operations per (pixel, splat) evaluation of the reference fragment and blend, exp counted as
# 1: eigenframe coordinates, falloff, alpha, and the blend of colour and alpha
RASTER_FLOPS_PER_EVAL = 20
# FP32 operations per (pixel, entry) evaluation of the blend backward, counted the same way (FMA = 1,
# exp and reciprocal = 1 each): eigenframe coordinates and the coverage test 9; falloff, alpha and its
# clamp, transmittance recovery T / (1 - alpha), dL/dalpha, the behind-colour update and the opacity
# and centre terms 25; the nine per-entry words (colour 3, conic 3) 6; their sum over the pixels 9
BACKWARD_OPS_PER_EVAL = 49
# the renderer's depth sort (integer keys fl(1 - z) * 2^24 in digits of 7, 8, 8, 2 bits) skips its last two
# passes when the frame's largest key is below this
DEPTH_PAIR_SKIP_BELOW = 1 << 15


def parse_args(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="c3_6m")
    ap.add_argument("--resolution", default="", metavar="WxH",
                    help="render the plain-forward workload at this size instead of its own (e.g. 1920x1080)")
    ap.add_argument("--tile", default="auto", help="WxH, or auto (vkgs_renderer_autotune)")
    ap.add_argument("--t-min", type=float, default=1.0 / 8192.0)
    ap.add_argument("--triangles", action="store_true",
                    help="triangle coverage (VKGS_RASTER_TRIANGLES), the reference's rasterization, for its cost")
    ap.add_argument("--timed-frames", type=int, default=20, help="eager frames with per-stage HIP events")
    ap.add_argument("--cpu-frames", type=int, default=3, help="oracle frames for cpu_baseline (0 = skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--fwd-bwd-steps", type=int, default=5, help="timed fwd+bwd steps of the batch workload (0 = skip)")
    ap.add_argument("--no-fwd-bwd-tune", action="store_true",
                    help="batched fwd+bwd forms with the default configuration instead of vkgs_diff_autotune's")
    ap.add_argument("--batch-scale", type=float, default=1.0, help="scales every scene size of the batch workload")
    ap.add_argument("--fwd-bwd-mode", choices=("both", "plain", "batched", "batched_eager"), default="both",
                    help="which fwd+bwd forms to time (one form alone is for profiling it)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="also run one device-to-device copy of MIB MiB (a known-size kernel for scaling the "
                         "FETCH_SIZE / WRITE_SIZE counters of a rocprofv3 --pmc run)")
    return ap.parse_args(argv)


TRAFFIC_NOTE = ("HBM bytes per launch from the committed rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE summary "
                "profiles/traffic_<workload>.json (not measured in this run; its header names the commit)")


def load_traffic(kernel: str, workload: str, tile: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc summary, if one exists for
    this workload and tile shape."""
    path = os.path.join(HERE, "profiles", f"traffic_{workload}.json")
    try:
        with open(path) as f:
            data = json.load(f)
        if data.get("tile", "8x8") != tile:
            return None
        return data["kernels"][kernel]["hbm_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def batch_workload(scale: float = 1.0, seed: int = 5):
    """32 views of a 3M-splat scene at resolutions drawn from {256^2, 512^2, 800^2, 1600x900} on orbit
    cameras of radius 8-30 m, plus 8 scenes of 50k-2M splats (log-uniform) with 2 views each."""
    import numpy as np
    from vkgs_amd import scene as S

    rng = np.random.default_rng(seed)
    big = S.clustered(max(1000, int(3_000_000 * scale)), seed=50, clusters=64, box=(40.0, 10.0, 40.0))
    scenes = [big]
    counts = np.exp(rng.uniform(np.log(50_000), np.log(2_000_000), 8))
    for i, c in enumerate(counts):
        scenes.append(S.gaussian_cloud(max(1000, int(c * scale)), seed=60 + i, pos_std=4.0))
    res = [(256, 256), (512, 512), (800, 800), (1600, 900)]
    views = []

    def cam(w, h):
        r = rng.uniform(8.0, 30.0)
        phi = rng.uniform(0.35, 1.35)
        theta = rng.uniform(0.0, 2 * np.pi)
        return S.orbit_camera(w, h, r, phi=phi, theta=theta)

    for _ in range(32):
        views.append((0, cam(*res[rng.integers(0, 4)])))
    for i in range(8):
        for _ in range(2):
            views.append((1 + i, cam(*res[rng.integers(0, 4)])))
    return scenes, views


def c4_perturbed(sc, seed: int = 9):
    """Target parameters of the differentiation workload: positions and log-scales perturbed by
    N(0, 0.01), SH and opacity logits by N(0, 0.1)."""
    import numpy as np
    from vkgs_amd import scene as S

    rng = np.random.default_rng(seed)
    rows = sc.rows.copy()
    c = S.COL
    n = rows.shape[0]
    rows[:, 0:3] += rng.normal(0.0, 0.01, (n, 3)).astype(np.float32)
    rows[:, c["scale_0"]:c["scale_0"] + 3] += rng.normal(0.0, 0.01, (n, 3)).astype(np.float32)
    rows[:, c["opacity"]] += rng.normal(0.0, 0.1, n).astype(np.float32)
    rows[:, c["f_dc_0"]:c["f_rest_0"] + 45] += rng.normal(0.0, 0.1, (n, 48)).astype(np.float32)
    return S.Scene(rows)


def bench_c4(steps: int):
    """ms/view of fwd+bwd on the differentiation workload: 500k splats, one 1024x1024 view, L1 loss
    against the image of the perturbed parameters (one view per launch set)."""
    import torch
    from vkgs_amd import diff as D
    from vkgs_amd import scene as S

    wl = S.WORKLOADS["c4_500k"]
    sc = wl.make()
    cam = S.orbit_camera(wl.width, wl.height, wl.radius)
    params = D.scene_tensors(sc)
    with torch.no_grad():
        target = D.render(**D.scene_tensors(c4_perturbed(sc), requires_grad=False), camera=cam)

    def step():
        for t in params.values():
            t.grad = None
        img = D.render(**params, camera=cam)
        (img - target).abs().mean().backward()

    step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    return {"c4_ms_per_view": round(dt * 1e3, 3), "c4_views_per_s": round(1.0 / dt, 2),
            "c4_visible": D.last_info["visible"], "c4_pairs": D.last_info["pairs"]}


def bench_fwd_bwd(steps: int, scale: float, mode: str = "both", tune: bool = True):
    """views/s of fwd+bwd (L1 loss, the mean over each view summed over views) for the batch
    workload: one launch set and one loss per view (plain, default 8x8 configuration); one launch
    set for all views with the loss formed over their flat (P, 4) output and the 32-view scene
    processed in spatial order (batched_eager); and that step captured once into a hipGraph and
    replayed (batched: one graph launch per step, its loss gradient formed elementwise inside the
    graph).  With tune, the batched forms use the configuration vkgs_diff_autotune picks for the
    batch (tile shape, LDS batch depth, resident-wave cap), chosen before any timed step."""
    import torch
    from vkgs_amd import diff as D

    scenes, views = batch_workload(scale)
    params = [D.scene_tensors(s) for s in scenes]
    targets = [torch.rand((c.screen_size[1], c.screen_size[0], 4), device="cuda") for _, c in views]
    # the batched form's loss over the flat (P, 4) output: the same sum of per-view L1 means, with
    # each pixel weighted by 1 / (4 H W) of its view
    tgt_flat = torch.cat([t.reshape(-1, 4) for t in targets])
    w_flat = torch.cat([torch.full((c.screen_size[0] * c.screen_size[1], 1),
                                   1.0 / (4 * c.screen_size[0] * c.screen_size[1]), device="cuda")
                        for _, c in views])

    def zero_grad():  # as an optimizer step's zero_grad(set_to_none=True) would
        for p in params:
            for t in p.values():
                t.grad = None

    def plain_step():
        zero_grad()
        for (si, c), tgt in zip(views, targets):
            img = D.render(**params[si], camera=c)
            (img - tgt).abs().mean().backward()

    def l1_grad(flat):  # d/d(flat) of the summed per-view L1 means: elementwise, no reduction
        return torch.sign(flat - tgt_flat) * w_flat

    graphed = {}
    out = {}
    cfg = None
    if tune and mode != "plain":
        t0 = time.perf_counter()
        cfg, tuned_ms = D.autotune(params, views, spatial_order=True, steps=2)
        out["tuned"] = D.describe(cfg)
        out["tuned_ms_per_step"] = round(tuned_ms, 3)
        log(f"fwd+bwd autotune: {out['tuned']} ({tuned_ms:.3f} ms/step, search {time.perf_counter() - t0:.1f} s)")

    def batched_eager_step():
        zero_grad()
        flat, _ = D.render_batch_flat(params, views, spatial_order=True, config=cfg)
        ((flat - tgt_flat).abs() * w_flat).sum().backward()

    def batched_step():  # the whole fwd + bwd step over all views as one hipGraph launch
        if "step" not in graphed:
            zero_grad()
            graphed["step"] = D.GraphedStep(params, views, l1_grad, spatial_order=True, config=cfg)
        graphed["step"].step()

    forms = (("plain", plain_step), ("batched_eager", batched_eager_step), ("batched", batched_step))
    for name, fn in forms:
        if mode not in ("both", name):
            continue
        fn()  # warm-up: sizes every buffer (and captures the graphed step)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        out[f"{name}_views_per_s"] = round(len(views) / dt, 2)
        out[f"{name}_ms_per_view"] = round(dt * 1e3 / len(views), 3)
        if name == "batched_eager":
            out["batched_visible_items"] = D.last_info["visible"]
            out["batched_pairs"] = D.last_info["pairs"]
    if "step" in graphed:
        graphed["step"].check()  # no replay overflowed its pair list
    # rooflines: one eager step of each form with per-stage events and the rasters' work counters
    # (outside the timed steps)
    try:
        rooflines = {}
        if mode in ("both", "plain"):
            profs = []
            for (si, c), tgt in zip(views, targets):
                w = 1.0 / (4 * c.screen_size[0] * c.screen_size[1])
                t2 = tgt.reshape(-1, 4)
                profs.append(D.profile_step([params[si]], [(0, c)], lambda f, t2=t2, w=w: torch.sign(f - t2) * w))
            rooflines["plain"] = diff_roofline(profs)
        if mode in ("both", "batched", "batched_eager"):
            zero_grad()
            rooflines["batched"] = diff_roofline([D.profile_step(params, views, l1_grad, spatial_order=True,
                                                                 config=cfg)])
        zero_grad()
        out["roofline"] = rooflines
    except Exception as e:  # reported, never hides the step figures
        out["roofline"] = {"error": repr(e)}
    if mode == "both":
        out["batched_over_plain"] = round(out["batched_views_per_s"] / out["plain_views_per_s"], 3)
        out["batched_eager_over_plain"] = round(out["batched_eager_views_per_s"] / out["plain_views_per_s"], 3)
    out["views"] = len(views)
    out["splats"] = [s.n for s in scenes]
    return out


def diff_roofline(profs):
    """Roofline of the fwd+bwd step's dominant kernel, the raster backward (VALU-bound), from
    vkgs_diff_stage_ms / vkgs_diff_raster_work of one eager step (profs: one profile per call, summed):
    executed (pixel, entry) evaluations = 64 x the entries the waves stepped through."""
    ms = {k: sum(p["ms"][k] for p in profs) for k in profs[0]["ms"]}
    f_eval = 64.0 * sum(p["forward_entries"] for p in profs)
    b_eval = 64.0 * sum(p["backward_entries"] for p in profs)
    b_tf = BACKWARD_OPS_PER_EVAL * b_eval / (ms["backward_raster"] * 1e-3) / 1e12
    f_tf = RASTER_FLOPS_PER_EVAL * f_eval / (ms["forward_raster"] * 1e-3) / 1e12
    return {"kernel": "diff_raster_backward_pair_kernel / diff_raster_backward_kernel", "bound": "valu",
            "achieved": round(b_tf, 3), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(b_tf / FP32_PEAK_TFLOPS, 4), "frac_of_issue_ceiling": round(b_tf / FP32_ISSUE_TFLOPS, 4),
            "traffic": None, "avg_ms": round(ms["backward_raster"], 4),
            "work": f"{BACKWARD_OPS_PER_EVAL} FP32 ops (FMA, exp, rcp = 1) x {b_eval:.4g} executed (pixel, entry) "
                    f"evaluations over {len(profs)} call(s)",
            "forward_raster": {"achieved": round(f_tf, 3), "frac": round(f_tf / FP32_PEAK_TFLOPS, 4),
                               "frac_of_issue_ceiling": round(f_tf / FP32_ISSUE_TFLOPS, 4),
                               "avg_ms": round(ms["forward_raster"], 4),
                               "work": f"{RASTER_FLOPS_PER_EVAL} FP32 ops x {f_eval:.4g} executed evaluations"},
            "stage_ms": {k: round(v, 4) for k, v in ms.items()}}


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def median(xs):
    xs = sorted(xs)
    if not xs:
        return float("nan")
    m = len(xs) // 2
    return xs[m] if len(xs) % 2 else 0.5 * (xs[m - 1] + xs[m])


def host_cpu():
    """(CPU model name, logical CPUs this process may run on, logical CPUs of the machine)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count() or 0
    return model, usable, os.cpu_count() or 0


def main(argv=None) -> int:
    args = parse_args(argv)
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus N with N > 1 is not supported",
                          "n_gpus": args.gpus}))
        return 2
    import numpy as np
    import torch

    if not torch.cuda.is_available():
        print(json.dumps({"error": "no HIP device"}))
        return 1
    from vkgs_amd import pipeline as P
    from vkgs_amd import scene as S

    wl = S.WORKLOADS[args.workload]
    if args.resolution:
        rw, rh = (int(v) for v in args.resolution.lower().split("x"))
        wl = dataclasses.replace(wl, width=rw, height=rh)
    log(f"generating {wl.name}: {wl.n} splats")
    t0 = time.perf_counter()
    sc = wl.make()
    log(f"generated in {time.perf_counter() - t0:.1f}s")
    cam = S.orbit_camera(wl.width, wl.height, wl.radius)

    r = P.Renderer(sc.n)
    r.load_buffer(sc.ply_buffer(), sc.n)
    out = torch.empty((wl.height, wl.width, 4), dtype=torch.float32, device="cuda")
    if args.tile == "auto":
        r.set_config(8, 8, args.t_min, triangles=args.triangles)
        tw, th, tuned_ms = r.autotune(cam, out, frames=10)
        log(f"autotuned tile {tw}x{th} ({tuned_ms:.3f} ms/frame)")
    else:
        tw, th = (int(x) for x in args.tile.split("x"))
        r.set_config(tw, th, args.t_min, triangles=args.triangles)
    for _ in range(3):  # grows the pair storage if the first frame overflowed it
        r.render(cam, out)
        status, visible, pairs = r.frame_info()
        if status == 0:
            break
    if status != 0:
        print(json.dumps({"error": f"pair list overflow (status {status})"}))
        return 1
    log(f"visible {visible} pairs {pairs}")

    # ---- timed region: graph replays ----
    for _ in range(args.warmup):
        r.render(cam, out)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r.render(cam, out)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / args.steps
    fps = 1e3 / ms
    log(f"{ms:.3f} ms/frame, {fps:.1f} FPS")
    # per-frame spread: HIP events around each replay on the renderer's stream
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    for a, b in evs:
        a.record()
        r.render(cam, out)
        b.record()
    torch.cuda.synchronize()
    fr = sorted(a.elapsed_time(b) for a, b in evs)
    frame_ms = {"median": round(fr[len(fr) // 2], 4), "p10": round(fr[len(fr) // 10], 4),
                "p90": round(fr[(9 * len(fr)) // 10], 4)}
    log("frame ms", frame_ms)

    # ---- per-stage split with HIP events on the renderer's stream ----
    r.set_timing(True)
    stages = {k: [] for k in P.STAGES}
    for _ in range(args.timed_frames):
        r.render(cam, out)
        for k, v in r.stage_ms().items():
            if k in stages:
                stages[k].append(v)
    wave_entries = r.raster_work()  # of the last timed frame: the same camera, so every frame's
    r.set_timing(False)
    stage_ms = {k: round(median(v), 5) for k, v in stages.items()}
    log("stage ms", stage_ms)
    r_variant = 0

    if args.calibrate_copy > 0:  # known-size kernel for scaling rocprofv3 --pmc byte counters
        nbytes = args.calibrate_copy << 20
        src = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy of {nbytes} bytes done")
        del src, dst

    # ---- algorithmic work of the frame ----
    ranges, _ = r.tile_lists()
    lens = (ranges[:, 1] - ranges[:, 0]).astype(np.float64)
    tiles_x = (wl.width + tw - 1) // tw
    tx = np.arange(ranges.shape[0]) % tiles_x
    ty = np.arange(ranges.shape[0]) // tiles_x
    pix = (np.minimum(tw, wl.width - tx * tw) * np.minimum(th, wl.height - ty * th)).astype(np.float64)
    p_eval = float((pix * lens).sum())  # whole tile lists
    p_exec = 64.0 * wave_entries  # evaluations the raster executed (lists up to each tile's early exit)
    n, V, K = sc.n, visible, pairs
    b_proj = 12.0 * n + 172.0 * V  # positions of all + cov3d/opacity/SH read + 48-B record per visible
    t_proj = stage_ms["project"] * 1e-3
    f_raster = RASTER_FLOPS_PER_EVAL * p_exec
    t_raster = stage_ms["raster"] * 1e-3
    proj_gbs = b_proj / t_proj / 1e9
    raster_tflops = f_raster / t_raster / 1e12
    T = ranges.shape[0]
    t_ps = (stage_ms["count"] + stage_ms["project"] + stage_ms["depth_sort"] + stage_ms["bin"]
            + stage_ms["tile_sort"]) * 1e-3
    # algorithmic bytes of a (tile | depth) key design, fixed by the workload alone: 12-B pairs emitted,
    # P = ceil((ceil(log2 T) + 32) / 8) radix passes each reading and writing every pair, ranges
    passes = -(-(int(math.ceil(math.log2(max(T, 2)))) + 32) // 8)
    b_emit, b_sort, b_ranges = 12.0 * K, 24.0 * K * passes, 8.0 * K + 8.0 * T
    ps_gbs = (b_proj + b_emit + b_sort + b_ranges) / t_ps / 1e9
    # bytes this design moves by its own algorithm: depth sort over V with 12-B (key, slot | packed tile
    # rectangle) items in 2 or 4 passes (the integer depth key's last two passes run only when the frame's
    # largest key reaches 2^15), tile sort over K 8-B pairs, emit and ranges
    _, fkeys = r.sorted_ids()
    k_max = int((fkeys[:V].view(np.float32).astype(np.float64) * 2.0 ** 24).max()) if V else 0
    depth_passes = 4 if k_max >= DEPTH_PAIR_SKIP_BELOW else 2
    own_bytes = b_proj + 24.0 * depth_passes * V + 32.0 * K + 20.0 * K + 8.0 * T
    own_gbs = own_bytes / t_ps / 1e9

    # the raster kernel this configuration launches (8x8 tiles go two per wave from 8192 tiles up
    # unless a pair flag says otherwise)
    cfg = r.tuned if r.tuned is not None else None
    flags = cfg.flags if cfg is not None else 0
    pairs_form = (tw, th) == (8, 8) and r_variant == 0 and not (flags & PAIRS_NEVER) and (
        (flags & PAIRS_ALWAYS) or T >= 8192)
    # (the timed frames, whose stage times price the roofline, run the pair kernel's work-counting
    # form rasterize_pair_kernel<true>; graph-replayed frames run <false>, which skips the counts)
    raster_name = ("rasterize_tri_kernel" if args.triangles else "rasterize_pair_kernel<true>" if pairs_form else
                   "rasterize_kernel" if r_variant == 0 else "rasterize_wave_kernel")
    if t_raster >= t_proj:
        # vector fp32 work (no stage is a dense contraction, MFMA is unused)
        roof = {"kernel": raster_name, "bound": "valu", "achieved": round(raster_tflops, 3),
                "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(raster_tflops / FP32_PEAK_TFLOPS, 4),
                "frac_of_issue_ceiling": round(raster_tflops / FP32_ISSUE_TFLOPS, 4),
                "traffic": load_traffic(raster_name, wl.name, f"{tw}x{th}"), "avg_ms": stage_ms["raster"],
                "work": f"{RASTER_FLOPS_PER_EVAL} FP32 ops (exp = 1) x {p_exec:.4g} executed (pixel, splat) "
                        f"evaluations (64 x {wave_entries} wave steps; whole lists: {p_eval:.4g})",
                "traffic_from": TRAFFIC_NOTE}
    else:
        roof = {"kernel": "fused_write_kernel", "bound": "hbm", "achieved": round(proj_gbs, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(proj_gbs / HBM_PEAK_GBS, 4),
                "traffic": load_traffic("fused_write_kernel", wl.name, f"{tw}x{th}"),
                "avg_ms": stage_ms["project"], "work": f"12N + 172V = {b_proj:.4g} B"}
    roof_proj = {"kernel": "fused_write_kernel", "bound": "hbm", "achieved": round(proj_gbs, 1),
                 "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(proj_gbs / HBM_PEAK_GBS, 4),
                 "traffic": load_traffic("fused_write_kernel", wl.name, f"{tw}x{th}"), "avg_ms": stage_ms["project"],
                 "work": f"12N + 172V = {b_proj:.4g} B", "traffic_from": TRAFFIC_NOTE}

    result = {
        "metric": METRIC,
        "value": round(fps, 2),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"plain forward {wl.name}: {wl.n} splats, {wl.width}x{wl.height}, orbit r={wl.radius} m",
                   "splats": n, "width": wl.width, "height": wl.height, "visible": V, "pairs": K,
                   "tile": f"{tw}x{th}",
                   "raster_kernel": raster_name.replace("<true>", "<false>"),  # the graph-replayed frames' form
                   "waves_per_cu": (cfg.waves_per_cu if cfg is not None else 0) or "max",
                   "t_min": args.t_min, "graph": True},
        "frame_ms": frame_ms,
        "stages_ms": stage_ms,
        "roofline": roof,
        "roofline_projection": roof_proj,
        "hbm_frac_projection_sort": round(ps_gbs / HBM_PEAK_GBS, 4),
        "projection_sort": {"ms": round(t_ps * 1e3, 4), "bytes": round(b_proj + b_emit + b_sort + b_ranges),
                            "passes": passes, "gbs": round(ps_gbs, 1),
                            "own_bytes": round(own_bytes), "own_gbs": round(own_gbs, 1),
                            "depth_key_max": k_max, "depth_passes": depth_passes,
                            "work": "12N + 172V + 12K + 24K*P + 8K + 8T (P radix passes of a (tile|depth) key)"},
    }

    # host copy of the splat storage for the CPU baseline, then free the renderer
    host_storage = None
    if not args.no_cpu_baseline and args.cpu_frames > 0:
        st = r.storage()
        host_storage = (st.position.cpu().numpy(), st.cov3d.cpu().numpy(), st.opacity.cpu().numpy(),
                        st.sh.cpu().numpy().view(np.uint16))
        del st
    r.close()
    torch.cuda.empty_cache()

    # ---- fwd+bwd views/s: plain (one view per launch set) vs batched ----
    if args.fwd_bwd_steps > 0:
        try:
            result["fwd_bwd"] = bench_fwd_bwd(args.fwd_bwd_steps, args.batch_scale, args.fwd_bwd_mode,
                                              tune=not args.no_fwd_bwd_tune)
            if args.fwd_bwd_mode == "both":
                result["fwd_bwd"].update(bench_c4(max(args.fwd_bwd_steps, 3)))
            log("fwd+bwd", result["fwd_bwd"])
        except Exception as e:  # reported, never hides the frame figure
            result["fwd_bwd"] = {"error": repr(e)}

    # ---- CPU baseline: the oracle's full frame on the host cores ----
    if host_storage is not None:
        try:
            import oracle

            pos, cov, opa, sh = host_storage
            oracle.lib()
            times = []
            for _ in range(args.cpu_frames):
                c0 = time.perf_counter()
                ref = oracle.frame(cam, pos, cov, opa, sh, wl.width, wl.height)
                times.append(time.perf_counter() - c0)
            assert ref["visible"] == V
            cpu_fps = 1.0 / median(times)
            model, usable, total = host_cpu()
            result["cpu_baseline"] = {
                "value": round(cpu_fps, 4), "unit": "frames/s", "cores": oracle.num_threads(), "kind": "port",
                "sample": f"median of {args.cpu_frames} full frames of the same {wl.name} workload (rank, stable "
                          "sort, inverse, projection, back-to-front draw) by the C oracle with OpenMP",
                "frame_s": [round(t, 3) for t in times], "cpu_model": model, "cpus_usable": usable,
                "cpus_machine": total}
            log(f"cpu oracle {median(times):.2f} s/frame on {oracle.num_threads()} threads ({model}, "
                f"{usable} of {total} CPUs usable)")
        except Exception as e:  # the baseline never blocks the GPU figure
            result["cpu_baseline"] = {"value": None, "error": str(e)}
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
