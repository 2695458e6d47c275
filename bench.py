"""Benchmark: power-cell walk forward + backward on one MI355X.

Workload (the 1M-cell 1080p configuration): 1,000,000 cells with sites from
an 8-component Gaussian mixture (centres U[-1,1]^3, sigma 0.2), power
weights U[0, 1e-4] (r = sqrt(w)), density lognormal(0, 1), SH degree 3
colour (48 coefficients per cell, N(0, 0.1), +0.5 offset), adjacency from the
lifted-hull power-diagram build, fp16 walk faces, one 1920x1080 pinhole
camera (60 deg) on the radius-3 sphere, per-pixel L2 loss against a uniform
random target.

One step = SH colour pre-pass + walk forward (recording each ray's
contributing segments, their transmittance and running colour) + loss
gradient + walk backward (the recorded segments sorted by cell and
differentiated in cell order; rays past the record's cap replayed from the
checkpoint in the same launch) + SH backward, inputs resident in HBM.  The
same step with the record read ray by ray, and with checkpoint-and-replay
and no record (ray by ray, or band by band through the grouped pipeline),
is timed beside it ("backward_forms").  Also reported: the
16-camera differentiation configuration (262,144 cells, 800x800), the tile
rasterizer on the same frame, and the two serving traces.  Prints ONE JSON
line.  Single GPU: --gpus above 1 exits with a message.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "dynamic-powerfoam_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
GRAD_KEYS = ("spheres", "nsigmas", "texel_sites", "texel_rgb", "texel_height")


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cells", type=int, default=1_000_000)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU-baseline work (0 = skip)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--serve-requests", type=int, default=1440,
                    help="requests in each serving trace (0 = skip); p99 is about the 15th worst of 1440")
    ap.add_argument("--serve-load", type=float, default=0.6, help="offered load as a fraction of plain capacity")
    ap.add_argument("--serve-capacity", type=int, default=16384, help="8x8 tiles per serving step")
    ap.add_argument("--serve-cells", type=int, default=3_000_000, help="cells of the single served scene")
    ap.add_argument("--raster", type=int, default=1, help="also time the tile rasterizer fwd+bwd (0 = skip)")
    ap.add_argument("--serve-scenes", type=int, default=8, help="scenes of the multi-scene serving leg (0 = skip)")
    ap.add_argument("--cell-order", choices=("morton", "input"), default="morton",
                    help="cells in Morton order (as the reference's training leaves them) or as generated")
    ap.add_argument("--adjacency", choices=("gpu", "qhull"), default="gpu",
                    help="power-diagram adjacency from the GPU builder or the CPU Qhull lift (same edges)")
    ap.add_argument("--serve-quantum", type=int, default=1024, help="tiles per request per step before leftovers")
    ap.add_argument("--serve-streams", type=int, default=3, help="server streams (table buffers)")
    ap.add_argument("--serve-backward", choices=("grouped", "recorded"), default="grouped",
                    help="server's forward+backward tiles: grouped by (request, cell) over the step, or tile by tile "
                         "from the tile's record")
    ap.add_argument("--serve-record-cap", type=int, default=128,
                    help="segments per ray the server records for forward+backward tiles (0 = replay)")
    ap.add_argument("--serve-oldest-share", type=float, default=1.0,
                    help="share of a step's capacity the oldest active request may take")
    ap.add_argument("--face-layout", choices=("packed", "f16"), default="f16",
                    help="walk faces as 12-byte records with the neighbour (packed) or the (E, 4) fp16 array "
                         "with the adjacency beside it (f16)")
    ap.add_argument("--record-cap", type=int, default=128,
                    help="contributing segments per ray the forward records for the backward (0 = replay)")
    ap.add_argument("--forms", default="grouped,recorded,replay,replay_grouped",
                    help="backward forms timed beside the step, comma-separated")
    ap.add_argument("--replay-band-tiles", type=int, default=0,
                    help="tiles per band of the banded replay backward form (0: the library default)")
    ap.add_argument("--backward", choices=("grouped", "recorded", "replay"), default="grouped",
                    help="walk backward: the record's segments sorted by cell (default), the record ray by ray, "
                         "or checkpoint-and-replay without a record")
    ap.add_argument("--c3", type=int, default=1,
                    help="also time the 16-camera differentiation configuration (262,144 cells, 800x800, 0 = skip)")
    ap.add_argument("--c1", type=int, default=1,
                    help="also run the plain forward configuration (4,096 cells, 128x128) in full on the GPU and "
                         "the CPU (0 = skip)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device-to-device copy 5 times (reference kernel for HBM byte counters)")
    ap.add_argument("--sweep-traces", type=int, default=3, help="--serve-sweep: independent traces per load")
    ap.add_argument("--sweep-requests", type=int, default=480, help="--serve-sweep: requests per trace")
    ap.add_argument("--serve-sweep", action="store_true",
                    help="only run both serving traces at loads 0.3/0.6/0.9 and at saturation, server and plain "
                         "path, and print that as one JSON object")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.serve_sweep:
        serve_sweep(args)
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: rays would shard over GPUs with "
              f"the scene replicated in each GPU's HBM, and only the single-MI355X path is built and reported",
              flush=True)
        sys.exit(2)

    import numpy as np
    import torch

    from powerfoam import _lib
    from powerfoam import raytrace as rt
    from powerfoam import shading, workloads
    from powerfoam.adjacency import build_power_adjacency
    from powerfoam.camera import cameras_to_device

    dev = "cuda"
    t0 = time.time()
    sc = workloads.gmm_scene(args.cells, seed=args.seed)
    cam = workloads.sphere_cameras(sc["rng"], 1, args.width, args.height)[0]
    target_np = sc["rng"].uniform(0.0, 1.0, (1, args.height, args.width, 3)).astype(np.float32)
    if args.cell_order == "morton":
        # sort_points (scene.py:197-233) leaves a trained scene in this order
        from powerfoam.geometry import morton_sort

        perm = morton_sort(torch.from_numpy(sc["points"]).to(dev)).long().cpu().numpy()
        for k in ("points", "radii", "density", "normals", "texel_sites", "texel_rgb", "texel_height", "sh"):
            if sc[k] is not None:
                sc[k] = np.ascontiguousarray(sc[k][perm])
    log(f"scene {args.cells} cells generated in {time.time() - t0:.1f}s; building power adjacency ({args.adjacency})")
    t1 = time.time()
    if args.adjacency == "gpu":
        from powerfoam.power import build_power_adjacency_gpu

        adj, off = build_power_adjacency_gpu(torch.from_numpy(sc["points"]).to(dev), torch.from_numpy(sc["radii"]).to(dev))
        torch.cuda.synchronize()
        adj, off = adj.cpu(), off.cpu()
    else:
        adj, off = build_power_adjacency(torch.from_numpy(sc["points"]), torch.from_numpy(sc["radii"]))
    adj_seconds = time.time() - t1
    log(f"adjacency: {adj.numel()} directed edges (mean degree {adj.numel() / args.cells:.2f}) in {adj_seconds:.1f}s")

    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    pts, rad = t(sc["points"]), t(sc["radii"])
    adj_d, off_d = adj.to(dev), off.to(dev)
    faces = rt.walk_faces(pts, rad, adj_d, off_d, torch.float16)
    sites = t(sc["texel_sites"])
    coeffs = t(sc["sh"])
    cams_dev = cameras_to_device([cam], None, dev)
    rt.start_cells_device(pts, rad, cams_dev)
    K = sites.shape[1]
    texel_rgb = torch.empty(1, args.cells, K, 3, dtype=torch.float32, device=dev)
    scene = rt.WalkScene(pts, rad, t(sc["density"]), t(sc["normals"]), sites, texel_rgb[0], t(sc["texel_height"]),
                         adj_d, off_d, faces, pack_faces=args.face_layout == "packed")
    target = t(target_np)
    W, H = args.width, args.height
    grads = dict(spheres=torch.zeros(args.cells, 4, device=dev), nsigmas=torch.zeros(args.cells, 4, device=dev),
                 texel_sites=torch.zeros(args.cells, K, 3, device=dev),
                 texel_rgb=torch.zeros(args.cells, K, 3, device=dev),
                 texel_height=torch.zeros(args.cells, K, device=dev))
    g_coeffs = torch.zeros_like(coeffs)
    if args.backward == "replay":
        args.record_cap = 0
    grouped = args.backward == "grouped"
    workspace = rt.grouped_workspace(args.cells, 1, W, H, args.record_cap, dev) if grouped else None
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    times = {"fwd": [], "bwd": []}

    replay_ws = []  # the banded replay's workspace, allocated on first use
    band = args.replay_band_tiles or rt.REPLAY_BAND_TILES

    def step(record, kind=None):
        kind = kind or args.backward
        cap = 0 if kind in ("replay", "replay_grouped") else args.record_cap
        shading.sh_forward(coeffs, 3, sites, cams_dev, out=texel_rgb)
        if record:
            ev[0].record()
        f = rt.launch_walk_forward(scene, cams_dev, W, H, want_depth=False, record_cap=cap)
        if record:
            ev[1].record()
        gc = 2.0 * (f["color"] - target)
        for k in GRAD_KEYS:
            grads[k].zero_()
        g_coeffs.zero_()
        if record:
            ev[2].record()
        if kind == "replay_grouped":
            if not replay_ws:
                n = _lib.load().pf_walk_replay_grouped_workspace_bytes(args.cells, 1, W, H, 128, band)
                replay_ws.append(torch.empty(n, dtype=torch.uint8, device=dev))
            b = rt.launch_walk_backward(scene, cams_dev, W, H, f["color"], f["log_t"], gc, None, grads=grads,
                                        grouped=True, workspace=replay_ws[0], band_tiles=band)
        else:
            b = rt.launch_walk_backward(scene, cams_dev, W, H, f["color"], f["log_t"], gc, None, grads=grads,
                                        record=f["record"], grouped=kind == "grouped",
                                        workspace=workspace if kind == "grouped" else None)
        if record:
            ev[3].record()
        shading.sh_backward(coeffs, 3, sites, cams_dev, grads["texel_rgb"][None], grad_coeffs=g_coeffs)
        return f, b

    log("warmup")
    for _ in range(args.warmup):
        step(False)
    torch.cuda.synchronize()
    # per-kernel durations with events on the launch stream (torch's current)
    for _ in range(3):
        step(True)
        torch.cuda.synchronize()
        times["fwd"].append(ev[0].elapsed_time(ev[1]))
        times["bwd"].append(ev[2].elapsed_time(ev[3]))
    log("timed steps")
    torch.cuda.synchronize()
    s0 = time.perf_counter()
    for _ in range(args.steps):
        step(False)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - s0) / args.steps
    rays = W * H
    mrays = rays / dt / 1e6

    # the other backward forms on the same step: the record read ray by ray,
    # and checkpoint-and-replay without a record (no segment kept from the
    # forward), ray by ray or through the grouped pipeline band by band
    forms = {}
    kinds = [k for k in args.forms.split(",") if k]
    unknown = set(kinds) - {"grouped", "recorded", "replay", "replay_grouped"}
    if unknown:
        raise SystemExit(f"--forms: unknown backward forms {sorted(unknown)}")
    for kind in kinds:
        if kind != "grouped" and args.record_cap <= 0:
            continue
        for _ in range(2):
            step(False, kind)
        fw, bw, st = [], [], []
        for _ in range(3):
            torch.cuda.synchronize()
            a0 = time.perf_counter()
            step(True, kind)
            torch.cuda.synchronize()
            st.append((time.perf_counter() - a0) * 1e3)
            fw.append(ev[0].elapsed_time(ev[1]))
            bw.append(ev[2].elapsed_time(ev[3]))
        forms[kind] = {"fwd_ms": round(float(np.median(fw)), 3), "bwd_ms": round(float(np.median(bw)), 3),
                       "step_ms": round(float(np.median(st)), 3),
                       "mrays_per_s": round(W * H / (float(np.median(st)) * 1e-3) / 1e6, 2)}
    log(f"backward forms: {forms}")

    # algorithmic bytes from the kernels' own per-ray counters (untimed pass)
    shading.sh_forward(coeffs, 3, sites, cams_dev, out=texel_rgb)
    f = rt.launch_walk_forward(scene, cams_dev, W, H, want_depth=False, want_sequences=True, want_bytes=True,
                               record_cap=args.record_cap, want_ray_status=True)
    gc = 2.0 * (f["color"] - target)
    for k in GRAD_KEYS:
        grads[k].zero_()
    b = rt.launch_walk_backward(scene, cams_dev, W, H, f["color"], f["log_t"], gc, None, grads=grads,
                                want_bytes=True, record=f["record"], grouped=grouped, workspace=workspace)
    torch.cuda.synchronize()
    segs = None
    if f["record"] is not None:
        cnt = rt.record_counts(f["record"], 1, W, H).double()
        cap = f["record"]["cap"]
        segs = dict(cap=cap, mean=round(cnt.mean().item(), 2),
                    p99=float(torch.quantile(cnt.flatten()[::7], 0.99).item()), max=int(cnt.max().item()),
                    replayed_rays=int((cnt > cap).sum().item()))
        log(f"record: contributing segments per ray {segs}")
    fwd_bytes = float(f["alg_bytes"].sum().item())
    bwd_bytes = float(b["alg_bytes"].sum().item())
    visits = f["visits"].double()
    status = int(f["status"].item()) | int(b["status"].item())
    # each status bit with the number of rays that set it in the forward
    rs = f["ray_status"]
    status_rays = {name: int(((rs & bit) != 0).sum().item())
                   for name, bit in (("step_cap", 1), ("cycle", 4), ("bad_record", 8)) if status & bit}
    fwd_ms = float(np.median(times["fwd"]))
    bwd_ms = float(np.median(times["bwd"]))
    dom = "walk_bwd" if bwd_ms >= fwd_ms else "walk_fwd"
    dom_bytes, dom_ms = (bwd_bytes, bwd_ms) if dom == "walk_bwd" else (fwd_bytes, fwd_ms)
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    log(f"step {dt * 1e3:.2f} ms = {mrays:.1f} Mrays/s; walk fwd {fwd_ms:.2f} ms "
        f"({fwd_bytes / fwd_ms / 1e6:.0f} GB/s alg), bwd {bwd_ms:.2f} ms ({bwd_bytes / bwd_ms / 1e6:.0f} GB/s alg); "
        f"visits/ray mean {visits.mean().item():.1f} max {int(visits.max().item())}; status {status} "
        f"(rays: {status_rays})")

    cpu = None
    if args.cpu_seconds > 0:
        cpu = cpu_baseline(sc, adj.numpy(), off.numpy(), faces, texel_rgb, cams_dev, f, gc, args.cpu_seconds)

    raster = None
    if args.raster:
        raster = raster_bench(sc, pts, rad, texel_rgb[0], target, cam, args)

    c3 = c3_bench(args) if args.c3 else None
    c1 = c1_bench(args) if args.c1 else None

    serving = serving_multi = None
    if args.serve_requests > 0:
        # the serving legs build their own scenes: free the benchmark scene
        del scene, grads, texel_rgb, faces, f, b, gc
        torch.cuda.empty_cache()
        serving = serve_bench(args)
        if args.serve_scenes > 0:
            serving_multi = serve_multi_bench(args)

    out = {
        "metric": "Mrays/s fwd+bwd (1M cells, 1080p) + p50/p99 render latency vs plain path",
        "value": round(mrays, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Gaussian-mixture scene, random target image)",
        "config": {
            "workload": f"walk fwd+bwd, {args.cells} cells GMM, {W}x{H}, SH3, L2 loss",
            "cells": args.cells,
            "width": W,
            "height": H,
            "rays_per_step": rays,
            "face_storage": "f16" if args.face_layout == "f16" else "f16 packed with the neighbour",
            "cell_order": args.cell_order,
            "adjacency": args.adjacency,
            "adjacency_build_s": round(adj_seconds, 2),
            "attribute_storage": "f32",
            "mean_visits_per_ray": round(visits.mean().item(), 2),
            "walk_fwd_ms": round(fwd_ms, 4),
            "walk_bwd_ms": round(bwd_ms, 4),
            "walk_backward": args.backward,
            "walk_fwd_alg_bytes": fwd_bytes,
            "walk_bwd_alg_bytes": bwd_bytes,
            "record_segments": segs,
            "status": status,
            "status_rays": status_rays,
            "backward_forms": forms,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": profiled_traffic(dom) if (args.cells, W, H) == (1_000_000, 1920, 1080) else None,
        },
        "cpu_baseline": cpu,
        "raster": raster,
        "differentiation_16_cameras": c3,
        "plain_forward_4096_cells": c1,
        "serving": serving,
        "serving_multi": serving_multi,
    }
    print(json.dumps(out), flush=True)


def c3_bench(args):
    """The differentiation configuration: 262,144 GMM cells, 16 cameras of
    800x800 on the radius-3 sphere, each with its own SH pre-pass colours,
    walk forward recording segments, the grouped backward (per-camera colour
    gradients) and the SH backward, an L2 loss against random targets."""
    import numpy as np
    import torch

    from powerfoam import raytrace as rt
    from powerfoam import shading, workloads
    from powerfoam.camera import cameras_to_device
    from powerfoam.geometry import morton_sort
    from powerfoam.power import build_power_adjacency_gpu

    n, B, W, H = 262_144, 16, 800, 800
    sc = workloads.gmm_scene(n, seed=args.seed + 7)
    cams = workloads.sphere_cameras(sc["rng"], B, W, H)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda")
    perm = morton_sort(t(sc["points"])).long().cpu().numpy()
    for k in ("points", "radii", "density", "normals", "texel_sites", "texel_height", "sh"):
        sc[k] = np.ascontiguousarray(sc[k][perm])
    pts, rad = t(sc["points"]), t(sc["radii"])
    adj, off = build_power_adjacency_gpu(pts, rad)
    faces = rt.walk_faces(pts, rad, adj, off, torch.float16)
    cams_dev = cameras_to_device(cams, None, "cuda")
    rt.start_cells_device(pts, rad, cams_dev)
    sites, coeffs = t(sc["texel_sites"]), t(sc["sh"])
    K = sites.shape[1]
    texel_rgb = torch.empty(B, n, K, 3, device="cuda")
    shading.sh_forward(coeffs, 3, sites, cams_dev, out=texel_rgb)
    scene = rt.WalkScene(pts, rad, t(sc["density"]), t(sc["normals"]), sites, texel_rgb, t(sc["texel_height"]), adj,
                         off, faces, pack_faces=args.face_layout == "packed")
    cap = rt.RayTracer(None, "cuda").record_cap_for(B, W, H)
    workspace = rt.grouped_workspace(n, B, W, H, cap, "cuda")
    target = torch.rand(B, H, W, 3, generator=torch.Generator().manual_seed(args.seed + 8)).to("cuda")
    grads = dict(spheres=torch.zeros(n, 4, device="cuda"), nsigmas=torch.zeros(n, 4, device="cuda"),
                 texel_sites=torch.zeros(n, K, 3, device="cuda"), texel_rgb=torch.zeros(B, n, K, 3, device="cuda"),
                 texel_height=torch.zeros(n, K, device="cuda"))
    g_coeffs = torch.zeros_like(coeffs)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]

    def step(want_bytes=False):
        shading.sh_forward(coeffs, 3, sites, cams_dev, out=texel_rgb)
        ev[0].record()
        f = rt.launch_walk_forward(scene, cams_dev, W, H, want_depth=False, record_cap=cap, want_bytes=want_bytes)
        ev[1].record()
        gc = 2.0 * (f["color"] - target)
        for k in GRAD_KEYS:
            grads[k].zero_()
        g_coeffs.zero_()
        ev[2].record()
        b = rt.launch_walk_backward(scene, cams_dev, W, H, f["color"], f["log_t"], gc, None, grads=grads,
                                    record=f["record"], grouped=True, workspace=workspace, want_bytes=want_bytes)
        ev[3].record()
        shading.sh_backward(coeffs, 3, sites, cams_dev, grads["texel_rgb"], grad_coeffs=g_coeffs)
        return f, b

    for _ in range(2):
        step()
    torch.cuda.synchronize()
    fw, bw = [], []
    for _ in range(3):
        step()
        torch.cuda.synchronize()
        fw.append(ev[0].elapsed_time(ev[1]))
        bw.append(ev[2].elapsed_time(ev[3]))
    reps = 5
    torch.cuda.synchronize()
    s0 = time.perf_counter()
    for _ in range(reps):
        step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - s0) / reps
    f, b = step(want_bytes=True)
    torch.cuda.synchronize()
    fwd_bytes, bwd_bytes = float(f["alg_bytes"].sum().item()), float(b["alg_bytes"].sum().item())
    cnt = rt.record_counts(f["record"], B, W, H).double()
    segs = dict(cap=cap, mean=round(cnt.mean().item(), 2), max=int(cnt.max().item()),
                replayed_rays=int((cnt > cap).sum().item()))
    fwd_ms, bwd_ms = float(np.median(fw)), float(np.median(bw))
    dom_bytes, dom_ms, dom = (fwd_bytes, fwd_ms, "walk_fwd") if fwd_ms >= bwd_ms else (bwd_bytes, bwd_ms, "walk_bwd")
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    rays = B * W * H
    out = {"workload": f"{n} GMM cells, {B} cameras {W}x{H}, per-camera SH3 pre-pass, walk fwd (record cap {cap}) "
                       f"+ grouped bwd + SH bwd, L2 loss",
           "ms_per_step": round(dt * 1e3, 3), "mrays_per_s": round(rays / dt / 1e6, 2), "walk_fwd_ms": round(fwd_ms, 3),
           "walk_bwd_ms": round(bwd_ms, 3), "walk_fwd_alg_bytes": fwd_bytes, "walk_bwd_alg_bytes": bwd_bytes,
           "record": segs, "status": int(f["status"].item()) | int(b["status"].item()),
           "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                        "note": "algorithmic bytes count every visit's reads of a scene small enough to stay in "
                                "the 256 MB Infinity Cache across the 16 cameras, so frac can pass 1"}}
    log(f"16 cameras: step {dt * 1e3:.2f} ms = {out['mrays_per_s']} Mrays/s; fwd {fwd_ms:.2f} ms, bwd {bwd_ms:.2f} ms; "
        f"record {segs}")
    return out


def raster_bench(sc, pts, rad, texel_rgb, target, cam, args):
    """The pinhole tile rasterizer (Rasterizer.forward and its backward,
    rasterize.py:892-1516) on the benchmark's scene, camera and loss, over
    the scene's Cech complex (AABBTree.build_cech_complex, bvh.py:96-123, on
    the GPU).  Timed with events on torch's stream, binning included."""
    import types

    import numpy as np
    import torch

    from powerfoam.bvh import AABBTree
    from powerfoam.rasterize import Rasterizer

    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda")
    t0 = time.time()
    tree = AABBTree("cuda")
    tree.update(pts, rad)
    adj, off = tree.build_cech_complex()
    torch.cuda.synchronize()
    cech_s = time.time() - t0
    ra = types.SimpleNamespace(render_objective="volume", num_texel_sites=8, is_pinhole=True,
                               disable_coop_prim_load=False, disable_coop_adj_load=False)
    rz = Rasterizer(ra, "cuda")
    params = [pts.clone().requires_grad_(True), rad.clone().requires_grad_(True),
              t(sc["density"]).requires_grad_(True), t(sc["normals"]).requires_grad_(True),
              t(sc["texel_sites"]).requires_grad_(True), texel_rgb.clone().requires_grad_(True),
              t(sc["texel_height"]).requires_grad_(True)]
    tgt = target[0]

    def run(bwd):
        out = rz.forward(cam, None, *params, adj, off, None, False)
        if bwd:
            ((out[0] - tgt) ** 2).sum().backward()
        return out

    for _ in range(2):
        run(True)
    torch.cuda.synchronize()
    times = {}
    for bwd in (False, True):
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ms = []
        for _ in range(3):
            for p in params:
                p.grad = None
            ev[0].record()
            run(bwd)
            ev[1].record()
            torch.cuda.synchronize()
            ms.append(ev[0].elapsed_time(ev[1]))
        times[bwd] = float(np.median(ms))
    W, H = cam.width, cam.height
    # algorithmic bytes, from the forward's own tile lists and stop positions:
    # per (tile, listed cell up to the tile's stop) the cell's sphere and
    # [n, sigma] 32 B, CSR bounds 8 B, fp16 faces 8 B each and K = 8 texels
    # 224 B, staged once per 8x8 tile, plus the list entry 4 B; per pixel
    # the outputs (colour, log T, normal distance, normal: 32 B); the
    # binning reads each sphere (16 B) and writes and sorts 12 B per entry.
    # The backward reads the same per (tile, cell) plus the neighbour ids
    # (4 B per face), and read-modify-writes the cell's gradient row
    # 2 (32 + 28 K) B; per pixel the incoming gradient and log T 16 B.
    run(False)
    offsets, prims, early, csr = rz.last_lists
    tiles = offsets.numel() - 1
    starts = offsets[:-1]
    idx = torch.arange(prims.numel(), device="cuda")
    tile_of = torch.repeat_interleave(torch.arange(tiles, device="cuda"), (offsets[1:] - starts))
    live = (idx - starts[tile_of]) < early.long()[tile_of]
    deg = (csr[1:] - csr[:-1]).long()[prims.long()]
    n_live = float(live.sum().item())
    faces_live = float(deg[live].sum().item())
    n_isect = float(prims.numel())
    fwd_bytes = n_live * (32 + 8 + 224 + 4) + faces_live * 8 + W * H * 32 + sc["points"].shape[0] * 16 + n_isect * 24
    bwd_bytes = n_live * (32 + 8 + 224 + 4 + 2 * (32 + 28 * 8)) + faces_live * 12 + W * H * 16
    achieved = (fwd_bytes + bwd_bytes) / (times[True] * 1e-3) / 1e9
    out = {"adjacency": "cech", "adjacency_edges": int(adj.numel()), "cech_build_s": round(cech_s, 2),
           "fwd_ms": round(times[False], 3), "fwd_bwd_ms": round(times[True], 3),
           "mrays_per_s": round(W * H / (times[True] * 1e-3) / 1e6, 3),
           "fwd_alg_bytes": fwd_bytes, "bwd_alg_bytes": bwd_bytes,
           "roofline": {"bound": "hbm", "kernel": "raster fwd+bwd", "achieved": round(achieved, 1),
                        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4)},
           "workload": f"Rasterizer.forward + backward of an L2 loss, same scene, camera and target as the walk"}
    log(f"raster: Cech {adj.numel()} edges in {cech_s:.2f}s; fwd {times[False]:.2f} ms, fwd+bwd {times[True]:.2f} ms")
    return out


def build_served_scene(n, seed):
    """A resident GMM scene of n cells (as the benchmark's: SH degree 3,
    Morton-ordered cells, GPU power-diagram adjacency, fp16 walk faces)."""
    import numpy as np
    import torch

    from powerfoam import raytrace as rt
    from powerfoam import workloads
    from powerfoam.geometry import morton_sort
    from powerfoam.power import build_power_adjacency_gpu
    from powerfoam.serving import ResidentScene

    sc = workloads.gmm_scene(n, seed=seed)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda")
    pts = t(sc["points"])
    perm = morton_sort(pts).long()
    pts, rad = pts[perm].contiguous(), t(sc["radii"])[perm].contiguous()
    adj, off = build_power_adjacency_gpu(pts, rad)
    faces = rt.walk_faces(pts, rad, adj, off)
    return ResidentScene(pts, rad, t(sc["density"])[perm], t(sc["normals"])[perm], t(sc["texel_sites"])[perm],
                         t(sc["texel_height"])[perm], adj, off, faces, sh=t(sc["sh"])[perm])


def plain_service(scenes, make):
    """Warm the plain path on a mix of every kind of request (the
    allocator's cache) and measure its mean service time per request (s).
    Returns (warm-up requests, mean service time)."""
    import torch

    from powerfoam.serving import PlainRenderer

    plain = PlainRenderer(scenes)
    warm = make(32)
    for r in warm:
        plain.render(r)
    torch.cuda.synchronize()
    probe = make(16)
    t0 = time.perf_counter()
    for r in probe:
        plain.render(r)
        torch.cuda.synchronize()
    return warm, (time.perf_counter() - t0) / len(probe)


_SERVERS = {}


def serve_trace(scenes, make, args, label, load=None, calibration=None):
    """Serve one trace with the continuous-batching server and with the
    plain per-request path (one request at a time, FIFO).  Poisson arrivals
    at load x the plain path's measured capacity (load defaults to
    --serve-load); load "saturation" submits every request at once (closed
    loop: throughput at saturation, nothing rejected for waiting).
    calibration: (warm-up requests, plain mean service) from plain_service.
    Returns the stats."""
    import copy

    import numpy as np
    import torch

    from powerfoam.serving import PlainRenderer, RenderServer, latency_stats

    load = args.serve_load if load is None else load
    rng = np.random.default_rng(2000 + args.seed)
    plain = PlainRenderer(scenes)
    warm, mean_service = calibration if calibration is not None else plain_service(scenes, make)
    trace = make(args.serve_requests)
    if load == "saturation":
        rate = None
        for r in trace:
            r.arrival = 0.0
    else:
        rate = load / mean_service
        arr = np.cumsum(rng.exponential(1.0 / rate, len(trace)))
        for r, a in zip(trace, arr):
            r.arrival = float(a)
    rays = sum(r.width * r.height for r in trace)
    log(f"{label}: {len(trace)} requests, load {load}, rate {rate if rate is None else round(rate, 1)}/s "
        f"(plain mean service {mean_service * 1e3:.2f} ms)")
    def consume(r):  # a client takes the gradients and lets them go
        r.grads = None

    plain.on_finish = consume
    p_trace = copy.deepcopy(trace)
    p_wall = plain.run(p_trace)
    p_stats = latency_stats(p_trace)
    # one server per scene set and configuration, warmed once and kept across
    # loads and traces, as a serving process keeps its server (the plain
    # path's allocator cache is warm the same way)
    key = (id(scenes), args.serve_capacity, args.serve_quantum, args.serve_oldest_share, args.serve_streams,
           args.serve_record_cap, args.serve_backward)
    srv = _SERVERS.get(key)
    if srv is None:
        srv = RenderServer("cuda", memory_budget=64 << 30, tile_capacity=args.serve_capacity,
                           quantum=args.serve_quantum, oldest_share=args.serve_oldest_share,
                           streams=args.serve_streams, record_cap=args.serve_record_cap, max_wait=1e9,
                           backward=args.serve_backward)
        for name, scene in scenes.items():
            srv.add_scene(name, scene)
        srv.on_finish = consume
        srv.run(copy.deepcopy(warm))  # warm-up (graph capture, the allocator's cache)
        _SERVERS.clear()
        _SERVERS[key] = srv
    srv.max_wait = 1e9 if load == "saturation" else 2.0
    s_trace = copy.deepcopy(trace)
    srv.finished, srv.rejected = [], []
    s_wall = srv.run(s_trace)
    s_stats = latency_stats([r for r in s_trace if not r.rejected])
    # identical images: the served result equals the plain path's
    same = all(torch.equal(a.color, b.color) for a, b in zip(s_trace, p_trace) if not a.rejected)
    out = {
        "requests": len(trace), "rate_hz": None if rate is None else round(rate, 2), "load": load,
        "p50_ms": round(s_stats["p50_ms"], 3), "p99_ms": round(s_stats["p99_ms"], 3),
        "mrays_per_s": round(rays / s_wall / 1e6, 2),
        "plain_p50_ms": round(p_stats["p50_ms"], 3), "plain_p99_ms": round(p_stats["p99_ms"], 3),
        "plain_mrays_per_s": round(rays / p_wall / 1e6, 2),
        "rejected": len([r for r in s_trace if r.rejected]), "status": int(srv.status.item()),
        "tile_capacity": args.serve_capacity, "quantum": args.serve_quantum, "colour": srv.colour,
        "results_equal_plain": bool(same),
        "results_compared": len([r for r in s_trace if not r.rejected]),
        "wall_s": round(s_wall, 3), "plain_wall_s": round(p_wall, 3),
        "plain_mean_service_ms": round(mean_service * 1e3, 3),
    }
    fb = [r for r in s_trace if not r.rejected and r.mode == "forward_backward"]
    if fb:  # the forward+backward requests alone
        fs = latency_stats(fb)
        fp = latency_stats([r for r in p_trace if r.mode == "forward_backward"])
        out["forward_backward"] = {"requests": len(fb), "p50_ms": round(fs["p50_ms"], 3),
                                   "p99_ms": round(fs["p99_ms"], 3), "plain_p50_ms": round(fp["p50_ms"], 3),
                                   "plain_p99_ms": round(fp["p99_ms"], 3), "backward": srv.backward}
    log(f"{label}: p50 {out['p50_ms']} ms p99 {out['p99_ms']} ms vs plain p50 {out['plain_p50_ms']} "
        f"p99 {out['plain_p99_ms']} ms, rejected {out['rejected']}")
    return out


RESOLUTIONS = [(256, 256), (512, 512), (1024, 1024), (1920, 1080)]


def single_scene_trace(args):
    """One resident scene of --serve-cells cells (3M), forward requests,
    resolutions uniform over {256^2, 512^2, 1024^2, 1920x1080}, cameras on
    the radius-3 sphere.  Returns (scenes, request maker, workload)."""
    import numpy as np

    from powerfoam import workloads
    from powerfoam.serving import RenderRequest

    t0 = time.time()
    scenes = {"s": build_served_scene(args.serve_cells, args.seed + 1)}
    log(f"serving scene: {args.serve_cells} cells resident in {time.time() - t0:.1f}s")
    rng = np.random.default_rng(1000 + args.seed)

    def make(n):
        reqs = []
        for k in range(n):
            w, h = RESOLUTIONS[int(rng.integers(0, 4))]
            reqs.append(RenderRequest(scene="s", camera=workloads.sphere_cameras(rng, 1, w, h)[0], rid=k))
        return reqs

    workload = (f"{args.serve_cells}-cell GMM scene resident, SH3 colour pre-pass per request, forward "
                "requests, resolutions uniform over 256^2/512^2/1024^2/1920x1080, Poisson arrivals")
    return scenes, make, workload


def multi_scene_trace(args):
    """Several resident scenes under memory admission: cell counts
    log-uniform in [1e5, 2e6], each request picks a scene uniformly,
    resolutions as above, a quarter of the requests forward+backward
    against a random target image, a 64 GB budget.  Returns (scenes,
    request maker, workload, cell counts)."""
    import numpy as np
    import torch

    from powerfoam import workloads
    from powerfoam.serving import RenderRequest

    rng = np.random.default_rng(3000 + args.seed)
    sizes = np.exp(rng.uniform(np.log(1e5), np.log(2e6), args.serve_scenes)).astype(int)
    t0 = time.time()
    scenes = {f"s{k}": build_served_scene(int(n), args.seed + 10 + k) for k, n in enumerate(sizes)}
    log(f"serving scenes: {', '.join(str(int(n)) for n in sizes)} cells resident in {time.time() - t0:.1f}s")

    def make(n):
        reqs = []
        for k in range(n):
            w, h = RESOLUTIONS[int(rng.integers(0, 4))]
            name = f"s{int(rng.integers(0, len(sizes)))}"
            bwd = rng.uniform() < 0.25
            tgt = torch.from_numpy(rng.uniform(0, 1, (h, w, 3)).astype(np.float32)) if bwd else None
            reqs.append(RenderRequest(scene=name, camera=workloads.sphere_cameras(rng, 1, w, h)[0], rid=k,
                                      mode="forward_backward" if bwd else "forward", target=tgt))
        return reqs

    workload = (f"{len(sizes)} GMM scenes resident (cells log-uniform in [1e5, 2e6]), a scene picked "
                "uniformly per request, 25% forward+backward requests against random targets, "
                "resolutions uniform over 256^2/512^2/1024^2/1920x1080, Poisson arrivals, 64 GB budget")
    return scenes, make, workload, [int(n) for n in sizes]


def serve_bench(args):
    """Serving form, one resident scene (single_scene_trace) at --serve-load."""
    scenes, make, workload = single_scene_trace(args)
    out = serve_trace(scenes, make, args, "serving")
    _SERVERS.clear()
    out["workload"] = workload
    return out


def serve_multi_bench(args):
    """Serving form, several scenes under admission (multi_scene_trace) at
    --serve-load."""
    scenes, make, workload, sizes = multi_scene_trace(args)
    out = serve_trace(scenes, make, args, "serving (multi-scene)")
    _SERVERS.clear()
    out["scene_cells"] = sizes
    out["workload"] = workload
    return out


SWEEP_LOADS = (0.3, 0.6, 0.9, "saturation")


def serve_sweep(args):
    """Both serving traces at offered loads 0.3, 0.6 and 0.9 of the plain
    path's capacity and at saturation (every request submitted at once):
    p50/p99 latency and throughput of the server and the plain path, on
    --sweep-traces independent traces of --sweep-requests requests per load
    (a trace's p99 swings with its bursts), each row and the mean over the
    traces.  Prints one JSON object (not the benchmark line)."""
    import numpy as np
    import torch

    out = {"loads": [str(x) for x in SWEEP_LOADS], "traces_per_load": args.sweep_traces,
           "requests_per_trace": args.sweep_requests, "serve_backward": args.serve_backward}
    args.serve_requests = args.sweep_requests
    keys = ("p50_ms", "p99_ms", "mrays_per_s", "plain_p50_ms", "plain_p99_ms", "plain_mrays_per_s")
    for name, builder in (("single_scene", single_scene_trace), ("multi_scene", multi_scene_trace)):
        if name == "multi_scene" and args.serve_scenes <= 0:
            continue
        built = builder(args)
        scenes, make, workload = built[:3]
        cal = plain_service(scenes, make)
        rows, means = [], []
        for ld in SWEEP_LOADS:
            runs = [serve_trace(scenes, make, args, f"{name} load {ld} trace {k}", load=ld, calibration=cal)
                    for k in range(args.sweep_traces)]
            rows.extend(runs)
            mean = {"load": ld, **{k: round(float(np.mean([r[k] for r in runs])), 3) for k in keys}}
            fb = [r["forward_backward"] for r in runs if "forward_backward" in r]
            if fb:
                mean["forward_backward"] = {k: round(float(np.mean([f[k] for f in fb])), 3)
                                            for k in ("p50_ms", "p99_ms", "plain_p50_ms", "plain_p99_ms")}
            means.append(mean)
            log(f"{name} load {ld} mean over {len(runs)} traces: {mean}")
        out[name] = {"workload": workload, "rows": rows, "means": means}
        if len(built) > 3:
            out[name]["scene_cells"] = built[3]
        _SERVERS.clear()
        del scenes, make, built, cal
        torch.cuda.empty_cache()
    print(json.dumps(out), flush=True)


TRAFFIC_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "walk_gmm1m_1080p_hbm_traffic.md")


def profiled_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary of this workload, or None."""
    try:
        with open(TRAFFIC_FILE) as fh:
            for line in fh:
                parts = line.split()
                if len(parts) == 3 and parts[0] == "traffic_bytes" and parts[1] == kernel:
                    return float(parts[2])
    except OSError:
        return None
    return None


def copy_calibration(nbytes=1 << 30, reps=5):
    """A plain copy of known size: under a byte counter (FETCH_SIZE /
    WRITE_SIZE) each launch should read and write nbytes, which gives the
    counter's scale on this device."""
    import torch

    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": nbytes, "launches": reps}), flush=True)


def c1_bench(args):
    """The plain forward configuration in full: 4,096 cells uniform in
    [-1, 1]^3, weights U[0, 0.01], density lognormal(0, 1), one colour per
    cell, the Qhull-lift adjacency, one 128x128 60-degree camera at distance
    3 looking at the origin, seed 0.  Every ray on the GPU (median of 200
    launches) and on the host's cores with the oracle; sequences compared
    bit for bit and colours to 1e-4."""
    import numpy as np
    import torch

    from oracle import oracle as orc
    from powerfoam import raytrace as rt
    from powerfoam import workloads
    from powerfoam.adjacency import build_power_adjacency
    from powerfoam.camera import pack_cameras

    sc = workloads.uniform_scene(4096, seed=0)
    adj, off = build_power_adjacency(torch.from_numpy(sc["points"]), torch.from_numpy(sc["radii"]))
    adj, off = adj.numpy(), off.numpy()
    cam = workloads.front_camera(128, 128)
    words = pack_cameras([cam], [orc.start_cell(sc["points"], sc["radii"], cam.eye.numpy())])
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    pts, rad, adj_d, off_d = t(sc["points"]), t(sc["radii"]), t(adj), t(off)
    faces = rt.walk_faces(pts, rad, adj_d, off_d)
    scene = rt.WalkScene(pts, rad, t(sc["density"]), t(sc["normals"]), t(sc["texel_sites"]), t(sc["texel_rgb"]),
                         t(sc["texel_height"]), adj_d, off_d, faces)
    cams = t(words)
    for _ in range(5):
        rt.launch_walk_forward(scene, cams, 128, 128, want_depth=False)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ms = []
    for _ in range(200):
        ev[0].record()
        rt.launch_walk_forward(scene, cams, 128, 128, want_depth=False)
        ev[1].record()
        torch.cuda.synchronize()
        ms.append(ev[0].elapsed_time(ev[1]))
    g = rt.launch_walk_forward(scene, cams, 128, 128, want_sequences=True)
    torch.cuda.synchronize()
    fh = faces.cpu().view(torch.int16).numpy().view(np.uint16)
    cs = orc.Scene(sc["points"], sc["radii"], sc["density"], sc["normals"], sc["texel_sites"], sc["texel_rgb"],
                   sc["texel_height"], adj, off, fh)
    nth = orc.num_threads()
    orc.walk_forward(cs, words, nthreads=nth)
    cpu_s = []
    for _ in range(5):
        t0 = time.perf_counter()
        c = orc.walk_forward(cs, words, nthreads=nth)
        cpu_s.append(time.perf_counter() - t0)
    same = bool(np.array_equal(g["visits"].cpu().numpy(), c["visits"]) and
                np.array_equal(g["seq_hash"].cpu().numpy().view(np.uint64), c["seq_hash"]))
    err = float(np.abs(g["color"].cpu().numpy() - c["color"]).max())
    gms, cs_ = float(np.median(ms)), float(np.median(cpu_s))
    out = {"workload": "plain walk forward, 4096 uniform cells, 128x128, 60 deg, seed 0 (every ray)",
           "rays": 128 * 128, "gpu_ms": round(gms, 4), "gpu_mrays_per_s": round(128 * 128 / (gms * 1e-3) / 1e6, 2),
           "cpu_ms": round(cs_ * 1e3, 3), "cpu_mrays_per_s": round(128 * 128 / cs_ / 1e6, 4), "cpu_cores": nth,
           "cpu_kind": "port", "sequences_equal": same, "max_abs_rgb": err,
           "mean_visits_per_ray": round(float(c["visits"].mean()), 2)}
    log(f"plain forward 4096 cells 128x128: GPU {gms:.3f} ms, CPU {cs_ * 1e3:.2f} ms on {nth} threads; "
        f"sequences equal {same}, max |rgb| {err:.1e}")
    return out


def cpu_baseline(sc, adj, off, faces, texel_rgb, cams_dev, fwd, gc, seconds):
    """The CPU restatement (oracle/) timed on this host's cores: walk forward
    + backward on a bounded, evenly spaced sample of the same frame's rays."""
    import numpy as np
    import torch

    from oracle import oracle as orc

    nth = orc.num_threads()
    cams = cams_dev.cpu().numpy()
    W, H = int(cams[0, 10]), int(cams[0, 11])
    fh = faces.cpu().view(torch.int16).numpy().view(np.uint16)
    cs = orc.Scene(sc["points"], sc["radii"], sc["density"], sc["normals"], sc["texel_sites"],
                   texel_rgb[0].cpu().numpy(), sc["texel_height"], adj, off, fh)
    color = fwd["color"].cpu().numpy()
    gcol = gc.cpu().numpy()

    def run(n):
        rays = np.linspace(0, W * H - 1, n).astype(np.int64)
        t0 = time.perf_counter()
        orc.walk_forward(cs, cams, rays=rays, nthreads=nth)
        g = np.zeros_like(gcol).reshape(-1, 3)
        g[rays] = gcol.reshape(-1, 3)[rays]
        orc.walk_backward(cs, cams, color, g.reshape(gcol.shape), nthreads=nth)
        return time.perf_counter() - t0

    probe = 4096
    tp = run(probe)
    n = int(min(W * H, max(probe, probe * seconds / max(tp, 1e-6))))
    log(f"cpu baseline: {nth} threads, probe {probe} rays in {tp:.2f}s; sample {n} rays")
    tt = run(n)
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    return {"value": round(n / tt / 1e6, 5), "unit": "Mrays/s", "cores": nth, "kind": "port",
            "host": {"nproc": os.cpu_count(), "cpu_model": model},
            "sample": f"walk fwd+bwd (C++ restatement, OpenMP) on {n} evenly spaced rays of the same 1080p frame, "
                      f"{tt:.1f}s; SH pre-pass excluded"}


if __name__ == "__main__":
    main()
