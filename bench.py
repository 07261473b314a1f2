"""Benchmark: primary+shadow Mrays/s at 1920x1080 on the 90M-point depth-12
octree (config 3: planes 60%, 200 boxes, 50 spheres, a 10M-point cluster of
15 ellipsoids; 2 point lights) on one MI355X, exact traversal, and the same
frames with the learned shadow surrogate.

A step is one frame: one camera pose of a seeded camera path rendered by the
hot path (primary trace, one shadow trace per light for every shaded pixel,
shading), inputs resident in HBM. value = (primary + shadow rays) / wall time
of the K timed exact frames.

Roofline. Every traversal launch is bracketed by HIP events on the library's
stream (qbt_profile_*), per kernel kind. Algorithmic bytes: 32 B per node
expansion (one 8 x int32 child block) + 4 B per primary leaf hit (its model
index) for the primary kernel; 32 B per expansion for a shadow kernel; per
frame additionally 24 B of colour + normal per primary hit and the outputs
(hit 4 + leaf 4 + RGBA8 4 + 1 per light per pixel). The roofline object is
for the dominant kernel (largest share of the frame's kernel time);
roofline.frame gives the whole frame. Peak: 8 TB/s HBM3E.

Surrogate leg: the MLP is trained online on the GPU (exact renders of a
separate seeded camera path with sample collection, Adam steps after each
frame), then the timed frames are rendered again with QBT_RENDER_SURROGATE
(fallback band [0.1, 0.9]); its Mrays/s counts the same rays (predicted or
traced), and its shadow-bit mismatch against the exact frames is reported.

CPU baseline: the C oracle (restatement of the reference traversal) on whole
frames of the timed camera path on the host's cores, each frame checked
bit-exactly against the GPU frame.

More legs, reported beside the headline: the same scene at 3840x2160
(uhd_3840x2160); the same scene with a populated dynamic tree walked in
lockstep (dynamic_tree; with the time to rebuild a 2M- and a 10M-point
dynamic tree from points already on the GPU, as the reference rebuilds its
dynamic tree every frame); and batched views (batch_views): the 90M-point scene with 4
lights, 64 views of resolutions drawn from 320x240 ... 3840x2160 at random
poses rendered as one hipGraph-captured batch with the autotuner's tuning,
two of its views checked bit for bit against single-view renders.

Prints ONE JSON line on stdout; progress goes to stderr.

``python bench.py --summarize-pmc DIR...`` instead reads rocprofv3 counter
CSVs (FETCH_SIZE / WRITE_SIZE / TCC hits of a bench run) and writes the HBM
traffic summary the roofline line quotes.
"""
import argparse
import csv
import glob
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
COPY_BYTES = 1 << 30
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pmc_traversal_config3_1080p.json")
KIND_PRIMARY, KIND_SHADOW, KIND_SURROGATE, KIND_FRAME = 0, 1, 2, 3


def host_cpu():
    """model name and logical CPU count of the host (the CPU baseline's machine)"""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "logical_cpus": os.cpu_count()}


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=3)
    ap.add_argument("--npoints", type=int, default=0, help="0 = the config's point count")
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--pose-seed", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="CPU-baseline sample budget")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = min(16, cores)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--sur-frames", type=int, default=256, help="exact training frames")
    ap.add_argument("--sur-samples", type=int, default=1 << 24, help="training samples collected")
    ap.add_argument("--sur-steps", type=int, default=32, help="Adam steps after each training frame")
    ap.add_argument("--sur-final", type=int, default=2000, help="Adam steps after the last frame")
    ap.add_argument("--sur-batch", type=int, default=4096)
    ap.add_argument("--sur-lr", type=float, default=2e-3)
    ap.add_argument("--sur-final-batch", type=int, default=0, help="batch of the final steps (0 = --sur-batch)")
    ap.add_argument("--sur-final-lr", type=float, default=0.0, help="rate of the final steps (0 = --sur-lr)")
    ap.add_argument("--sur-decay", type=float, default=0.0,
                    help="final steps: cosine decay of the rate to this fraction (0 = constant)")
    ap.add_argument("--sur-band", type=float, nargs=2, default=None,
                    help="fixed uncertainty band (default: calibrated on held-out frames to --sur-target)")
    ap.add_argument("--sur-target", type=float, default=0.01,
                    help="shadow-bit mismatch the calibrated band allows on the held-out frames")
    ap.add_argument("--sur-cal-frames", type=int, default=8, help="held-out exact frames of the band calibration")
    ap.add_argument("--uhd-frames", type=int, default=5, help="frames of the 3840x2160 leg (0 = skip)")
    ap.add_argument("--dyn-frames", type=int, default=5, help="frames of the dynamic-tree leg (0 = skip)")
    ap.add_argument("--batch-views", type=int, default=64, help="views of the batched leg (0 = skip)")
    ap.add_argument("--edit-shots", type=int, default=5, help="shots of the edit leg (0 = skip)")
    ap.add_argument("--c5-frames", type=int, default=10,
                    help="timed frames of the config-5 surrogate leg (0 = skip)")
    ap.add_argument("--batch-reps", type=int, default=3, help="timed batches of the batched leg")
    ap.add_argument("--traffic", default=TRAFFIC_FILE)
    ap.add_argument("--summarize-pmc", nargs="*", default=None, metavar="DIR")
    ap.add_argument("--summarize-mfma", nargs="*", default=None, metavar="DIR",
                    help="print the MLP kernels' matrix-core counters (rocprofv3 output folders)")
    ap.add_argument("--path", type=int, default=None, help="render path (0 fused, 1 wavefront); default: library's")
    ap.add_argument("--block", type=int, default=None, help="threads per block (64/128/256)")
    ap.add_argument("--refill", type=int, default=None, help="idle lanes that trigger a wave refill (1..64)")
    ap.add_argument("--layout", type=int, default=None, help="0 compact node layout, 1 direct")
    ap.add_argument("--stack", type=int, default=None, help="0 LDS traversal stacks, 1 global memory")
    ap.add_argument("--reorder", type=int, default=None, help="1 shadow rays in direction-octant order")
    ap.add_argument("--graph", type=int, default=None, help="1 replay captured hipGraphs")
    ap.add_argument("--strip", type=int, default=None, help="tile order: 0 row-major, S strips of S tiles")
    ap.add_argument("--merge", type=int, default=None,
                    help="camera and shadow rays of exact renders in one launch (k_frame): 0 never, "
                         "1 where it pays (library default), 2 always")
    return ap.parse_args()


TRAVERSAL_KERNELS = ("k_render", "k_primary", "k_shadow", "k_frame")


def summarize_pmc(dirs, out_path):
    """Average per-dispatch counters of each traversal kernel (non-counting
    variants) and of the copy calibration kernel; byte counters are scaled by
    the copy kernel's known traffic."""
    vals = {}
    for d in dirs:
        for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            with open(f) as fh:
                for row in csv.DictReader(fh):
                    name = row.get("Kernel_Name", "")
                    kern = None
                    if "k_copy16" in name:
                        kern = "copy"
                    m = re.search(r"\b(k_render|k_primary|k_shadow|k_frame)<([^>]*)>", name)
                    if m:  # not the counting variant: COUNT is k_render's 3rd, the others' 4th argument
                        targs = [a.strip() for a in m.group(2).split(",")]
                        ci = 2 if m.group(1) == "k_render" else 3
                        if len(targs) > ci and targs[ci] == "false":
                            kern = m.group(1)
                    if kern is None:
                        continue
                    key = (kern, row["Counter_Name"])
                    vals.setdefault(key, {})[(f, row.get("Dispatch_Id"))] = float(row["Counter_Value"])
    avg = {k: float(np.mean(list(v.values()))) for k, v in vals.items()}
    n = {k: len(v) for k, v in vals.items()}
    res = {"copy_bytes": COPY_BYTES}
    fc, wc = avg.get(("copy", "FETCH_SIZE")), avg.get(("copy", "WRITE_SIZE"))
    ff = (COPY_BYTES / 1024.0) / fc if fc else None
    wf = (COPY_BYTES / 1024.0) / wc if wc else None
    res["fetch_scale"], res["write_scale"] = ff, wf
    per_kernel = {}
    for k in TRAVERSAL_KERNELS:
        e = {}
        if (k, "FETCH_SIZE") in avg and ff:
            e["fetch_bytes_per_launch"] = avg[(k, "FETCH_SIZE")] * 1024.0 * ff
            e["launches"] = n[(k, "FETCH_SIZE")]
        if (k, "WRITE_SIZE") in avg and wf:
            e["write_bytes_per_launch"] = avg[(k, "WRITE_SIZE")] * 1024.0 * wf
        if "fetch_bytes_per_launch" in e and "write_bytes_per_launch" in e:
            e["hbm_bytes_per_launch"] = e["fetch_bytes_per_launch"] + e["write_bytes_per_launch"]
        h, m = avg.get((k, "TCC_HIT_sum")), avg.get((k, "TCC_MISS_sum"))
        if h is not None and m is not None and h + m > 0:
            e["l2_hit_rate"] = h / (h + m)
        if e:
            per_kernel[k] = e
    res["kernels"] = per_kernel
    res["counters_kib_or_count"] = {"%s:%s" % k: v for k, v in avg.items()}
    res["dispatches"] = {"%s:%s" % k: v for k, v in n.items()}
    os.makedirs(os.path.dirname(out_path), exist_ok=True)
    with open(out_path, "w") as fh:
        json.dump(res, fh, indent=1, sort_keys=True)
    print(json.dumps(res, indent=1, sort_keys=True))


MLP_KERNELS = ("k_mlp_shadow", "k_mlp_infer", "k_mlp_train")
MFMA_BF16_PEAK_TFLOPS = 2500.0
SIMDS = 1024  # 256 CUs x 4 SIMDs


def summarize_mfma(dirs):
    """Matrix-core counters of the MLP kernels: per-dispatch averages of
    SQ_INSTS_VALU_MFMA_MOPS_BF16 (units of 512 flops), SQ_INSTS_MFMA,
    SQ_VALU_MFMA_BUSY_CYCLES and GRBM_GUI_ACTIVE from counter_collection.csv,
    against the average durations in kernel_stats.csv of a --kernel-trace
    --stats run. Prints a markdown table."""
    vals, durs = {}, {}
    pat = re.compile(r"\b(%s)\b" % "|".join(MLP_KERNELS))
    for d in dirs:
        for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            with open(f) as fh:
                for row in csv.DictReader(fh):
                    m = pat.search(row.get("Kernel_Name", ""))
                    if m:
                        key = (m.group(1), row["Counter_Name"])
                        vals.setdefault(key, {})[(f, row.get("Dispatch_Id"))] = float(row["Counter_Value"])
        for f in glob.glob(os.path.join(d, "**", "*kernel_stats.csv"), recursive=True):
            with open(f) as fh:
                for row in csv.DictReader(fh):
                    m = pat.search(row.get("Name", ""))
                    if m:
                        c, tot = int(row["Calls"]), float(row["TotalDurationNs"])
                        a = durs.setdefault(m.group(1), [0, 0.0])
                        a[0], a[1] = a[0] + c, a[1] + tot
    avg = {k: float(np.mean(list(v.values()))) for k, v in vals.items()}
    n = {k: len(v) for k, v in vals.items()}
    lines = ["| kernel | dispatches | bf16 flops / dispatch (MOPS x 512) | MFMA insts / dispatch | avg duration (us) "
             "| achieved TFLOP/s | frac of 2.5 PFLOP/s | MFMA busy (busy / (GUI_ACTIVE/8 x 1024 SIMDs)) |",
             "|---|---|---|---|---|---|---|---|"]
    for k in MLP_KERNELS:
        mops = avg.get((k, "SQ_INSTS_VALU_MFMA_MOPS_BF16"))
        if mops is None:
            continue
        flops = mops * 512.0
        ins = avg.get((k, "SQ_INSTS_MFMA"))
        busy, gui = avg.get((k, "SQ_VALU_MFMA_BUSY_CYCLES")), avg.get((k, "GRBM_GUI_ACTIVE"))
        dc = durs.get(k)
        us = dc[1] / dc[0] / 1e3 if dc and dc[0] else None
        tf = flops / (us * 1e-6) / 1e12 if us else None
        util = busy / (gui / 8.0 * SIMDS) if busy is not None and gui else None
        f = lambda x, fmt: (fmt % x) if x is not None else "n/a"
        lines.append(f"| {k} | {n[(k, 'SQ_INSTS_VALU_MFMA_MOPS_BF16')]} | {flops:.4g} | {f(ins, '%.4g')} | "
                     f"{f(us, '%.1f')} | {f(tf, '%.1f')} | {f(tf / MFMA_BF16_PEAK_TFLOPS if tf else None, '%.4f')} | "
                     f"{f(util, '%.4f')} |")
    print("\n".join(lines))


def load_traffic(path):
    if not os.path.exists(path):
        return {}
    try:
        with open(path) as f:
            return json.load(f).get("kernels", {})
    except (OSError, ValueError):
        return {}


def roofline(args, prof, K, W, H, nl, exps, exps_p, leafhits, copy_gbps, with_traffic=True):
    """Per-kernel and per-frame algorithmic bytes against the measured kernel times."""
    km, kl = prof["kind_ms"], prof["kind_launches"]
    e_p = float(np.mean(exps_p))
    e_s = float(np.mean(exps)) - e_p
    h = float(np.mean(leafhits))
    kern = {}
    if kl[KIND_PRIMARY]:
        name = "k_primary" if kl[KIND_SHADOW] else "k_render"
        b = 32.0 * e_p + 4.0 * h if name == "k_primary" else 32.0 * (e_p + e_s) + 28.0 * h + W * H * (12 + nl)
        kern[name] = {"alg_bytes_per_launch": b, "avg_ms": km[KIND_PRIMARY] / kl[KIND_PRIMARY],
                      "share": km[KIND_PRIMARY], "launches": kl[KIND_PRIMARY]}
    if kl[KIND_SHADOW]:
        kern["k_shadow"] = {"alg_bytes_per_launch": 32.0 * e_s * K / kl[KIND_SHADOW],
                            "avg_ms": km[KIND_SHADOW] / kl[KIND_SHADOW],
                            "share": km[KIND_SHADOW], "launches": kl[KIND_SHADOW]}
    if kl[KIND_FRAME]:  # camera and shadow rays in one launch: both kernels' bytes
        kern["k_frame"] = {"alg_bytes_per_launch": (32.0 * (e_p + e_s) + 4.0 * h) * K / kl[KIND_FRAME],
                           "avg_ms": km[KIND_FRAME] / kl[KIND_FRAME],
                           "share": km[KIND_FRAME], "launches": kl[KIND_FRAME]}
    traffic = load_traffic(args.traffic) if with_traffic else {}
    for name, e in kern.items():
        e["achieved_gbs"] = e["alg_bytes_per_launch"] / (e["avg_ms"] * 1e-3) / 1e9
        e["frac"] = e["achieved_gbs"] / HBM_PEAK_GBS
        t = traffic.get(name, {})
        e["traffic"] = round(t["hbm_bytes_per_launch"]) if "hbm_bytes_per_launch" in t else None
        e["l2_hit_rate"] = t.get("l2_hit_rate")
    dom = max(kern, key=lambda k: kern[k]["share"])
    d = kern[dom]
    frame_bytes = 32.0 * (e_p + e_s) + 28.0 * h + W * H * (12 + nl)
    frame_ms = (km[KIND_PRIMARY] + km[KIND_SHADOW] + km[KIND_FRAME]) / K
    roof = {"bound": "hbm", "achieved": round(d["achieved_gbs"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(d["frac"], 4), "traffic": d["traffic"], "kernel": dom,
            "alg_bytes_per_launch": round(d["alg_bytes_per_launch"]), "kernel_avg_ms": round(d["avg_ms"], 4),
            "l2_hit_rate": d["l2_hit_rate"], "copy_kernel_gbps": round(copy_gbps, 1),
            "kernels": {k: {kk: (round(vv, 4) if isinstance(vv, float) else vv) for kk, vv in e.items()
                            if kk != "share"} for k, e in kern.items()},
            "frame": {"alg_bytes": round(frame_bytes), "traversal_ms": round(frame_ms, 4),
                      "achieved": round(frame_bytes / (frame_ms * 1e-3) / 1e9, 1),
                      "frac": round(frame_bytes / (frame_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
            "expansions_per_frame": {"primary": round(e_p), "shadow": round(e_s)}}
    return roof


def surrogate_leg(args, q, ctx, sc, views, Wm, K, mask, exact_wall, exact_rays, queries_per_frame):
    from qubatron_amd import surrogate as S
    W, H = args.width, args.height
    lights = sc.lights
    tpos, tang = sc.poses(args.sur_frames, seed=args.pose_seed + 7919)
    tviews = [q.view(tpos[i], tang[i], W, H, sc.basecube, sc.depth) for i in range(args.sur_frames)]
    stride = max(1, int(round(args.sur_frames * queries_per_frame / args.sur_samples)))
    t0 = time.time()
    info = S.train_online(ctx, tviews, lights, sc.basecube, nsamples=args.sur_samples, stride=stride,
                          steps_per_frame=args.sur_steps, batch=args.sur_batch, lr=args.sur_lr, seed=args.seed,
                          final_steps=args.sur_final, final_batch=args.sur_final_batch or None,
                          final_lr=args.sur_final_lr or None, decay=args.sur_decay)
    ctx.synchronize()
    train_s = time.time() - t0
    log(f"surrogate: trained on {info['samples']} samples from {args.sur_frames} frames (stride {stride}), "
        f"{info['steps']} Adam steps in {train_s:.2f}s, last loss {info['loss_last']:.4f}")
    cal = None
    if args.sur_band is None:  # the widest band within the mismatch target on frames not trained on
        cpos, cang = sc.poses(args.sur_cal_frames, seed=args.pose_seed + 4243)
        cviews = [q.view(cpos[i], cang[i], W, H, sc.basecube, sc.depth) for i in range(args.sur_cal_frames)]
        cal = S.calibrate_band(ctx, cviews, lights, sc.basecube, target=args.sur_target)
        cal["pose_seed"] = args.pose_seed + 4243
        band = cal["band"]
        log(f"surrogate: band [{band[0]:.3f}, {band[1]:.3f}] calibrated on {args.sur_cal_frames} held-out frames "
            f"({cal['samples']} queries): decided {cal['held_out_decided']:.3f}, mismatch {cal['held_out_mismatch']:.4f}")
    else:
        band = list(args.sur_band)
    ctx.surrogate_set_band(*band)
    smask = mask | q.RENDER_SURROGATE
    for i in range(Wm):
        ctx.render(views[i], lights, smask)
    ctx.synchronize()
    ctx.profile_begin()
    ctx.synchronize()
    t0 = time.perf_counter()
    for i in range(Wm, K + Wm):
        ctx.render(views[i], lights, smask)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    prof = ctx.profile_end()
    # untimed: mismatch against the exact frames, predicted share
    bad = tot = pred = traced = 0
    resolved = 0
    fields = ("flags", "shadow")
    for i in range(Wm, K + Wm):
        ctx.render(views[i], lights, q.OUT_FLAGS | q.OUT_SHADOW)
        ex = ctx.read_frame(0, fields=fields)
        ctx.render(views[i], lights, q.OUT_FLAGS | q.OUT_SHADOW | q.RENDER_SURROGATE)
        su = ctx.read_frame(0, fields=fields)
        st = ctx.stats()
        pred += st["predicted"]
        traced += st["shadow_rays"]
        resolved += st["primary_rays"] + st["predicted"] + st["shadow_rays"]
        b, t = S.shadow_mismatch(ex, su, len(lights))
        bad, tot = bad + b, tot + t
    # the standalone inference entry point (k_mlp_infer) on 4 M queries: hit
    # points uniform in the cube, random unit normals, the scene's lights
    rng = np.random.default_rng(args.seed + 5)
    nq = 1 << 22
    bx, by, bz, bw = (float(v) for v in sc.basecube)
    ip = (np.array([bx, by - bw, bz - bw], np.float32) + rng.random((nq, 3), np.float32) * np.float32(bw))
    nv = rng.normal(size=(nq, 3)).astype(np.float32)
    nv /= np.linalg.norm(nv, axis=1, keepdims=True)
    lq = np.asarray(lights, np.float32)[np.arange(nq) % len(lights)]
    ctx.synchronize()
    t0 = time.perf_counter()
    pq = ctx.surrogate_predict(ip, nv, lq, sc.basecube)
    infer_s = time.perf_counter() - t0
    value = exact_rays / wall / 1e6
    km, kl = prof["kind_ms"], prof["kind_launches"]
    res = {"value": round(value, 3), "unit": "Mrays/s", "speedup_vs_exact": round(exact_wall / wall, 4),
           "ms_per_step": round(wall * 1e3 / K, 4), "mismatch": round(bad / max(tot, 1), 6),
           "mismatched_queries": bad, "shadow_queries": tot,
           "predicted_fraction": round(pred / max(pred + traced, 1), 4),
           "rays_resolved_per_frame": round(resolved / K), "band": [round(b, 4) for b in band],
           "band_calibration": cal,
           "kernel_ms_per_frame": {"primary": round(km[KIND_PRIMARY] / K, 4),
                                   "shadow": round(km[KIND_SHADOW] / K, 4),
                                   "mlp_decision": round(km[KIND_SURROGATE] / K, 4)},
           "mlp_decision_launches": kl[KIND_SURROGATE],
           "predict_entry": {"queries": nq, "seconds_with_copies": round(infer_s, 4),
                             "mean_p": round(float(np.mean(pq)), 4)},
           "train": {"frames": args.sur_frames, "samples": info["samples"], "stride": stride,
                     "lit_fraction": round(info["lit_fraction"], 4), "adam_steps": info["steps"],
                     "batch": args.sur_batch, "lr": args.sur_lr, "final_steps": args.sur_final,
                     "final_batch": args.sur_final_batch or args.sur_batch,
                     "final_lr": args.sur_final_lr or args.sur_lr, "decay": args.sur_decay,
                     "loss_last": round(info["loss_last"], 5),
                     "seconds": round(train_s, 2), "pose_seed": args.pose_seed + 7919},
           "workload": f"same scene and timed camera path as the exact form, {W}x{H}"}
    log(f"surrogate: {value:.1f} Mrays/s ({res['speedup_vs_exact']:.3f}x exact), mismatch {res['mismatch']:.5f}, "
        f"predicted {res['predicted_fraction']:.3f}")
    return res


def count_frames(q, ctx, views, lights, mask):
    """Untimed counting renders: per view, rays, expansions (all, primary)
    and primary leaf hits."""
    rays, exps, exps_p, leafhits = [], [], [], []
    for v in views:
        ctx.render(v, lights, mask | q.OUT_FLAGS | q.RENDER_COUNT)
        st = ctx.stats()
        fl = ctx.read_frame(0, fields=("flags",))["flags"]
        rays.append(st["primary_rays"] + st["shadow_rays"])
        exps.append(st["expansions"])
        exps_p.append(st["expansions_primary"])
        leafhits.append(int(np.count_nonzero(fl & q.PX_LEAF)))
    return rays, exps, exps_p, leafhits


def uhd_leg(args, q, ctx, sc, mask, copy_gbps):
    """The same scene and lights at 3840x2160: K frames of a seeded path."""
    W, H, K = 3840, 2160, args.uhd_frames
    lights = sc.lights
    pos, ang = sc.poses(K + 1, seed=args.pose_seed + 2160)
    views = [q.view(pos[i], ang[i], W, H, sc.basecube, sc.depth) for i in range(K + 1)]
    rays, exps, exps_p, leafhits = count_frames(q, ctx, views[1:], lights, mask)
    ctx.render(views[0], lights, mask)  # warmup
    ctx.synchronize()
    ctx.profile_begin()
    ctx.synchronize()
    t0 = time.perf_counter()
    for v in views[1:]:
        ctx.render(v, lights, mask)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    prof = ctx.profile_end()
    roof = roofline(args, prof, K, W, H, len(lights), exps, exps_p, leafhits, copy_gbps, with_traffic=False)
    value = float(sum(rays)) / wall / 1e6
    log(f"3840x2160: {value:.1f} Mrays/s, {wall * 1e3 / K:.2f} ms/frame, {roof['kernel']} at "
        f"{roof['frac']:.4f} of HBM peak")
    return {"value": round(value, 3), "unit": "Mrays/s", "ms_per_step": round(wall * 1e3 / K, 4), "steps": K,
            "rays_per_frame": round(float(np.mean(rays))), "pose_seed": args.pose_seed + 2160,
            "roofline": {k: roof[k] for k in ("kernel", "achieved", "frac", "kernel_avg_ms", "frame")}}


def edit_leg(args, q, ctx, sc, pnt, nrm, nodes, mask):
    """The reference's shot edit (modelutil_punch_hole, modelutil.c:345-500) on
    the static tree, then the next frame: every point within `half` of a hit
    point is removed (its parent node uploaded, 48 bytes) and re-inserted up
    to 10 units into the surface along its normal (every node it creates
    uploaded), as many calls as the reference makes. Timed: the upload calls,
    and the first frame after them (layout patch included) against a frame
    without edits and a frame after a whole-array upload (full relayout)."""
    W, H, K = args.width, args.height, args.edit_shots
    lights = sc.lights
    pos, ang = sc.poses(K + 1, seed=args.pose_seed)
    views = [q.view(pos[i], ang[i], W, H, sc.basecube, sc.depth) for i in range(K + 1)]
    tree = q.HostOctree.from_nodes(nodes, sc.basecube, sc.depth)
    bt = q._ffi.BUFFER_STATIC_OCTREE

    def frame_ms(v):
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.render(v, lights, mask)
        ctx.synchronize()
        return (time.perf_counter() - t0) * 1e3

    ctx.render(views[0], lights, mask | q.OUT_ISP | q.OUT_FLAGS)
    fr = ctx.read_frame(0, fields=("isp", "flags"))
    shaded = np.argwhere(fr["flags"] & q.PX_SHADED)
    plain = [frame_ms(v) for v in views[1:]]
    rel0, pat0, _ = ctx.layout_stats()
    shots, upload_ms, uploads, edit_frames = [], 0.0, 0, []
    half = np.float32(15.0)  # guntype 2: dist up to 30 units, half 15 (modelutil.c:378-379)
    for k in range(K):
        cy, cx = shaded[(k * 7919) % len(shaded)]
        centre = fr["isp"][cy, cx, :3].astype(np.float32)
        d = pnt - centre
        near = np.flatnonzero(((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]) < half * half)
        ranges, moved = [], []
        for i in near:  # remove (octree_remove_point), one node upload each
            m, o = tree.remove_point(pnt[i])
            if o >= 0:
                ranges.append((o, tree.len))
                dist = np.float32(np.sqrt(float(((pnt[i] - centre) ** 2).sum())))
                rat = (half - dist) / half * np.float32(-10.0)
                moved.append((m, (pnt[i] + nrm[m] * rat).astype(np.float32)))
        created = []
        for m, p_ in moved:  # re-insert (octree_insert_point with octindarr)
            oa = np.zeros(sc.depth + 1, np.int32)
            tree.insert_point(m, p_, octindarr=oa)
            created.append([int(o) for o in oa[oa > 0]])
        nv = tree.nodes_view()
        ctx.synchronize()
        t0 = time.perf_counter()
        for o, ln in ranges:
            ctx.upload(nv, q._ffi.TYPE_INT, ln * 48, 16, o * 48, (o + 1) * 48, bt)
        for cl in created:
            for o in cl:
                ctx.upload(nv, q._ffi.TYPE_INT, tree.len * 48, 16, o * 48, (o + 1) * 48, bt)
        t1 = time.perf_counter()
        upload_ms += (t1 - t0) * 1e3
        uploads += len(ranges) + sum(len(c) for c in created)
        edit_frames.append(frame_ms(views[1 + k]))
        shots.append({"removed": len(ranges), "nodes_created": sum(len(c) for c in created)})
    rel1, pat1, recs = ctx.layout_stats()
    # the same frame after a whole-array upload: a full relayout of the tree
    nv = tree.nodes_view()
    ctx.upload(nv, q._ffi.TYPE_INT, tree.len * 48, 16, 0, tree.len * 48, bt)
    ctx.synchronize()
    relayout_frame = frame_ms(views[1])
    rel2 = ctx.layout_stats()[0]
    res = {"shots": shots, "upload_calls": uploads, "upload_us_per_call": round(upload_ms * 1e3 / max(uploads, 1), 2),
           "plain_frame_ms": round(float(np.median(plain)), 4),
           "edit_frame_ms": [round(x, 4) for x in edit_frames],
           "frame_after_whole_upload_ms": round(relayout_frame, 3),
           "layout_patches": pat1 - pat0, "relayouts_during_edits": rel1 - rel0,
           "relayouts_after_whole_upload": rel2 - rel1, "records": recs,
           "workload": f"{K} shots (radius {float(half)}) on the static tree, each followed by one {W}x{H} frame"}
    log(f"edits: {K} shots, {uploads} node uploads at {res['upload_us_per_call']} us per call; frame after a shot "
        f"{np.median(edit_frames):.2f} ms (plain {res['plain_frame_ms']:.2f} ms, after a whole-array upload "
        f"{relayout_frame:.1f} ms); patches {pat1 - pat0}, relayouts {rel1 - rel0}")
    return res


def dynamic_leg(args, q, ctx, sc, mask, copy_gbps):
    """The same scene with a populated dynamic tree walked in lockstep with
    the static one (octree_fsh.c:223-241, 313-317): a seeded 2M-point shell
    (a standing figure, 240 x 600 x 240 units) near the cube centre, its own
    colours and normals; K frames of the timed path at 1920x1080."""
    W, H, K = args.width, args.height, args.dyn_frames
    rng = np.random.default_rng(args.seed + 77)
    m = 2000000
    u = rng.normal(size=(m, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    dp = (np.array([900.0, 700.0, 900.0]) + u * np.array([120.0, 300.0, 120.0])).astype(np.float32)
    ctx.upload_points(dp, dynamic=True)
    ctx.upload_model(rng.uniform(0.0, 1.0, (m, 3)).astype(np.float32), u.astype(np.float32), dynamic=True)
    dn = ctx.build_octree(sc.basecube, sc.depth, m, dynamic=True)
    lights = sc.lights
    pos, ang = sc.poses(K + 1, seed=args.pose_seed)
    views = [q.view(pos[i], ang[i], W, H, sc.basecube, sc.depth) for i in range(K + 1)]
    rays, exps, exps_p, leafhits = count_frames(q, ctx, views[1:], lights, mask)
    ctx.render(views[0], lights, mask)
    ctx.synchronize()
    ctx.profile_begin()
    ctx.synchronize()
    t0 = time.perf_counter()
    for v in views[1:]:
        ctx.render(v, lights, mask)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    prof = ctx.profile_end()
    dual = bool(ctx.stats()["dual"])
    # per-frame rebuild of the dynamic tree from its resident points (the
    # reference rebuilds its dynamic tree on the CPU every frame,
    # qubatron.c:439-452): this figure, then a 10M-point one
    rebuild = {}
    for npt in (m, 10000000):
        if npt != m:
            u2 = rng.normal(size=(npt, 3))
            u2 /= np.linalg.norm(u2, axis=1, keepdims=True)
            ctx.upload_points((np.array([900.0, 700.0, 900.0]) + u2 * np.array([120.0, 300.0, 120.0]))
                              .astype(np.float32), dynamic=True)
            ctx.upload_model(rng.uniform(0.0, 1.0, (npt, 3)).astype(np.float32), u2.astype(np.float32), dynamic=True)
            del u2
        ctx.build_octree(sc.basecube, sc.depth, npt, dynamic=True)  # warm
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(K):
            nodes_b = ctx.build_octree(sc.basecube, sc.depth, npt, dynamic=True)
        ctx.synchronize()
        rebuild[str(npt)] = {"ms": round((time.perf_counter() - t0) * 1e3 / K, 3), "nodes": nodes_b}
        log(f"dynamic tree rebuild from resident points: {npt} points -> {nodes_b} nodes in "
            f"{rebuild[str(npt)]['ms']:.2f} ms")
    # the reference's per-frame sequence with the dynamic tree rebuilt on the
    # GPU: rebuild from the resident points, then the frame, timed together
    ctx.build_octree(sc.basecube, sc.depth, npt, dynamic=True)
    ctx.render(views[0], lights, mask)
    ctx.synchronize()
    t0 = time.perf_counter()
    for v in views[1:]:
        ctx.build_octree(sc.basecube, sc.depth, npt, dynamic=True)
        ctx.render(v, lights, mask)
    ctx.synchronize()
    rebuild_frame_ms = (time.perf_counter() - t0) * 1e3 / K
    log(f"dynamic tree rebuilt every frame ({npt} points): {rebuild_frame_ms:.2f} ms per frame (rebuild + render)")
    ctx.upload_octree(np.zeros((1, 12), np.int32), dynamic=True)  # back to the static-only scene
    value = float(sum(rays)) / wall / 1e6
    km, kl = prof["kind_ms"], prof["kind_launches"]
    log(f"dynamic tree ({m} points, {dn} nodes): {value:.1f} Mrays/s, {wall * 1e3 / K:.2f} ms/frame")
    return {"value": round(value, 3), "unit": "Mrays/s", "ms_per_step": round(wall * 1e3 / K, 4), "steps": K,
            "dynamic_points": m, "dynamic_nodes": dn, "dual_walk": dual, "rebuild_per_frame": rebuild,
            "frame_with_rebuild_ms": {"points": npt, "ms": round(rebuild_frame_ms, 4),
                                      "what": "qbt_build_octree of the dynamic tree from resident points + the frame"},
            "rays_per_frame": round(float(np.mean(rays))),
            "expansions_per_frame": round(float(np.mean(exps))),
            "kernel_ms_per_frame": {"primary": round(km[KIND_PRIMARY] / K, 4),
                                    "shadow": round(km[KIND_SHADOW] / K, 4),
                                    "camera_and_shadow": round(km[KIND_FRAME] / K, 4)}}


def config5_leg(args, q, ctx, nthreads, mask):
    """The learned shadow surrogate on its own workload: the config-2 scene
    (10 M points, depth 12, one light) at the bench resolution; a seeded camera
    path of c5_frames exact frames timed, then the network trained online on
    16 M samples from 256 exact frames of a separate path and the same frames
    timed with it (surrogate_leg), with the mismatch against the exact frames."""
    sc = q.Scene(5, seed=args.seed)
    t0 = time.time()
    pnt, nrm, col = sc.points(nthreads=nthreads)
    ctx.upload_points(pnt)
    ctx.upload_model(col, nrm)
    ctx.upload_octree(np.zeros((1, 12), np.int32), dynamic=True)
    nnodes = ctx.build_octree(sc.basecube, sc.depth, sc.npoints)
    ctx.synchronize()
    del pnt, nrm, col
    log(f"config 5 leg: config-2 scene ({sc.npoints} points, {nnodes} nodes, {len(sc.lights)} light) in "
        f"{time.time() - t0:.2f}s")
    W, H, K, Wm = args.width, args.height, args.c5_frames, args.warmup
    lights = sc.lights
    pos, ang = sc.poses(K + Wm, seed=args.pose_seed)
    views = [q.view(pos[i], ang[i], W, H, sc.basecube, sc.depth) for i in range(K + Wm)]
    rays, exps, exps_p, leafhits = count_frames(q, ctx, views[Wm:], lights, mask)
    for i in range(Wm):
        ctx.render(views[i], lights, mask)
    ctx.synchronize()
    ctx.profile_begin()
    ctx.synchronize()
    t0 = time.perf_counter()
    for i in range(Wm, K + Wm):
        ctx.render(views[i], lights, mask)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    prof = ctx.profile_end()
    total_rays = float(sum(rays))
    km = prof["kind_ms"]
    log(f"config 5 leg: exact {total_rays / wall / 1e6:.1f} Mrays/s, {wall * 1e3 / K:.3f} ms/frame")
    qpf = (total_rays - K * W * H) / K
    res = surrogate_leg(args, q, ctx, sc, views, Wm, K, mask, wall, total_rays, qpf)
    res["exact"] = {"value": round(total_rays / wall / 1e6, 3), "ms_per_step": round(wall * 1e3 / K, 4),
                    "kernel_ms_per_frame": {"primary": round(km[KIND_PRIMARY] / K, 4),
                                            "shadow": round(km[KIND_SHADOW] / K, 4),
                                            "camera_and_shadow": round(km[KIND_FRAME] / K, 4)}}
    res["workload"] = (f"config 5: the config-2 scene ({sc.npoints} points, depth {sc.depth}, {len(lights)} light), "
                       f"{W}x{H}, {K} frames of a seeded camera path, exact and with the network")
    res["points"], res["nodes"], res["steps"] = sc.npoints, nnodes, K
    return res


BATCH_RES = ((320, 240), (640, 480), (1280, 720), (1920, 1080), (3840, 2160))


def batch_leg(args, q, ctx, nthreads, mask):
    """Batched views: the 90M-point scene with 4 lights, N views of
    resolutions drawn from BATCH_RES at random poses, rendered as one batch
    (one persistent launch per stage, hipGraph-captured and replayed), with
    the tuning the autotuner picks on the first 8 views; every view checked
    bit for bit against its single-view render."""
    sc = q.Scene(4, seed=args.seed)
    lights = sc.lights
    t0 = time.time()
    pnt, nrm, col = sc.points(nthreads=nthreads)
    ctx.upload_points(pnt)
    ctx.upload_model(col, nrm)
    nnodes = ctx.build_octree(sc.basecube, sc.depth, sc.npoints)
    ctx.synchronize()
    del pnt, nrm, col
    log(f"batch leg: config 4 scene ({sc.npoints} points, {nnodes} nodes, {len(lights)} lights) in "
        f"{time.time() - t0:.2f}s")
    n = args.batch_views
    rng = np.random.default_rng(args.pose_seed + 4)
    sizes = [BATCH_RES[i] for i in rng.integers(0, len(BATCH_RES), n)]
    pos, ang = sc.poses(n, seed=args.pose_seed + 4)
    views = [q.view(pos[i], ang[i], w, h, sc.basecube, sc.depth) for i, (w, h) in enumerate(sizes)]
    t0 = time.time()
    best, table = ctx.autotune(views[:8], lights, mask, reps=1)
    tune_s = time.time() - t0
    ctx.set_tuning(graph=1)
    log(f"batch leg: autotuned on 8 views in {tune_s:.1f}s ({len(table)} candidates): block "
        f"{best['block_threads']}, refill {best['refill']}, stack {best['stack']}, reorder {best['reorder']}")
    ctx.render(views, lights, mask)  # capture
    st = ctx.stats()
    rays = st["primary_rays"] + st["shadow_rays"]
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.batch_reps):
        ctx.render(views, lights, mask)
    ctx.synchronize()
    wall = (time.perf_counter() - t0) / args.batch_reps
    # every view against its single-view render, bit for bit (untimed)
    batch_frames = [ctx.read_frame(i) for i in range(n)]
    same = True
    for i, bf in enumerate(batch_frames):
        ctx.render(views[i], lights, mask)
        one = ctx.read_frame(0)
        same = same and all(np.array_equal(bf[k], one[k]) for k in bf)
    check = list(range(n))
    value = rays / wall / 1e6
    log(f"batch leg: {n} views, {sum(w * h for w, h in sizes) / 1e6:.1f} M pixels, {rays / 1e6:.1f} M rays per "
        f"batch in {wall * 1e3:.1f} ms = {value:.1f} Mrays/s; bit-exact vs single views: {same}")
    if not same:
        raise SystemExit("batched frames differ from single-view frames")
    # the same batch with the shadow rays reordered by direction octant (the
    # workload's stated order), whatever the autotuner picked; every view
    # against the frames above (equal to the single-view renders)
    ctx.set_tuning(reorder=1, graph=1)
    ctx.render(views, lights, mask)  # capture
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.batch_reps):
        ctx.render(views, lights, mask)
    ctx.synchronize()
    wall_o = (time.perf_counter() - t0) / args.batch_reps
    same_o = True
    for i, bf in enumerate(batch_frames):
        fo = ctx.read_frame(i)
        same_o = same_o and all(np.array_equal(bf[k], fo[k]) for k in bf)
    del batch_frames
    ctx.set_tuning(reorder=best["reorder"], graph=1)
    log(f"batch leg, shadow rays in direction-octant order: {wall_o * 1e3:.1f} ms = {rays / wall_o / 1e6:.1f} "
        f"Mrays/s; bit-exact vs single views: {same_o}")
    if not same_o:
        raise SystemExit("octant-ordered batch frames differ from single-view frames")
    return {"value": round(value, 3), "unit": "Mrays/s", "ms_per_batch": round(wall * 1e3, 3),
            "reps": args.batch_reps, "views": n, "pixels": int(sum(w * h for w, h in sizes)),
            "rays_per_batch": int(rays), "lights": int(len(lights)), "points": sc.npoints, "nodes": nnodes,
            "resolutions": {f"{w}x{h}": int(sum(1 for s_ in sizes if s_ == (w, h))) for w, h in BATCH_RES},
            "tuning": best, "autotune_candidates": len(table), "autotune_s": round(tune_s, 2),
            "graph": True, "bit_exact_vs_single_view": same, "checked_views": len(check),
            "octant_reordered": {"value": round(rays / wall_o / 1e6, 3), "ms_per_batch": round(wall_o * 1e3, 3),
                                 "bit_exact_vs_single_view": same_o, "checked_views": n},
            "pose_seed": args.pose_seed + 4}


def main():
    args = parse()
    if args.summarize_pmc is not None:
        summarize_pmc(args.summarize_pmc, args.traffic)
        return
    if args.summarize_mfma is not None:
        summarize_mfma(args.summarize_mfma)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on exactly one GPU (asked for --gpus {args.gpus}); "
              "the traversal shards by view or screen tile with no per-frame collective, "
              "but multi-GPU runs are not part of it.", file=sys.stderr)
        sys.exit(2)

    import qubatron_amd as q

    t_all = time.time()
    sc = q.Scene(args.config, seed=args.seed, npoints=args.npoints)
    lights = sc.lights
    nl = len(lights)
    W, H = args.width, args.height
    nthreads = args.cpu_threads or min(16, os.cpu_count() or 1)

    t0 = time.time()
    pnt, nrm, col = sc.points(nthreads=nthreads)
    log(f"generated {sc.npoints} points in {time.time() - t0:.2f}s (config {args.config}, seed {args.seed})")

    ctx = q.Context(0)
    ctx.set_tuning(path=args.path, block_threads=args.block, refill=args.refill, layout=args.layout,
                   stack=args.stack, reorder=args.reorder, graph=args.graph, strip=args.strip, merge=args.merge)
    tuning = ctx.tuning()
    copy_ms, copy_gbps = ctx.copy_bandwidth(COPY_BYTES, 5)
    log(f"copy kernel: {copy_gbps:.0f} GB/s (read+write) on {COPY_BYTES >> 20} MiB")
    t0 = time.time()
    ctx.upload_points(pnt)
    ctx.upload_model(col, nrm)
    ctx.upload_octree(np.zeros((1, 12), np.int32), dynamic=True)
    log(f"uploaded points/colours/normals in {time.time() - t0:.2f}s")
    t0 = time.time()
    nnodes = ctx.build_octree(sc.basecube, sc.depth, sc.npoints)
    ctx.synchronize()
    t_build = time.time() - t0
    log(f"built octree on GPU: {nnodes} nodes ({nnodes * 48 / 1e9:.2f} GB reference layout) in {t_build:.3f}s")
    if args.edit_shots <= 0:
        del pnt

    K, Wm = args.steps, args.warmup
    pos, ang = sc.poses(K + Wm, seed=args.pose_seed)
    if args.config == 1:  # the reference demo camera: cube centre, looking +z (yaw pi), every frame
        pos[:] = (900.0, 900.0, 900.0)
        ang[:] = (np.pi, 0.0)
    views = [q.view(pos[i], ang[i], W, H, sc.basecube, sc.depth) for i in range(K + Wm)]
    mask = q.OUT_DEFAULT

    # untimed counting pass: per-frame rays, expansions and primary leaf hits
    rays, exps, exps_p, leafhits = count_frames(q, ctx, views[Wm:K + Wm], lights, mask)
    tuning["compact_layout_used"] = bool(ctx.stats()["compact"])
    log(f"counting pass: {np.mean(rays) / 1e6:.3f} Mrays/frame, {np.mean(exps) / 1e6:.2f} M expansions/frame "
        f"({np.mean(exps_p) / 1e6:.2f} M primary)")

    for i in range(Wm):  # warmup
        ctx.render(views[i], lights, mask)
    ctx.synchronize()

    # timed region: K frames back to back, synchronised on both sides
    ctx.profile_begin()
    ctx.synchronize()
    t0 = time.perf_counter()
    for i in range(Wm, K + Wm):
        ctx.render(views[i], lights, mask)
    ctx.synchronize()
    t1 = time.perf_counter()
    prof = ctx.profile_end()
    wall = t1 - t0
    total_rays = float(sum(rays))
    value = total_rays / wall / 1e6
    km, kl = prof["kind_ms"], prof["kind_launches"]
    log(f"timed {K} frames: {wall * 1e3:.2f} ms wall; primary {km[0] / max(kl[0], 1):.3f} ms x {kl[0]}, "
        f"shadow {km[1] / max(kl[1], 1):.3f} ms x {kl[1]}, camera and shadow {km[3] / max(kl[3], 1):.3f} ms x {kl[3]}")
    roof = roofline(args, prof, K, W, H, nl, exps, exps_p, leafhits, copy_gbps)
    log(f"roofline: {roof['kernel']} {roof['achieved']} GB/s = {roof['frac']:.4f} of peak; frame "
        f"{roof['frame']['achieved']} GB/s")

    sur = None
    if not args.no_surrogate:
        qpf = (total_rays - K * W * H) / K  # shadow queries per frame
        sur = surrogate_leg(args, q, ctx, sc, views, Wm, K, mask, wall, total_rays, qpf)

    # CPU baseline: the oracle on whole frames of the timed camera path
    cpu = None
    nodes = None
    if not args.no_cpu:
        from oracle import pyoracle as po
        t0 = time.time()
        nodes = ctx.download_octree(nnodes)
        log(f"downloaded node array for the CPU oracle in {time.time() - t0:.2f}s")
        osc = po.OracleScene(nodes, col, nrm, sc.basecube, sc.depth)
        tc, cpu_rays, frames, same = 0.0, 0, 0, True
        for i in range(Wm, K + Wm):
            ctx.render(views[i], lights, mask | q.OUT_FLAGS | q.OUT_ISP)
            g = ctx.read_frame(0)
            ts = time.time()
            o = osc.render(pos[i], ang[i], W, H, lights, nthreads=nthreads)
            tc += time.time() - ts
            cpu_rays += W * H + int(np.count_nonzero(o["flags"] & po.PX_SHADED)) * nl
            frames += 1
            same = same and all(np.array_equal(g[k], o[k]) for k in ("hit", "leaf", "shadow", "flags", "isp",
                                                                    "rgba8"))
            if tc >= args.cpu_seconds:
                break
        log(f"CPU oracle: {frames} frames in {tc:.2f}s on {nthreads} threads, bit-exact vs GPU: {same}")
        if not same:
            raise SystemExit("CPU oracle and GPU frames differ")
        cpu = {"value": round(cpu_rays / tc / 1e6, 4), "unit": "Mrays/s", "cores": nthreads, "kind": "port",
               "sample": f"{frames} whole {W}x{H} frames of the timed camera path ({cpu_rays} rays), "
                         f"C oracle -O2 -ffp-contract=off, pthreads",
               "seconds": round(tc, 2), "bit_exact_vs_gpu": same, "host_cpu": host_cpu()}

    uhd = uhd_leg(args, q, ctx, sc, mask, copy_gbps) if args.uhd_frames > 0 else None
    dyn = dynamic_leg(args, q, ctx, sc, mask, copy_gbps) if args.dyn_frames > 0 else None
    edits = None
    if args.edit_shots > 0:
        if nodes is None:
            nodes = ctx.download_octree(nnodes)
        edits = edit_leg(args, q, ctx, sc, pnt, nrm, nodes, mask)
        del pnt
    batch = batch_leg(args, q, ctx, nthreads, mask) if args.batch_views > 0 else None  # replaces the scene
    c5 = None
    if args.c5_frames > 0 and not args.no_surrogate:
        c5 = config5_leg(args, q, ctx, nthreads, mask)  # replaces the scene

    res = {
        "metric": "primary+shadow Mrays/s @1080p, 90M-pt depth-12 octree; surrogate vs exact",
        "value": round(value, 3), "unit": "Mrays/s", "n_gpus": 1, "steps": K, "warmup": Wm,
        "ms_per_step": round(wall * 1e3 / K, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded scene generator, no assets)",
        "config": {"workload": f"config{args.config} exact primary+shadow traversal, {sc.npoints} points, "
                               f"depth {sc.depth}, {W}x{H}, {nl} lights, "
                               + ("the demo camera (900, 900, 900) looking +z" if args.config == 1
                                  else "seeded camera path"),
                   "points": sc.npoints, "nodes": nnodes, "width": W, "height": H, "lights": int(nl),
                   "poses": K, "seed": args.seed, "pose_seed": args.pose_seed,
                   "rays_per_frame": round(total_rays / K), "expansions_per_frame": round(float(np.mean(exps)))},
        "roofline": roof,
        "cpu_baseline": cpu,
        "surrogate": sur,
        "uhd_3840x2160": uhd,
        "dynamic_tree": dyn,
        "edits": edits,
        "batch_views": batch,
        "surrogate_config5": c5,
        "tuning": tuning,
        "build_s": round(t_build, 3),
        "total_s": round(time.time() - t_all, 1),
    }
    print(json.dumps(res), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
