"""Benchmark: primary+shadow Mrays/s at 1920x1080 on the 90M-point depth-12
octree (config 3: planes 60%, 200 boxes, 50 spheres, a 10M-point cluster of
15 ellipsoids; 2 point lights), exact traversal on one MI355X.

A step is one frame: one camera pose of a seeded camera path rendered by the
hot path (primary trace, one shadow trace per light for every shaded pixel,
shading), inputs resident in HBM. value = (primary + shadow rays) / wall time
of the K timed frames. The traversal kernel's average launch duration is
measured with HIP events on the library's stream (qbt_profile_*), and its
algorithmic bytes (32 B per node expansion + 28 B per primary leaf hit + the
per-pixel outputs) give the roofline fraction against 8 TB/s. The CPU oracle
(the C restatement of the reference traversal) is timed beside it on whole
frames of the same camera path, using the host's cores, and each sampled
frame is checked bit-exactly against the GPU frame.

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
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_k_render_config3_1080p.json")
RENDER_KERNEL = "k_render<false, true, false>"


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
    ap.add_argument("--traffic", default=TRAFFIC_FILE)
    ap.add_argument("--summarize-pmc", nargs="*", default=None, metavar="DIR")
    ap.add_argument("--path", type=int, default=None, help="render path (0 fused, 1 wavefront); default: library's")
    ap.add_argument("--block", type=int, default=None, help="threads per block (64/128/256)")
    ap.add_argument("--refill", type=int, default=None, help="idle lanes that trigger a wave refill (1..64)")
    ap.add_argument("--layout", type=int, default=None, help="0 compact node layout, 1 direct")
    return ap.parse_args()


TRAVERSAL_KERNELS = ("k_render", "k_primary", "k_shadow")


def summarize_pmc(dirs, out_path):
    """Average per-dispatch counters of each traversal kernel (non-counting
    variants) and of the copy calibration kernel; byte counters are scaled by
    the copy kernel's known traffic. The traversal kernels of one frame are
    summed into a per-frame figure."""
    vals = {}
    for d in dirs:
        for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            with open(f) as fh:
                for row in csv.DictReader(fh):
                    name = row.get("Kernel_Name", "")
                    kern = None
                    if "k_copy16" in name:
                        kern = "copy"
                    m = re.search(r"\b(k_render|k_primary|k_shadow)<([^>]*)>", name)
                    if m and m.group(2).split(",")[-1].strip() == "false":  # not the counting variant
                        kern = m.group(1)
                    if kern is None:
                        continue
                    key = (kern, row["Counter_Name"])
                    vals.setdefault(key, {})[(f, row.get("Dispatch_Id"))] = float(row["Counter_Value"])
    avg = {k: float(np.mean(list(v.values()))) for k, v in vals.items()}
    tot = {k: float(np.sum(list(v.values()))) for k, v in vals.items()}
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
            e["fetch_bytes_total"] = tot[(k, "FETCH_SIZE")] * 1024.0 * ff
        if (k, "WRITE_SIZE") in avg and wf:
            e["write_bytes_per_launch"] = avg[(k, "WRITE_SIZE")] * 1024.0 * wf
            e["write_bytes_total"] = tot[(k, "WRITE_SIZE")] * 1024.0 * wf
        h, m = avg.get((k, "TCC_HIT_sum")), avg.get((k, "TCC_MISS_sum"))
        if h is not None and m is not None and h + m > 0:
            e["l2_hit_rate"] = h / (h + m)
        if e:
            per_kernel[k] = e
    res["kernels"] = per_kernel
    # per frame: all traversal launches of the timed frames
    frames = None
    for k, e in per_kernel.items():
        if k in ("k_primary", "k_render") and "launches" in e:
            frames = e["launches"]
    if frames:
        fb = sum(e.get("fetch_bytes_total", 0.0) for e in per_kernel.values()) / frames
        wb = sum(e.get("write_bytes_total", 0.0) for e in per_kernel.values()) / frames
        res["hbm_bytes_per_launch"] = round(fb + wb)
        res["hbm_bytes_per_launch_note"] = "traversal kernels of one frame (one launch of the fused kernel, or " \
                                           "the primary + per-light shadow launches of the wavefront path)"
        hits = [e["l2_hit_rate"] for e in per_kernel.values() if "l2_hit_rate" in e]
        if hits:
            res["l2_hit_rate"] = float(np.mean(hits))
    res["counters_kib_or_count"] = {"%s:%s" % k: v for k, v in avg.items()}
    res["dispatches"] = {"%s:%s" % k: v for k, v in n.items()}
    os.makedirs(os.path.dirname(out_path), exist_ok=True)
    with open(out_path, "w") as fh:
        json.dump(res, fh, indent=1, sort_keys=True)
    print(json.dumps(res, indent=1, sort_keys=True))


def main():
    args = parse()
    if args.summarize_pmc is not None:
        summarize_pmc(args.summarize_pmc, args.traffic)
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
    W, H = args.width, args.height
    nthreads = args.cpu_threads or min(16, os.cpu_count() or 1)

    t0 = time.time()
    pnt, nrm, col = sc.points(nthreads=nthreads)
    log(f"generated {sc.npoints} points in {time.time() - t0:.2f}s (config {args.config}, seed {args.seed})")

    ctx = q.Context(0)
    ctx.set_tuning(path=args.path, block_threads=args.block, refill=args.refill, layout=args.layout)
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
    del pnt

    K, Wm = args.steps, args.warmup
    pos, ang = sc.poses(K + Wm, seed=args.pose_seed)
    views = [q.view(pos[i], ang[i], W, H, sc.basecube, sc.depth) for i in range(K + Wm)]
    mask = q.OUT_DEFAULT

    # untimed counting pass: per-frame rays, expansions and primary leaf hits
    rays, exps, leafhits = [], [], []
    for i in range(Wm, K + Wm):
        ctx.render(views[i], lights, mask | q.OUT_FLAGS | q.RENDER_COUNT)
        st = ctx.stats()
        fl = ctx.read_frame(0, fields=("flags",))["flags"]
        rays.append(st["primary_rays"] + st["shadow_rays"])
        exps.append(st["expansions"])
        tuning["compact_layout_used"] = bool(st["compact"])
        leafhits.append(int(np.count_nonzero(fl & q.PX_LEAF)))
    log(f"counting pass: {np.mean(rays) / 1e6:.3f} Mrays/frame, {np.mean(exps) / 1e6:.2f} M expansions/frame")

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
    log(f"timed {K} frames: {wall * 1e3:.2f} ms wall, kernel avg {prof['avg_ms']:.3f} ms "
        f"(min {prof['min_ms']:.3f}, max {prof['max_ms']:.3f}) over {prof['launches']} launches")

    # roofline of the traversal kernel (one launch = one frame)
    out_bytes = W * H * (4 + 4 + len(lights) + 4)
    alg_bytes = [32.0 * e + 28.0 * h + out_bytes for e, h in zip(exps, leafhits)]
    achieved = (sum(alg_bytes) / len(alg_bytes)) / (prof["avg_ms"] * 1e-3) / 1e9
    traffic, l2 = None, None
    if os.path.exists(args.traffic):
        try:
            with open(args.traffic) as f:
                tj = json.load(f)
            traffic, l2 = tj.get("hbm_bytes_per_launch"), tj.get("l2_hit_rate")
        except Exception:
            traffic = None
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "alg_bytes_per_launch": round(float(np.mean(alg_bytes))), "kernel": "k_render",
            "kernel_avg_ms": round(prof["avg_ms"], 4), "l2_hit_rate": l2,
            "copy_kernel_gbps": round(copy_gbps, 1)}

    # CPU baseline: the oracle on whole frames of the timed camera path
    cpu = None
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
            cpu_rays += W * H + int(np.count_nonzero(o["flags"] & po.PX_SHADED)) * len(lights)
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
               "seconds": round(tc, 2), "bit_exact_vs_gpu": same}

    res = {
        "metric": "primary+shadow Mrays/s @1080p, 90M-pt depth-12 octree; surrogate vs exact",
        "value": round(value, 3), "unit": "Mrays/s", "n_gpus": 1, "steps": K, "warmup": Wm,
        "ms_per_step": round(wall * 1e3 / K, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded scene generator, no assets)",
        "config": {"workload": f"config{args.config} exact primary+shadow traversal, {sc.npoints} points, "
                               f"depth {sc.depth}, {W}x{H}, {len(lights)} lights, seeded camera path",
                   "points": sc.npoints, "nodes": nnodes, "width": W, "height": H, "lights": int(len(lights)),
                   "poses": K, "seed": args.seed, "pose_seed": args.pose_seed,
                   "rays_per_frame": round(total_rays / K), "expansions_per_frame": round(float(np.mean(exps)))},
        "roofline": roof,
        "cpu_baseline": cpu,
        "tuning": tuning,
        "build_s": round(t_build, 3),
        "total_s": round(time.time() - t_all, 1),
    }
    print(json.dumps(res), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
