"""Benchmark: batched FP32 tile-pair correlations/s on one MI355X.

Workloads (--workload):
  strips  (default) 256 pairs of 2048 x 256 uint8 strips (band-limited noise,
          per-pair seeds, fractional shifts by bilinear resampling, N(0,10)
          noise), CorrPatches semantics (normalised cross-correlation, FP32
          FFTs at 4096 x 512, exact integer summed-area tables -- the FP64
          tables' sums without their rounding -- and variance terms exact
          until their final FP64 scaling, legal overlap olw >= W/2,
          olh >= H/2).  The same batch with FP64 (S, S^2) table entries and
          with FP32 ones is timed beside it and reported as variants.  One step = one pass of the
          whole path over the batch, starting from the raw uint8 strips
          resident in HBM: statistics, summed-area tables, row r2c, fused
          column pass, row c2r + NCC + peak search, subpixel fit.
  sweep   CThmScan rotation sweep: 64 pairs of 2048 x 2048 uint16 fields
          (mean 30000, sd 6000), B rotated by U[-2.5, 2.5] deg and shifted by
          U[-60, 60] px, both decimated 8x to 256 x 256 and normalised;
          21 angles -5..5 deg each, CorrImagesS via RFromAngle: 1344
          correlations in one plan.
  uneven  4096 pairs, w, h ~ U{128..1024}, uint8 with 0-3 elliptical holes
          per mask, shifts U{-50..50}, N(0,12); CorrPatches with an
          EnoughPoints count rule (c1, c2 > w h / 4); pairs bucketed by
          padded size.
  serve   Serving loop: Poisson arrivals at 2000, 8000 and 20000 req/s and
          a burst; 70 % uneven-size and 30 % strip requests over tiles on
          4 x 4 tori (each tile referenced by four requests); device budget
          200 GB (capped by free memory); per-request p50/p99 latency and
          sustained pairs/s.

Printed: one JSON line per run with the throughput, accuracy against the
FP64 oracle and the plain FP64 GPU path, the roofline of the dominant kernel
(HIP events on the library stream) and a CPU baseline (the oracle on the host
cores, FFTW3 API where available, on a bounded sample: one thread, and one
pair per thread on all cores).  The default strips run also carries the
accuracy of the sweep (all 1344 correlations) and of 1024 uneven pairs
against the plain FP64 GPU path, and a short serving leg (p50 / p99 per
request at 2000, 8000 and 20000 requests/s).
"""
import argparse
import json
import os
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

# the CPU baseline's FFT library runs single-threaded per call, as the reference's install notes prescribe
os.environ.setdefault("MKL_NUM_THREADS", "1")

HBM_PEAK_GBS = 8000.0
STAGES = ["prep", "row_r2c", "col_fused", "row_c2r_eval", "finalize"]
KERNELS = {"prep": "k_tp_*", "row_r2c": "k_row_r2c", "col_fused": "k_col_fused",
           "row_c2r_eval": "k_row_c2r_eval", "finalize": "k_finalize"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _device_sync():
    """torch.cuda.synchronize if torch sees the device (initialised before the library), else a no-op:
    every plan run already waits on the library's stream."""
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.init()
            return torch.cuda.synchronize
    except Exception as e:  # pragma: no cover - depends on the box
        log(f"torch device sync unavailable: {e}")
    return lambda: None


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=["strips", "sweep", "uneven", "serve"], default="strips")
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--pairs", type=int, default=0, help="pairs (default: the workload's own count)")
    ap.add_argument("--width", type=int, default=2048)
    ap.add_argument("--height", type=int, default=256)
    ap.add_argument("--oracle-sample", type=int, default=16, help="pairs checked against the FP64 oracle")
    ap.add_argument("--cpu-sample", type=int, default=192, help="pairs timed on the CPU baseline")
    ap.add_argument("--plain-sample", type=int, default=256, help="pairs run through the plain FP64 GPU path")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-variant", action="store_true", help="skip the FP64- and FP32-table variants of the strips")
    ap.add_argument("--no-accuracy", action="store_true",
                    help="skip the accuracy legs of the strips run (sweep and uneven against the plain FP64 path)")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg of the strips run")
    ap.add_argument("--cpu-threads", type=int, default=0, help="threads of the all-core CPU baseline (0: the host's)")
    ap.add_argument("--stage-iters", type=int, default=5)
    ap.add_argument("--serve-seconds", type=float, default=2.0, help="duration of each offered load")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="run one device-to-device copy of MIB MiB first (a known-size kernel for counter scaling)")
    return ap.parse_args()


# ---------------------------------------------------------------------------
# shared measurement pieces
# ---------------------------------------------------------------------------

def time_plan(plan, steps, warmup, sync):
    for _ in range(warmup):
        plan.run(want=False)
    lat = []
    sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        t1 = time.perf_counter()
        plan.run(want=False)
        lat.append(time.perf_counter() - t1)
    sync()
    return time.perf_counter() - t0, lat


def stage_roofline(plan, iters, ms_step):
    """Per-stage times (HIP events on the library stream) and the dominant stage's roofline."""
    ms, by = plan.stage_times(iters)
    model = plan.model_bytes
    dom = int(np.argmax(ms))
    ach = by[dom] / (ms[dom] * 1e-3) / 1e9
    stages = {STAGES[i]: {"ms": round(ms[i], 4), "alg_bytes": by[i],
                          "GBs": round(by[i] / (ms[i] * 1e-3) / 1e9, 1) if ms[i] > 0 else None}
              for i in range(5)}
    kern = KERNELS[STAGES[dom]]
    if STAGES[dom] == "row_r2c" and all(b["r2c_np"] > 0 for b in plan.buckets()):
        kern = "k_row_r2c_u8"  # every bucket transforms its rows with the 8-bit row r2c
    roof = {"bound": "hbm", "kernel": kern, "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
            "alg_bytes_per_launch": by[dom],
            "pipeline": {"alg_bytes_per_step": model,
                         "achieved_GBs": round(model / (ms_step * 1e-3) / 1e9, 1),
                         "frac": round(model / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}}
    return stages, roof


def measured_traffic(name, kernel):
    """HBM bytes per launch of `kernel` measured with rocprofv3 counters on this workload (the summary
    committed under profiles/, which gives the command and the scaling), or None."""
    try:
        with open(os.path.join(ROOT, "profiles", name)) as f:
            for ln in f:
                p = ln.split()
                if len(p) == 3 and p[0] == "traffic_bytes_per_launch" and p[1] == kernel:
                    return float(p[2])
    except OSError:
        pass
    return None


def accuracy(res, ref):
    """res: XcResult list; ref: list of (status, r, dx, dy, XcResult) for the same correlations."""
    agree, sub, rerr = 0, [], []
    for r, o in zip(res, ref):
        if (r.lag_x, r.lag_y) == (o[4].lag_x, o[4].lag_y):
            agree += 1
            sub.append(max(abs(r.dx - o[2]), abs(r.dy - o[3])))
        rerr.append(abs(r.r - o[1]))
    sub = np.array(sub) if sub else np.array([np.nan])
    rerr = np.array(rerr)
    return {"n": len(ref), "peak_agree": agree / max(1, len(ref)), "subpx_max": float(sub.max()),
            "subpx_p99": float(np.percentile(sub, 99)), "subpx_mean": float(sub.mean()),
            "r_err_max": float(rerr.max()), "r_err_p99": float(np.percentile(rerr, 99))}


def accuracy_recs(res, recs):
    """Batched sweep results against the plain FP64 path's RFromAngle records (A, R, X, Y): the integer peak
    agrees when both displacements are within half a pixel; subpixel error is the larger displacement
    difference; the score error is relative."""
    agree, sub, rerr = 0, [], []
    for r, o in zip(res, recs):
        dx, dy = abs(r.dx - o.X), abs(r.dy - o.Y)
        if r.status == 0 and dx < 0.5 and dy < 0.5:
            agree += 1
            sub.append(max(dx, dy))
        rerr.append(abs(r.r - o.R) / max(1.0, abs(o.R)))
    sub = np.array(sub) if sub else np.array([np.nan])
    rerr = np.array(rerr)
    return {"n": len(recs), "peak_agree": agree / max(1, len(recs)), "subpx_max": float(sub.max()),
            "subpx_p99": float(np.percentile(sub, 99)), "subpx_mean": float(sub.mean()),
            "r_rel_err_max": float(rerr.max()), "r_rel_err_p99": float(np.percentile(rerr, 99))}


def merge_accuracy(parts):
    """Pooled agreement over several accuracy records (counts add; worst cases are the maxima)."""
    n = sum(p["n"] for p in parts)
    agree = sum(p["peak_agree"] * p["n"] for p in parts)
    return {"n": n, "peak_agree": agree / max(1, n), "peaks_differing": int(round(n - agree)),
            "subpx_max": max(p["subpx_max"] for p in parts)}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_threads(args):
    """Threads of the all-core baseline: the host's cores, at most 16 (one GPU's share of a shared box)."""
    n = args.cpu_threads or min(16, os.cpu_count() or 1)
    try:
        n = min(n, len(os.sched_getaffinity(0)))
    except AttributeError:
        pass
    return max(1, n)


def fft_provider(oracle):
    try:
        return oracle.set_fft_provider(1)
    except RuntimeError:
        return oracle.set_fft_provider(0)


def line_base(metric, value, unit, args, ms_step, config, dtype="f32"):
    return {"metric": metric, "value": round(value, 1), "unit": unit, "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(ms_step, 3), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": dtype, "data": "synthetic", "config": config}


# ---------------------------------------------------------------------------
# accuracy legs: the other workloads' batched results against the plain FP64 GPU path
# ---------------------------------------------------------------------------

def accuracy_legs(ctx, sync):
    """All 64 x 21 correlations of the thumbnail sweep and the first 1024 pairs of the uneven batch, batched
    (FP32) against the plain FP64 GPU path (itself bit-exact in its integer peak against the oracle)."""
    from alignment_projects_amd import abi, synth, workloads, xcorr
    out = {}
    t0 = time.time()
    A, B, _ = workloads.thumb_pairs(ctx, 64)
    angles = synth.sweep_angles()
    h, w = A[0].shape
    req, olap = int(0.25 * w * h), 64
    rgn, cnt = abi.rgn_big_enough(olap, req), abi.cnt_enough_points(req)
    A32 = [a.astype(np.float32) for a in A]
    B32 = [b.astype(np.float32) for b in B]
    tiles = [(ctx.tile_put(A32[i], normalize=False), ctx.tile_put(B32[i], normalize=False)) for i in range(len(A))]
    pairs = [xcorr.make_pair(ta, tb, mode=abi.XC_MODE_S, rgn=rgn, cnt=cnt, angle=ang) for ta, tb in tiles
             for ang in angles]
    plan = ctx.plan(pairs)
    res = plan.run(want=True)
    plan.destroy()
    for ta, tb in tiles:
        ta.release()
        tb.release()
    recs = []
    for i in range(len(A)):
        pa, va = synth.raster_points(A32[i].astype(np.float64))
        pb, vb = synth.raster_points(B32[i].astype(np.float64))
        recs += [ctx.rfromangle(ang, pa, va, pb, vb, olap, req) for ang in angles]
    out["sweep_vs_plain_fp64_gpu"] = accuracy_recs(res, recs)
    log(f"accuracy, sweep: {out['sweep_vs_plain_fp64_gpu']} ({time.time() - t0:.1f}s)")
    t0 = time.time()
    data = workloads.uneven_pairs(1024)
    sync()
    tiles, pairs = [], []
    for a, ma, b, mb, _ in data:
        ta, tb = _adopt(ctx, a, ma), _adopt(ctx, b, mb)
        tiles.append((ta, tb))
        hh, ww = a.shape
        pairs.append(xcorr.make_pair(ta, tb, cnt=abi.cnt_enough_points(ww * hh // 4)))
    plan = ctx.plan(pairs, refresh_tiles=True)
    res = plan.run(want=True)
    plan.destroy()
    for ta, tb in tiles:
        ta.release()
        tb.release()
    ref = []
    for a, ma, b, mb, _ in data:
        hh, ww = a.shape
        pa, va = synth.raster_points(a.cpu().numpy(), ma.cpu().numpy())
        pb, vb = synth.raster_points(b.cpu().numpy(), mb.cpu().numpy())
        ref.append(ctx.corr_patches(pa, synth.normalize(va), pb, synth.normalize(vb), 0, 0, 4096,
                                    cnt=abi.cnt_enough_points(ww * hh // 4)))
    out["uneven_vs_plain_fp64_gpu"] = accuracy(res, ref)
    log(f"accuracy, uneven: {out['uneven_vs_plain_fp64_gpu']} ({time.time() - t0:.1f}s)")
    del data
    return out


# ---------------------------------------------------------------------------
# strips: batched same-layer overlaps (the headline line)
# ---------------------------------------------------------------------------

def run_strips(args, sync):
    from alignment_projects_amd import abi, synth, xcorr

    W, H, NP = args.width, args.height, args.pairs or 256
    t0 = time.time()
    A, B, S = synth.strip_batch(NP, seed0=1000, w=W, h=H)
    log(f"generated {NP} strip pairs {W}x{H} in {time.time() - t0:.1f}s")
    # BigEnough-style legal overlap for same-layer strips: olw >= W/2, olh >= H/2
    rgn = abi.rgn_min(W // 2, H // 2)
    if args.calibrate_copy:
        import torch
        src = torch.ones(args.calibrate_copy * 2 ** 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy: {args.calibrate_copy} MiB read + {args.calibrate_copy} MiB written")
        del src, dst
    ctx = xcorr.Context(0)
    variants = {}
    for key, form, desc in (("variant_f64_tables", "f64", "f64 (S, S^2) pairs"),
                            ("variant_f32_tables", "f32", "f32 (S, S^2) pairs")):
        if args.no_variant:
            break
        # the same batch with FP64 (the reference's double tables) or FP32 table entries; reported beside the
        # headline, never as it
        vt = [(ctx.tile_put(A[i], tables=form), ctx.tile_put(B[i], tables=form)) for i in range(NP)]
        vp = ctx.plan([xcorr.make_pair(ta, tb, rgn=rgn) for ta, tb in vt], refresh_tiles=True, autotune=True)
        ve, _ = time_plan(vp, args.steps, args.warmup, sync)
        vres = vp.run(want=True)
        variants[key] = ({"tables": desc, "value": round(NP * args.steps / ve, 1), "unit": "pairs/s",
                          "ms_per_step": round(1e3 * ve / args.steps, 3)}, vres)
        vp.destroy()
        for ta, tb in vt:
            ta.release()
            tb.release()
        log(f"variant with {desc} tables: {variants[key][0]['value']} pairs/s")
    tiles = [(ctx.tile_put(A[i]), ctx.tile_put(B[i])) for i in range(NP)]
    pairs = [xcorr.make_pair(ta, tb, rgn=rgn) for ta, tb in tiles]
    plan = ctx.plan(pairs, refresh_tiles=True, autotune=True)
    log("plan ready")
    elapsed, lat = time_plan(plan, args.steps, args.warmup, sync)
    res = plan.run(want=True)
    value = NP * args.steps / elapsed
    ms_step = 1e3 * elapsed / args.steps
    log(f"batched: {value:.1f} pairs/s, {ms_step:.3f} ms/step")
    pcie = None
    try:
        # the batch's pixels crossing PCIe from pinned host memory (not part of the value)
        import torch
        ha, hb = torch.from_numpy(A).pin_memory(), torch.from_numpy(B).pin_memory()
        da, db = torch.empty_like(ha, device="cuda"), torch.empty_like(hb, device="cuda")
        da.copy_(ha, non_blocking=True)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        da.copy_(ha, non_blocking=True)
        db.copy_(hb, non_blocking=True)
        torch.cuda.synchronize()
        h2d = time.perf_counter() - t1
        pcie = {"h2d_ms": round(1e3 * h2d, 3), "bytes": int(A.nbytes + B.nbytes),
                "pairs_s_with_upload": round(NP / (h2d + ms_step * 1e-3), 1)}
        del ha, hb, da, db
    except Exception as e:  # pragma: no cover - depends on the box
        log(f"PCIe measurement skipped: {e}")
    stages, roofline = stage_roofline(plan, args.stage_iters, ms_step)
    log("stages:", json.dumps(stages))
    roofline["traffic"] = measured_traffic("traffic_strips_2048x256_256pairs.md", roofline["kernel"]) \
        if (W, H, NP) == (2048, 256, 256) else None

    # ---- plain FP64 GPU path on the same pairs (one pair per call) ----
    nplain = min(args.plain_sample, NP)
    if not args.no_cpu:
        nplain = NP  # the CPU baselines take their samples from these point lists
    pts = [(synth.raster_points(A[i]), synth.raster_points(B[i])) for i in range(nplain)]
    norm = [((pa, synth.normalize(va)), (pb, synth.normalize(vb))) for (pa, va), (pb, vb) in pts]
    ctx.corr_patches(norm[0][0][0], norm[0][0][1], norm[0][1][0], norm[0][1][1], 0, 0, 4096, rgn=rgn)
    t1 = time.perf_counter()
    plain = []
    for (pa, va), (pb, vb) in norm:
        plain.append(ctx.corr_patches(pa, va, pb, vb, 0, 0, 4096, rgn=rgn))
    plain_rate = nplain / (time.perf_counter() - t1)
    log(f"plain FP64 GPU: {plain_rate:.1f} pairs/s")
    acc = {"vs_plain_fp64_gpu": accuracy(res[:nplain], plain)}

    # ---- FP64 oracle on a sample (checker) and CPU baseline ----
    cpu = None
    if not args.no_cpu:
        import oracle
        nor = min(args.oracle_sample, NP)
        ref = [oracle.corr_patches(pa, va, pb, vb, 0, 0, 4096, rgn=rgn) for (pa, va), (pb, vb) in norm[:nor]]
        acc["vs_fp64_oracle"] = accuracy(res[:nor], ref)
        truth = [float(np.hypot(res[i].dx - S[i][0], res[i].dy - S[i][1])) for i in range(NP)]
        acc["vs_true_shift_px"] = {"max": max(truth), "mean": float(np.mean(truth))}
        prov = fft_provider(oracle)
        ncpu = min(args.cpu_sample, nplain)
        oracle.corr_patches(norm[0][0][0], norm[0][0][1], norm[0][1][0], norm[0][1][1], 0, 0, 4096, rgn=rgn)
        per = []
        t1 = time.perf_counter()
        for (pa, va), (pb, vb) in norm[:ncpu]:
            t2 = time.perf_counter()
            oracle.corr_patches(pa, va, pb, vb, 0, 0, 4096, rgn=rgn)
            per.append(time.perf_counter() - t2)
        cs = time.perf_counter() - t1
        single = {"value": round(ncpu / cs, 3), "unit": "pairs/s", "cores": 1, "seconds": round(cs, 2),
                  "latency_ms": {"p50": round(1e3 * float(np.percentile(per, 50)), 2),
                                 "p99": round(1e3 * float(np.percentile(per, 99)), 2)}}
        log(f"cpu baseline, 1 thread: {single['value']} pairs/s")
        # all cores: one pair per thread (the EZThreads / TCDGet pattern), FFT calls serialised by one mutex as
        # in the reference's FFTW build; the oracle's calls release the interpreter lock
        from concurrent.futures import ThreadPoolExecutor
        nth = cpu_threads(args)
        nall = min(NP, len(norm))
        t1 = time.perf_counter()
        with ThreadPoolExecutor(max_workers=nth) as ex:
            list(ex.map(lambda q: oracle.corr_patches(q[0][0], q[0][1], q[1][0], q[1][1], 0, 0, 4096, rgn=rgn),
                        norm[:nall]))
        ca = time.perf_counter() - t1
        oracle.set_fft_provider(0)
        cpu = {"value": round(nall / ca, 3), "unit": "pairs/s", "cores": nth, "kind": "port",
               "sample": f"{nall} of the {NP} strip pairs, one pair per thread on {nth} threads, CorrPatches "
                         f"restatement, FFT: {prov} under one mutex, MKL_NUM_THREADS=1",
               "seconds": round(ca, 2), "cpu_model": cpu_model(), "nproc": os.cpu_count(),
               "single_thread": single}
        log(f"cpu baseline, {nth} threads: {cpu['value']} pairs/s")

    line = line_base("tile-pair correlations/sec (batched FP32, 2048x256 strip pairs)", value, "pairs/s", args,
                     ms_step, {"workload": "batched same-layer strip overlaps", "pairs": NP, "tile": f"{W}x{H} u8",
                               "fft": f"{2 * W}x{2 * H}",
                               "tables": "exact integer (S, S^2) sums (u64: raw sum << 36 + sum of squares)",
                               "rule": f"olw >= {W // 2}, olh >= {H // 2}"})
    line["latency_ms"] = {"p50": round(1e3 * float(np.percentile(lat, 50)), 3),
                          "p99": round(1e3 * float(np.percentile(lat, 99)), 3), "unit": "per batch"}
    line["plain_fp64_gpu"] = {"value": round(plain_rate, 1), "unit": "pairs/s", "pairs": nplain,
                              "batched_over_plain": round(value / plain_rate, 1)}
    line["stages"] = stages
    line["autotune"] = plan.buckets()
    sw, sw_us = plan.sweep_waves()
    line["sweep_waves"] = {"waves_per_cu": sw, "tuned_us": round(sw_us, 1)}
    line["roofline"] = roofline
    line["cpu_baseline"] = cpu
    line["accuracy"] = acc
    line["pcie_inclusive"] = pcie
    for key, (variant, vres) in variants.items():
        variant["accuracy_vs_plain_fp64_gpu"] = accuracy(vres[:nplain], plain)
        line[key] = variant
    if not args.no_accuracy:
        acc.update(accuracy_legs(ctx, sync))
        acc["all_vs_plain_fp64_gpu"] = merge_accuracy([acc["vs_plain_fp64_gpu"], acc["sweep_vs_plain_fp64_gpu"],
                                                       acc["uneven_vs_plain_fp64_gpu"]])
    if not args.no_serve:
        runs, st, cfg = serve_leg(ctx, sync, [2000, 8000, 20000], 0.5, n_uneven=16, n_strips=6, burst=False)
        line["serving"] = {"config": cfg, "runs": runs, "server_stats": st,
                           "latency_ms_at_20000_req_s": {"p50": runs[2]["p50_ms"], "p99": runs[2]["p99_ms"]}}
    if cpu:
        line["batched_over_cpu"] = round(value / cpu["value"], 1)
    plan.destroy()
    for ta, tb in tiles:
        ta.release()
        tb.release()
    ctx.close()
    return line


# ---------------------------------------------------------------------------
# sweep: CThmScan rotation sweep over thumbnails
# ---------------------------------------------------------------------------

def run_sweep(args, sync):
    from alignment_projects_amd import abi, synth, workloads, xcorr

    NP = args.pairs or 64
    t0 = time.time()
    ctx = xcorr.Context(0)
    A, B, truth = workloads.thumb_pairs(ctx, NP)
    log(f"generated {NP} thumbnail pairs in {time.time() - t0:.1f}s")
    angles = synth.sweep_angles()
    h, w = A[0].shape
    req, olap = int(0.25 * w * h), 64
    rgn, cnt = abi.rgn_big_enough(olap, req), abi.cnt_enough_points(req)
    tiles = [(ctx.tile_put(A[i].astype(np.float32), normalize=False),
              ctx.tile_put(B[i].astype(np.float32), normalize=False)) for i in range(NP)]
    pairs = [xcorr.make_pair(ta, tb, mode=abi.XC_MODE_S, rgn=rgn, cnt=cnt, angle=ang)
             for ta, tb in tiles for ang in angles]
    plan = ctx.plan(pairs, autotune=True)
    log(f"plan ready: {len(pairs)} correlations")
    elapsed, lat = time_plan(plan, args.steps, args.warmup, sync)
    res = plan.run(want=True)
    value = len(pairs) * args.steps / elapsed
    ms_step = 1e3 * elapsed / args.steps
    log(f"sweep: {value:.1f} correlations/s, {ms_step:.3f} ms/step")
    stages, roofline = stage_roofline(plan, args.stage_iters, ms_step)
    na = len(angles)
    best = [angles[int(np.argmax([res[i * na + k].r if res[i * na + k].status == abi.XC_OK else -1e300
                                  for k in range(na)]))] for i in range(NP)]
    ang_err = np.array([abs(best[i] - truth[i][0]) for i in range(NP)])
    acc = {"best_angle_within_0.5deg_of_truth": float(np.mean(ang_err <= 0.5 + 1e-9)),
           "best_angle_err_max_deg": float(ang_err.max())}
    cpu = None
    if not args.no_cpu:
        import oracle
        prov = fft_provider(oracle)
        ns = min(NP, 8)
        ref, agree_best = [], 0
        t1 = time.perf_counter()
        for i in range(ns):
            pa, va = synth.raster_points(A[i])
            pb, vb = synth.raster_points(B[i])
            ri = [oracle.rfromangle(ang, pa, va, pb, vb, olap, req) for ang in angles]
            ob = angles[int(np.argmax([o[1] if o[0] == abi.XC_OK else -1e300 for o in ri]))]
            agree_best += ob == best[i]
            ref += ri
        cs = time.perf_counter() - t1
        oracle.set_fft_provider(0)
        acc["vs_fp64_oracle"] = accuracy(res[:ns * na], ref)
        acc["best_angle_agrees_with_oracle"] = agree_best / ns
        cpu = {"value": round(ns * na / cs, 3), "unit": "correlations/s", "cores": 1, "kind": "port",
               "sample": f"pairs 0..{ns - 1} at all {na} angles, RFromAngle restatement, FFT: {prov}, 1 thread",
               "seconds": round(cs, 2)}
    line = line_base("thumbnail sweep correlations/sec (batched FP32, 64 pairs x 21 angles)", value,
                     "correlations/s", args, ms_step,
                     {"workload": "CThmScan rotation sweep", "pairs": NP, "angles": na, "tile": f"{w}x{h} f32",
                      "source": "2048x2048 u16 decimated 8x", "rule": f"BigEnough({olap}, {req}), EnoughPoints({req})"})
    line["latency_ms"] = {"p50": round(1e3 * float(np.percentile(lat, 50)), 3),
                          "p99": round(1e3 * float(np.percentile(lat, 99)), 3), "unit": "per batch"}
    line["stages"] = stages
    line["autotune"] = plan.buckets()
    sw, sw_us = plan.sweep_waves()
    line["sweep_waves"] = {"waves_per_cu": sw, "tuned_us": round(sw_us, 1)}
    line["roofline"] = roofline
    line["cpu_baseline"] = cpu
    line["accuracy"] = acc
    plan.destroy()
    for ta, tb in tiles:
        ta.release()
        tb.release()
    ctx.close()
    return line


# ---------------------------------------------------------------------------
# uneven: thousands of masked pairs of random size
# ---------------------------------------------------------------------------

def _adopt(ctx, t, m):
    import torch  # noqa: F401  (tensors own the memory)
    h, w = t.shape
    assert m.stride(0) == t.stride(0)  # a device mask has the picture's row pitch
    return ctx.tile_put((h, w), mask=m.data_ptr(), device_ptr=t.data_ptr(), dtype=0, stride=t.stride(0))


def run_uneven(args, sync):
    from alignment_projects_amd import abi, synth, workloads, xcorr

    NP = args.pairs or 4096
    t0 = time.time()
    data = workloads.uneven_pairs(NP)
    sync()
    log(f"generated {NP} uneven pairs in {time.time() - t0:.1f}s")
    ctx = xcorr.Context(0)
    tiles, pairs = [], []
    for a, ma, b, mb, _ in data:
        ta, tb = _adopt(ctx, a, ma), _adopt(ctx, b, mb)
        tiles.append((ta, tb))
        h, w = a.shape
        pairs.append(xcorr.make_pair(ta, tb, cnt=abi.cnt_enough_points(w * h // 4)))
    plan = ctx.plan(pairs, refresh_tiles=True, autotune=True)
    log("plan ready")
    elapsed, lat = time_plan(plan, args.steps, args.warmup, sync)
    res = plan.run(want=True)
    value = NP * args.steps / elapsed
    ms_step = 1e3 * elapsed / args.steps
    log(f"uneven: {value:.1f} pairs/s, {ms_step:.3f} ms/step")
    stages, roofline = stage_roofline(plan, args.stage_iters, ms_step)
    ok = [i for i in range(NP) if res[i].status == abi.XC_OK]
    exact = np.mean([(res[i].lag_x, res[i].lag_y) == data[i][4] for i in ok]) if ok else 0.0
    acc = {"ok": len(ok), "integer_shift_exact": float(exact)}
    cpu = None
    if not args.no_cpu:
        import oracle
        wrong = [i for i in ok if (res[i].lag_x, res[i].lag_y) != data[i][4]][:16]
        sample = list(range(min(NP, 48))) + [i for i in wrong if i >= 48]
        nor = len(sample)
        norm = []
        for i in sample:
            a, ma, b, mb, _ = data[i]
            pa, va = synth.raster_points(a.cpu().numpy(), ma.cpu().numpy())
            pb, vb = synth.raster_points(b.cpu().numpy(), mb.cpu().numpy())
            norm.append((pa, synth.normalize(va), pb, synth.normalize(vb), a.shape))
        prov = fft_provider(oracle)
        t1 = time.perf_counter()
        ref = [oracle.corr_patches(pa, va, pb, vb, 0, 0, 4096, cnt=abi.cnt_enough_points(hw[0] * hw[1] // 4))
               for pa, va, pb, vb, hw in norm]
        cs = time.perf_counter() - t1
        oracle.set_fft_provider(0)
        acc["vs_fp64_oracle"] = accuracy([res[i] for i in sample], ref)
        acc["oracle_checked_pairs_off_true_shift"] = len(wrong)
        cpu = {"value": round(nor / cs, 3), "unit": "pairs/s", "cores": 1, "kind": "port",
               "sample": f"{nor} of the {NP} pairs (the first 48 and those off their true shift), "
                         f"CorrPatches restatement, FFT: {prov}, 1 thread", "seconds": round(cs, 2)}
    line = line_base("uneven masked pair correlations/sec (batched FP32, 4096 pairs 128..1024)", value, "pairs/s",
                     args, ms_step, {"workload": "uneven-size batch with irregular masks", "pairs": NP,
                                     "tile": "w,h U{128..1024} u8, 0-3 holes", "rule": "EnoughPoints(w h / 4)",
                                     "buckets": "next power of two per axis"})
    line["latency_ms"] = {"p50": round(1e3 * float(np.percentile(lat, 50)), 3),
                          "p99": round(1e3 * float(np.percentile(lat, 99)), 3), "unit": "per batch"}
    line["stages"] = stages
    line["autotune"] = plan.buckets()
    sw, sw_us = plan.sweep_waves()
    line["sweep_waves"] = {"waves_per_cu": sw, "tuned_us": round(sw_us, 1)}
    line["roofline"] = roofline
    line["cpu_baseline"] = cpu
    line["accuracy"] = acc
    plan.destroy()
    for ta, tb in tiles:
        ta.release()
        tb.release()
    ctx.close()
    return line


# ---------------------------------------------------------------------------
# serve: continuous batching under Poisson arrivals
# ---------------------------------------------------------------------------

def serve_leg(ctx, sync, rates, seconds, n_uneven=40, n_strips=12, burst=True):
    """Poisson arrivals at each rate (and a burst) over the 70/30 mix of uneven tiles and 2048 x 256 strips on
    small tori of neighbouring tiles.  Returns (runs, server stats, config)."""
    import torch

    from alignment_projects_amd import abi, workloads, xcorr

    t0 = time.time()
    rng = np.random.default_rng(5000)
    keep, tiles, pools = [], [], {"uneven": [], "strips": []}
    for li in range(n_uneven + n_strips):
        kind = "uneven" if li < n_uneven else "strips"
        if kind == "uneven":
            w, h = (int(v) for v in rng.integers(128, 1025, size=2))
            tl, reqs = workloads.lattice(6000 + li, w, h, jitter=25)
        else:
            w, h = 2048, 256
            tl, reqs = workloads.lattice(6000 + li, w, h, jitter=20)
        keep += tl
        th = [ctx.tile_put(t.shape, device_ptr=t.data_ptr(), dtype=0) for t in tl]
        tiles += th
        for ia, ib, s in reqs:
            if kind == "uneven":
                p = xcorr.make_pair(th[ia], th[ib], cnt=abi.cnt_enough_points(w * h // 4))
            else:
                p = xcorr.make_pair(th[ia], th[ib], rgn=abi.rgn_min(w // 2, h // 2))
            pools[kind].append((p, s))
    sync()
    log(f"serving tiles: {len(tiles)}, requests: {len(pools['uneven'])} + {len(pools['strips'])} "
        f"in {time.time() - t0:.1f}s")
    free, _ = torch.cuda.mem_get_info()
    budget = int(min(200e9, 0.85 * free))
    srv = xcorr.Server(ctx, hbm_budget=budget, max_batch=64, streams=3, max_wait_us=200)

    def offered(rate, seconds, seed):
        n = max(200, int(rate * seconds)) if rate else 4000
        r2 = np.random.default_rng(seed)
        kinds = r2.random(n) < 0.7
        picks = [pools["uneven"][r2.integers(len(pools["uneven"]))] if k else
                 pools["strips"][r2.integers(len(pools["strips"]))] for k in kinds]
        at = workloads.poisson_times(rate, n, seed) if rate else np.zeros(n)
        tickets = [None] * n
        lat = [None] * n
        wrong = [0]
        done = threading.Event()

        def collect():
            for i in range(n):
                while tickets[i] is None:
                    time.sleep(20e-6)
                r, l = srv.wait(tickets[i])
                lat[i] = l
                s = picks[i][1]
                if r.status == abi.XC_OK and (r.lag_x, r.lag_y) != s:
                    wrong[0] += 1
            done.set()

        th = threading.Thread(target=collect)
        th.start()
        late = 0.0
        t_start = time.perf_counter()
        for i in range(n):
            if rate:
                while True:
                    now = time.perf_counter() - t_start
                    if now >= at[i]:
                        break
                    if at[i] - now > 200e-6:
                        time.sleep(50e-6)
                late = max(late, time.perf_counter() - t_start - at[i])
            tickets[i] = srv.submit(picks[i][0])
        done.wait()
        t_end = time.perf_counter()
        th.join()
        lat_ms = np.array(lat) / 1e3
        return {"offered_req_s": rate if rate else "burst", "requests": n,
                "sustained_pairs_s": round(n / (t_end - t_start), 1),
                "p50_ms": round(float(np.percentile(lat_ms, 50)), 3),
                "p99_ms": round(float(np.percentile(lat_ms, 99)), 3),
                "max_submit_lag_ms": round(1e3 * late, 3), "wrong_shift": wrong[0]}

    offered(2000, 0.3, 1)  # warm-up: kernels loaded, spectra resident
    runs = []
    for k, rate in enumerate(list(rates) + ([0] if burst else [])):
        r = offered(rate, seconds, 10 + k)
        log(json.dumps(r))
        runs.append(r)
    st = srv.stats()
    srv.close()
    for t in tiles:
        t.release()
    del keep
    cfg = {"mix": "70% w,h U{128..1024} / 30% 2048x256 strips", "tiles": len(tiles),
           "budget_GB": round(budget / 1e9, 1), "max_batch": 64, "streams": 3, "max_wait_us": 200,
           "seconds_per_rate": seconds}
    return runs, st, cfg


def run_serve(args, sync):
    from alignment_projects_amd import xcorr

    ctx = xcorr.Context(0)
    runs, st, cfg = serve_leg(ctx, sync, [2000, 8000, 20000], args.serve_seconds)
    ctx.close()
    value = max(r["sustained_pairs_s"] for r in runs)
    line = line_base("served tile-pair correlations/sec (continuous batching, FP32)", value, "pairs/s", args, 0.0,
                     dict({"workload": "serving stream"}, **cfg))
    line["ms_per_step"] = None
    line["serving"] = runs
    line["server_stats"] = st
    line["latency_ms"] = {"p50": runs[2]["p50_ms"], "p99": runs[2]["p99_ms"], "unit": "per request at 20000 req/s"}
    line["roofline"] = None
    line["cpu_baseline"] = None
    return line


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: measures one GPU only (got --gpus {args.gpus}); pairs are independent, so a batch "
              "shards across GPUs by pair with no exchange", file=sys.stderr)
        sys.exit(2)
    sync = _device_sync()
    run = {"strips": run_strips, "sweep": run_sweep, "uneven": run_uneven, "serve": run_serve}[args.workload]
    line = run(args, sync)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
