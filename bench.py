"""Benchmark: separable Gaussian on one MI355X.

Headline workload: Image<f32,3> 512^3, spacing 1, sigma 2 (radius 6), input
U[0,1) + 32 Gaussian blobs (seed 1), already resident in HBM.  One step = one
GaussianFilter::generate_region over the whole volume.  `value` is
Gvoxel/s over the K timed steps.

The headline runs in FAST arithmetic (every tap a fused multiply-add; the
reference's weights, truncation, clamping and pass order; within 1e-5 of the
reference, the accuracy the target is stated at).  Also reported (under
"extras", one line of JSON in total):
  - EXACT arithmetic (bit-identical to itk-rs) on the same volume,
  - the max error of both modes against an f64 evaluation of the same
    operator (and of FAST against EXACT), relative to the max |output|,
  - forward + adjoint at 256^3 (sigma 1.8) vs forward only,
  - the batched form (64 uneven volumes, persistent launches per radius
    group in one captured graph, tile shapes and z runs picked by the
    autotuner) vs a loop of plain single-volume launches, and the batched
    adjoint over 16 uneven gradients vs a loop of plain adjoint launches,
  - 512^2 2-D and the 256x256x160 VecPixel<f32,3> configuration,
  - MeanFilter (radius 2, FAST and EXACT), BinomialBlurFilter (4
    repetitions), UnsharpMaskFilter (sigma 2) and RecursiveGaussianFilter
    (sigma 2) on a 512^3 volume; CannyEdgeDetectionFilter on a 2048^2 image,
  - host buffers end to end: the host-buffer entry point on the 512^3
    volume through the pinned staging ring and through pageable copies, and
    the u8 -> f32 -> Gaussian -> u8 chain of examples/blur_png.rs on a 512^3
    u8 volume and on one 512 x 512 image (the example's own use).

roofline: algorithmic bytes = 8 B per output voxel (one f32 read, one f32
write) / the average kernel duration from HIP events on the launch stream,
against 8 TB/s.  cpu_baseline: the oracle's threaded restatement of the
reference's loop structure on the host cores, over the same whole volume;
beside it (cpu_baseline.optimised) the same arithmetic in a cache-friendly
loop order, the comparison an optimised CPU port gives.

Before the W warmup steps the headline launch runs untimed for about 200 ms
of GPU time (`settle`; 50 ms before each extras leg): the GPU needs tens of
milliseconds of sustained load to reach its steady clocks (after 3 launches
the kernel still runs about 12% slower, and a gap of tens of milliseconds
drops them again), and 1000 timed launches after the settle average the same
as 100.

    python bench.py [--gpus 1] [--steps 100] [--warmup 200] [--arith fast|exact]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--arith", choices=["exact", "fast"], default="fast")
    ap.add_argument("--no-extras", action="store_true", help="headline loop only (profiling)")
    ap.add_argument("--extras", default="",
                    help="comma list of extras groups to run (default all): "
                         "c2, c5, c4, c5_batch, c1_c3, filters, aniso, host")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--keep-gc", action="store_true",
                    help="leave Python's garbage collector as it is (no collection before a "
                         "timed region, on inside it) and report its pauses (gc_pauses_ms)")
    ap.add_argument("--traffic-json", default=os.path.join(ROOT, "profiles", "traffic.json"),
                    help="rocprof-derived HBM bytes per launch, if measured")
    ap.add_argument("--trace-loops", type=int, default=0, metavar="N",
                    help="instead of the bench: N reps each of the BinomialBlurFilter call and "
                         "the 16-gradient plain adjoint loop with the collector as it is, "
                         "recording each rep's host and GPU time and every collector pause; "
                         "prints the reps slower than twice the median and the pauses inside "
                         "them")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one 512 MiB device-to-device copy (a counter reference "
                         "of known size for rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE)")
    return ap.parse_args()


def settle(fn, stream, gpu_ms=200.0, max_calls=4000):
    """Runs fn untimed until about `gpu_ms` of GPU time has passed (the
    clocks ramp up under sustained load)."""
    import torch
    done, calls = 0.0, 0
    while done < gpu_ms and calls < max_calls:
        # chunks of 20 calls, timed on the stream (the first call of a
        # process also loads code objects, so no single call sizes the rest)
        t0 = torch.cuda.Event(enable_timing=True)
        t1 = torch.cuda.Event(enable_timing=True)
        t0.record(stream)
        for _ in range(20):
            fn()
        t1.record(stream)
        torch.cuda.synchronize()
        done += t0.elapsed_time(t1)
        calls += 20


def max_rel_err(got, ref64):
    """max |got - ref| / max |ref| (f64 reference)."""
    return float((got.double() - ref64).abs().max()) / max(float(ref64.abs().max()), 1e-300)


def host_threads():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


GC_PAUSES_MS = []  # Python GC pauses inside timed regions (recorded with --keep-gc)
KEEP_GC = False


def _gc_callback(phase, info, _t={}):
    if phase == "start":
        _t["s"] = time.perf_counter()
    elif "s" in _t:
        GC_PAUSES_MS.append((time.perf_counter() - _t.pop("s")) * 1e3)


def timed_loop(fn, steps, warmup, stream, flush=None, settle_ms=50.0):
    """Runs fn untimed for about `settle_ms` of GPU time and `warmup` times,
    then `steps` timed times; returns (total wall seconds between
    synchronisations, per-launch event milliseconds).

    Python's cyclic garbage collector is off inside the timed region (as in
    the standard library's timeit): a full collection of this process takes
    tens of milliseconds, and one that lands between the start event and a
    rep's launches leaves the GPU idle for that long.  A collection runs just
    before the region instead; it idles the GPU too, and its clocks drop after
    so long a gap, so the settle runs after it.  --keep-gc leaves the
    collector as it is and records each collection's pause."""
    import gc
    import torch
    if not KEEP_GC:
        gc.collect()
        gc.disable()
    elif _gc_callback not in gc.callbacks:
        gc.callbacks.append(_gc_callback)
    try:
        if settle_ms > 0:
            settle(fn, stream, settle_ms)
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        evs = []
        t0 = time.perf_counter()
        for _ in range(steps):
            if flush is not None:
                flush()
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record(stream)
            fn()
            b.record(stream)
            evs.append((a, b))
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
    finally:
        gc.enable()
        if _gc_callback in gc.callbacks:
            gc.callbacks.remove(_gc_callback)
    return wall, [a.elapsed_time(b) for a, b in evs]


def trace_loops(itk, wl, dev, stream, reps):
    """Per-rep host and GPU times of two loop legs with Python's collector
    left as it is, and the collector's pauses that overlap slow reps."""
    import gc
    import numpy as np
    import torch
    pauses = []
    mark = {}

    def cb(phase, info):
        if phase == "start":
            mark["t"] = time.perf_counter()
        elif "t" in mark:
            pauses.append((mark.pop("t"), time.perf_counter(), info.get("generation")))
    gc.callbacks.append(cb)

    def run(name, fn):
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        pauses.clear()
        recs = []
        for _ in range(reps):
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            h0 = time.perf_counter()
            a.record(stream)
            fn()
            b.record(stream)
            recs.append((h0, time.perf_counter(), a, b))
        torch.cuda.synchronize()
        gpu = np.array([a.elapsed_time(b) for _, _, a, b in recs])
        host = np.array([(h1 - h0) * 1e3 for h0, h1, _, _ in recs])
        slow = [i for i in range(reps) if gpu[i] > 2 * np.median(gpu)]
        print(json.dumps({
            "leg": name, "reps": reps, "gpu_ms_median": float(np.median(gpu)),
            "gpu_ms_mean": float(gpu.mean()), "gpu_ms_max": float(gpu.max()),
            "host_ms_median": float(np.median(host)), "host_ms_max": float(host.max()),
            "collections": len(pauses),
            "longest_pause_ms": max([(e - s) * 1e3 for s, e, _ in pauses], default=0.0),
            "slow_reps": [{"rep": i, "gpu_ms": float(gpu[i]), "host_ms": float(host[i]),
                           "pauses_inside_ms": [round((e - s) * 1e3, 3) for s, e, _ in pauses
                                                if s < recs[i][1] and e > recs[i][0]]}
                          for i in slow[:8]]}), flush=True)

    xm = torch.rand((512, 512, 512), device=dev)
    im = itk.Image(itk.Region([0, 0, 0], [512] * 3), (1.0,) * 3, (0.0,) * 3, xm)
    om = itk.Image(im.region, (1.0,) * 3, (0.0,) * 3, torch.empty_like(xm))
    bf = itk.BinomialBlurFilter(im, 4)
    bf.arith = itk.FAST
    run("binomial_n4_512", lambda: bf.generate_into(im, om))
    del xm, im, om
    torch.cuda.empty_cache()
    grads, gparams = wl.c5_batch(dev)
    gouts = [torch.empty_like(q) for q in grads]

    def gloop():
        for q, o, p in zip(grads, gouts, gparams):
            itk.gaussian_backward(q, p[1], p[2], out=o)
    run("c5_plain_adjoint_loop_16", gloop)
    gc.callbacks.remove(cb)


def main():
    args = parse()
    global KEEP_GC
    KEEP_GC = args.keep_gc
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are not part "
                                   "of it (--gpus must be 1)"}))
        sys.exit(2)
    import numpy as np
    import torch

    from __graft_entry__ import load_package
    itk = load_package()
    from itk_rs_amd import workloads as wl

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream(dev)
    arith = itk.EXACT if args.arith == "exact" else itk.FAST
    if args.trace_loops:
        trace_loops(itk, wl, dev, stream, args.trace_loops)
        return

    # ---- headline: 512^3 f32, sigma 2 -------------------------------------
    x, sigma, sp = wl.c2(dev)
    n = x.numel()
    out = torch.empty_like(x)
    img = itk.Image(itk.Region([0, 0, 0], [512, 512, 512]), sp, (0.0,) * 3, x)
    oimg = itk.Image(img.region, sp, (0.0,) * 3, out)
    f = itk.GaussianFilter(img, sigma)
    f.arith = arith
    step = lambda: f.generate_into(img, oimg)  # noqa: E731
    if args.copy_calibration:
        # 512 MiB read + 512 MiB written by one elementwise kernel (a plain
        # copy_ may go to a DMA engine, which the counters do not see)
        torch.mul(x, 1.0, out=out)
        torch.cuda.synchronize()
    wall, ms = timed_loop(step, args.steps, args.warmup, stream, settle_ms=200.0)
    ms_avg = float(np.mean(ms))
    gvox = n * args.steps / wall / 1e9
    achieved = 8.0 * n / (ms_avg * 1e-3) / 1e9
    traffic = None
    if os.path.exists(args.traffic_json):
        with open(args.traffic_json) as fh:
            tj = json.load(fh)
        traffic = tj.get("c2_" + args.arith, {}).get("bytes_per_launch")
    result = {
        "metric": "Gvoxel/s and %HBM roofline, 512^3 f32 Gaussian (sigma 2), plain forward",
        "value": round(gvox, 3),
        "unit": "Gvoxel/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded U[0,1) + 32 Gaussian blobs, generated on device)",
        "config": {"workload": "GaussianFilter::generate_region, Image<f32,3> 512x512x512, "
                               "spacing 1, sigma 2 (radius 6), whole image",
                   "arith": args.arith, "voxels": n},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "kernel_ms": round(ms_avg, 4),
                     "algorithmic_bytes_per_launch": 8 * n},
    }

    extras = {}
    chosen = set(args.extras.split(",")) if args.extras else None

    def want(name):
        return chosen is None or name in chosen

    if not args.no_extras:
        if want("c2"):
            # reference point: one device copy of the same 512 MiB (read + write),
            # the rate a bare stream of these bytes reaches on this device
            _, msc = timed_loop(lambda: out.copy_(x), args.steps, 2, stream)
            copy_gbs = 8.0 * n / (float(np.mean(msc)) * 1e-3) / 1e9
            extras["device_copy_512MiB"] = {"gb_s": round(copy_gbs, 1),
                                            "hbm_frac": round(copy_gbs / HBM_PEAK_GBS, 4),
                                            "headline_vs_copy": round(achieved / copy_gbs, 4)}
            # the other arithmetic mode on the same volume
            other = itk.FAST if arith == itk.EXACT else itk.EXACT
            f2 = itk.GaussianFilter(img, sigma)
            f2.arith = other
            w2, ms2 = timed_loop(lambda: f2.generate_into(img, oimg), args.steps, 2, stream)
            extras["c2_" + ("fast" if other == itk.FAST else "exact")] = {
                "gvox_s": round(n * args.steps / w2 / 1e9, 3),
                "kernel_ms": round(float(np.mean(ms2)), 4),
                "hbm_frac": round(8.0 * n / (np.mean(ms2) * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
            # 8-piece streamed form (device-resident stream_full)
            ws, _ = timed_loop(lambda: itk.stream_full(f, 8), 5, 1, stream)
            extras["c2_stream_full_8"] = {"gvox_s": round(n * 5 / ws / 1e9, 3)}
            # accuracy of both modes against the f64 evaluation of the operator
            outs = {}
            for name, mode in (("fast", itk.FAST), ("exact", itk.EXACT)):
                outs[name] = itk.gaussian(x, sigma, sp, arith=mode)
            ref64 = itk.gaussian(x.double(), sigma, sp)
            scale = float(ref64.abs().max())
            extras["c2_max_rel_err"] = {
                "fast_vs_f64": float((outs["fast"].double() - ref64).abs().max()) / scale,
                "exact_vs_f64": float((outs["exact"].double() - ref64).abs().max()) / scale,
                "fast_vs_exact": float((outs["fast"] - outs["exact"]).abs().max()) /
                                 float(outs["exact"].abs().max())}
            del outs, ref64, out, oimg
            torch.cuda.empty_cache()

        if want("c5"):
            # forward + adjoint at 256^3 (MALL-resident sizes: flush between reps)
            x5, g5, s5, sp5 = wl.c5(dev)
            n5 = x5.numel()
            flushbuf = torch.empty(512 * 1024 * 1024 // 4, dtype=torch.float32, device=dev)
            flush = lambda: flushbuf.fill_(1.0)  # noqa: E731
            y5 = torch.empty_like(x5)
            gi5 = torch.empty_like(x5)
            im5 = itk.Image(itk.Region([0, 0, 0], [256] * 3), sp5, (0.0,) * 3, x5)
            om5 = itk.Image(im5.region, sp5, (0.0,) * 3, y5)
            go5 = itk.Image(im5.region, sp5, (0.0,) * 3, g5)
            gin5 = itk.Image(im5.region, sp5, (0.0,) * 3, gi5)
            f5 = itk.GaussianFilter(im5, s5)
            f5.arith = arith
            _, msf = timed_loop(lambda: f5.generate_into(im5, om5), 10, 2, stream, flush)

            def fwd_adj():
                f5.generate_into(im5, om5)
                f5.backward_into(go5, gin5, False)
            _, msfa = timed_loop(fwd_adj, 10, 2, stream, flush)
            fwd_g = n5 / (np.mean(msf) * 1e-3) / 1e9
            fa_g = n5 / (np.mean(msfa) * 1e-3) / 1e9
            extras["c5_fwd_adjoint_256"] = {"fwd_gvox_s": round(fwd_g, 3),
                                            "fwd_plus_adj_gvox_s": round(fa_g, 3),
                                            "ratio": round(fa_g / fwd_g, 3)}
            # error of the timed f32 forward and adjoint against f64 evaluations
            ref_f = itk.gaussian(x5.double(), s5, sp5)
            ref_a = itk.gaussian_backward(g5.double(), s5, sp5)
            extras["c5_max_rel_err_vs_f64"] = {"forward": max_rel_err(y5, ref_f),
                                               "adjoint": max_rel_err(gi5, ref_a)}
            del ref_f, ref_a
            del x5, g5, y5, gi5, im5, om5, go5, gin5, flushbuf
            torch.cuda.empty_cache()

        def rep_stats(wall, ms, reps, voxels):
            """Gvox/s from the wall clock and from the median rep, with the
            mean and median rep times (a slow outlier shows as mean >> median)."""
            return {"gvox_s": round(voxels * reps / wall / 1e9, 3),
                    "gvox_s_median": round(voxels / (float(np.median(ms)) * 1e-3) / 1e9, 3),
                    "ms_mean": round(float(np.mean(ms)), 4),
                    "ms_median": round(float(np.median(ms)), 4)}

        if want("c4"):
            # batched: 64 uneven volumes vs a loop of plain launches
            vols, params = wl.c4(dev)
            outs = [torch.empty_like(v) for v in vols]
            plan = itk.BatchPlan(vols, outs, [p[1] for p in params], [p[2] for p in params],
                                 direction=itk.BATCH_FORWARD, arith=arith)
            nb = plan.voxels
            plan.autotune(stream)  # z run per radius group, measured on these volumes
            wb, msb4 = timed_loop(lambda: plan.run(stream), 10, 2, stream)
            imgs = [itk.Image(itk.Region([0, 0, 0], list(reversed(v.shape))), p[2], (0.0,) * 3, v)
                    for v, p in zip(vols, params)]
            oims = [itk.Image(i.region, i.spacing, i.origin, o) for i, o in zip(imgs, outs)]
            fl = [itk.GaussianFilter(i, p[1]) for i, p in zip(imgs, params)]
            for q in fl:
                q.arith = arith

            def loop():
                for q, i, o in zip(fl, imgs, oims):
                    q.generate_into(i, o)
            wl_, msl4 = timed_loop(loop, 10, 2, stream)
            bs, ls = rep_stats(wb, msb4, 10, nb), rep_stats(wl_, msl4, 10, nb)
            extras["c4_batched_64"] = {"voxels": nb, "batched_gvox_s": bs["gvox_s"],
                                       "loop_gvox_s": ls["gvox_s"],
                                       "speedup": round(wl_ / wb, 3),
                                       "speedup_median": round(ls["ms_median"] / bs["ms_median"], 3),
                                       "batched": bs, "loop": ls}
            # error of the batched outputs against f64 evaluations, max over volumes
            plan.run(stream)
            err = 0.0
            for v, o, p in zip(vols, outs, params):
                err = max(err, max_rel_err(o, itk.gaussian(v.double(), p[1], p[2])))
            extras["c4_max_rel_err_vs_f64"] = err
            plan.close()
            del vols, outs, imgs, oims, fl
            torch.cuda.empty_cache()

        if want("c5_batch"):
            # batched adjoint: 16 uneven gradients (extents 48..192), sigma 1.8
            grads, gparams = wl.c5_batch(dev)
            gouts = [torch.empty_like(q) for q in grads]
            gplan = itk.BatchPlan(grads, gouts, [p[1] for p in gparams], [p[2] for p in gparams],
                                  direction=itk.BATCH_BACKWARD)
            ng = gplan.voxels
            gplan.autotune(stream)
            wgb, msgb = timed_loop(lambda: gplan.run(stream), 10, 2, stream)

            def gloop():
                for q, o, p in zip(grads, gouts, gparams):
                    itk.gaussian_backward(q, p[1], p[2], out=o)
            wgl, msgl = timed_loop(gloop, 10, 2, stream)
            bs, ls = rep_stats(wgb, msgb, 10, ng), rep_stats(wgl, msgl, 10, ng)
            extras["c5_batched_adjoint_16"] = {"voxels": ng, "batched_gvox_s": bs["gvox_s"],
                                               "loop_gvox_s": ls["gvox_s"],
                                               "speedup": round(wgl / wgb, 3),
                                               "speedup_median": round(ls["ms_median"] /
                                                                       bs["ms_median"], 3),
                                               "batched": bs, "loop": ls}
            gplan.run(stream)
            err = 0.0
            for q, o, p in zip(grads, gouts, gparams):
                err = max(err, max_rel_err(o, itk.gaussian_backward(q.double(), p[1], p[2])))
            extras["c5_batch_max_rel_err_vs_f64"] = err
            gplan.close()
            del grads, gouts
            torch.cuda.empty_cache()

        if want("c1_c3"):
            # 512^2 2-D and the VecPixel configuration
            x1, s1, sp1 = wl.c1(dev)
            w1, ms1 = timed_loop(lambda: itk.gaussian(x1, s1, sp1, arith=arith), 20, 3, stream)
            extras["c1_512sq"] = {"gvox_s": round(x1.numel() * 20 / w1 / 1e9, 3),
                                  "ms_median": round(float(np.median(ms1)), 4)}
            # the same calls back to back: wall time between two
            # synchronisations, without timed_loop's two events a call (for
            # an image this small they are most of a step's host time)
            import gc
            if not KEEP_GC:
                gc.collect()
                gc.disable()
            try:
                torch.cuda.synchronize()
                tb0 = time.perf_counter()
                for _ in range(500):
                    itk.gaussian(x1, s1, sp1, arith=arith)
                torch.cuda.synchronize()
                tb = (time.perf_counter() - tb0) / 500
            finally:
                if not KEEP_GC:
                    gc.enable()
            extras["c1_512sq"]["us_per_call_back_to_back"] = round(tb * 1e6, 2)
            extras["c1_512sq"]["gvox_s_back_to_back"] = round(x1.numel() / tb / 1e9, 3)
            extras["c1_max_rel_err_vs_f64"] = max_rel_err(itk.gaussian(x1, s1, sp1, arith=arith),
                                                          itk.gaussian(x1.double(), s1, sp1))
            v3, s3, sp3 = wl.c3(dev)
            o3 = torch.empty_like(v3)
            i3 = itk.Image(itk.Region([0, 0, 0], [256, 256, 160]), sp3, (0.0,) * 3, v3)
            oi3 = itk.Image(i3.region, sp3, (0.0,) * 3, o3)
            f3 = itk.GaussianFilter(i3, s3)
            f3.arith = arith
            # 252 MB in + out fits in the 256 MB MALL: flush it before every rep
            flushbuf = torch.empty(512 * 1024 * 1024 // 4, dtype=torch.float32, device=dev)
            w3, ms3 = timed_loop(lambda: f3.generate_into(i3, oi3), 10, 2, stream,
                                 lambda: flushbuf.fill_(1.0))
            n3 = v3.numel() // 3
            extras["c3_vec3_256x256x160"] = {"gvox_s": round(n3 / (np.mean(ms3) * 1e-3) / 1e9, 3),
                                             "kernel_ms": round(float(np.mean(ms3)), 4),
                                             "hbm_frac_24B": round(24.0 * n3 / (np.mean(ms3) * 1e-3)
                                                                   / 1e9 / HBM_PEAK_GBS, 4)}
            extras["c3_max_rel_err_vs_f64"] = max_rel_err(o3, itk.gaussian(v3.double(), s3, sp3,
                                                                           ndim=3))
            del v3, o3, i3, oi3, flushbuf
            torch.cuda.empty_cache()

        if want("filters"):
            # the other convolve_axis filters on the same 512^3 volume: MeanFilter
            # (radius 2, one fused round trip) and BinomialBlurFilter (4 rounds,
            # one radius-1 round trip each)
            xm = torch.rand((512, 512, 512), device=dev)
            ym = torch.empty_like(xm)
            im = itk.Image(itk.Region([0, 0, 0], [512] * 3), (1.0,) * 3, (0.0,) * 3, xm)
            om = itk.Image(im.region, (1.0,) * 3, (0.0,) * 3, ym)
            nm = xm.numel()
            for name, mode in (("fast", itk.FAST), ("exact", itk.EXACT)):
                mf = itk.MeanFilter(im, 2)
                mf.arith = mode
                wm, msm = timed_loop(lambda: mf.generate_into(im, om), 20, 3, stream)
                extras["mean_r2_512_" + name] = {
                    "gvox_s": round(nm * 20 / wm / 1e9, 3), "kernel_ms": round(float(np.mean(msm)), 4),
                    "hbm_frac": round(8.0 * nm / (np.mean(msm) * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
            bf = itk.BinomialBlurFilter(im, 4)
            bf.arith = arith
            wbn, msb = timed_loop(lambda: bf.generate_into(im, om), 20, 3, stream)
            extras["binomial_n4_512"] = {
                "gvox_s": round(nm * 20 / wbn / 1e9, 3), "ms": round(float(np.mean(msb)), 4),
                "ms_median": round(float(np.median(msb)), 4), "ms_min": round(float(np.min(msb)), 4),
                "hbm_frac_4_round_trips": round(4 * 8.0 * nm / (np.mean(msb) * 1e-3) / 1e9
                                                / HBM_PEAK_GBS, 4)}
            # UnsharpMaskFilter, sigma 2 in pixels (radius 6): one launch, the
            # per-pixel step in the blur's epilogue; the source pixel at each
            # output's index is read again there, from memory (the window
            # that held it left L2 steps before): 4 + 4 + 4 B per voxel
            uf = itk.UnsharpMaskFilter(im, 2.0)
            uf.arith = arith
            wu, msu = timed_loop(lambda: uf.generate_into(im, om), 20, 3, stream)
            extras["unsharp_s2_512"] = {
                "gvox_s": round(nm * 20 / wu / 1e9, 3), "ms": round(float(np.mean(msu)), 4),
                "hbm_frac_12B": round(12.0 * nm / (np.mean(msu) * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
            # RecursiveGaussianFilter, sigma 2: a causal and an anticausal
            # line pass per axis over an f64 work image, the first reading the
            # f32 input and the last writing the f32 output; each pass reads
            # its input twice (the anticausal pass recomputes the causal
            # values chunk by chunk from checkpointed states: 48 B written and
            # read per chunk of 32 samples) and writes once:
            # (4 + 4 + 8 + 1.5) + (8 + 8 + 8 + 1.5) + (8 + 8 + 4 + 1.5)
            # = 64.5 B per voxel
            rf = itk.RecursiveGaussianFilter(im, 2.0)
            wr, msr = timed_loop(lambda: rf.generate_into(im, om), 10, 2, stream)
            extras["recursive_s2_512"] = {
                "gvox_s": round(nm * 10 / wr / 1e9, 3), "ms": round(float(np.mean(msr)), 4),
                "bytes_per_voxel": 64.5,
                "hbm_frac": round(64.5 * nm / (np.mean(msr) * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
            del xm, ym, im, om
            torch.cuda.empty_cache()
            # CannyEdgeDetectionFilter on a 2048^2 disc with noise (sigma 1.5):
            # smoothing, gradient, suppression, hysteresis passes to the fixed
            # point (each pass waits on the stream)
            yy, xx = torch.meshgrid(torch.arange(2048, device=dev), torch.arange(2048, device=dev),
                                    indexing="ij")
            disc = (((xx - 1024.0) ** 2 + (yy - 1000.0) ** 2) < 600.0 ** 2).float()
            xc = disc + 0.05 * torch.rand((2048, 2048), device=dev)
            ic = itk.Image(itk.Region([0, 0], [2048, 2048]), (1.0, 1.0), (0.0, 0.0), xc)
            oc = itk.Image(ic.region, (1.0, 1.0), (0.0, 0.0), torch.empty_like(xc))
            cf = itk.CannyEdgeDetectionFilter(ic, 1.5, 0.05, 0.2)
            wc, msc2 = timed_loop(lambda: cf.generate_into(ic, oc), 10, 2, stream)
            extras["canny_2048sq"] = {"mpix_s": round(xc.numel() * 10 / wc / 1e6, 1),
                                      "ms": round(wc / 10 * 1e3, 4),
                                      "edge_pixels": int(oc.data.sum().item())}
            del xc, oc, ic, disc, xx, yy
            torch.cuda.empty_cache()

        if want("aniso"):
            # anisotropic spacing: slices twice the in-plane spacing give radii
            # (6, 6, 3) at sigma 2 (512 x 512 x 256 f32), on the fused kernel and
            # on the per-axis path (three round trips) that served such
            # volumes before; and the per-axis path on the 512^3 headline volume
            xa = torch.rand((256, 512, 512), device=dev)
            ya = torch.empty_like(xa)
            spa = (1.0, 1.0, 2.0)
            ia = itk.Image(itk.Region([0, 0, 0], [512, 512, 256]), spa, (0.0,) * 3, xa)
            oa = itk.Image(ia.region, spa, (0.0,) * 3, ya)
            na = xa.numel()
            res = {}
            for name, path in (("fused", itk.PATH_FUSED), ("per_axis", itk.PATH_PER_AXIS)):
                fa = itk.GaussianFilter(ia, 2.0)
                fa.arith = arith
                itk.set_path(path)
                try:
                    wa, msa = timed_loop(lambda: fa.generate_into(ia, oa), 20, 3, stream)
                finally:
                    itk.set_path(itk.PATH_AUTO)
                res[name] = {"gvox_s": round(na * 20 / wa / 1e9, 3),
                             "kernel_ms": round(float(np.mean(msa)), 4),
                             "hbm_frac": round(8.0 * na / (np.mean(msa) * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
            extras["aniso_r663_512x512x256"] = res
            extras["aniso_max_rel_err_vs_f64"] = max_rel_err(ya, itk.gaussian(xa.double(), 2.0, spa))
            del xa, ya, ia, oa
            torch.cuda.empty_cache()
            xp = torch.rand((512, 512, 512), device=dev)
            yp = torch.empty_like(xp)
            ip = itk.Image(itk.Region([0, 0, 0], [512] * 3), (1.0,) * 3, (0.0,) * 3, xp)
            op = itk.Image(ip.region, (1.0,) * 3, (0.0,) * 3, yp)
            fp = itk.GaussianFilter(ip, 2.0)
            fp.arith = arith
            itk.set_path(itk.PATH_PER_AXIS)
            try:
                wp, msp = timed_loop(lambda: fp.generate_into(ip, op), 10, 3, stream)
            finally:
                itk.set_path(itk.PATH_AUTO)
            extras["per_axis_r6_512"] = {"gvox_s": round(xp.numel() * 10 / wp / 1e9, 3),
                                         "kernel_ms": round(float(np.mean(msp)), 4)}
            del xp, yp, ip, op
            torch.cuda.empty_cache()

        if want("host"):
            # host buffers end to end (the drop-in's itkg_gaussian_generate_region_host
            # on the whole 512^3 f32 volume: 512 MiB in over PCIe, the kernels,
            # 512 MiB out), through the pinned staging ring and through pageable
            # copies; then the examples/blur_png.rs chain (u8 -> f32, Gaussian,
            # round -> u8) on a 512^3 u8 volume, host to host
            import ctypes
            hx = x.cpu().numpy()
            hy = np.empty_like(hx)
            reg_c = itk.Region([0, 0, 0], [512] * 3).to_c()
            sp_c = (ctypes.c_double * 3)(*sp)

            def host_call():
                rc = itk.lib.itkg_gaussian_generate_region_host(
                    3, itk.F32, 1, float(sigma), sp_c, ctypes.byref(reg_c), hx.ctypes.data,
                    ctypes.byref(reg_c), hy.ctypes.data, arith)
                if rc:
                    raise RuntimeError(itk.lib.itkg_last_error())

            def host_times(fn, reps):
                fn()  # creates the staging ring / maps the pool on the first call
                ts = []
                for _ in range(reps):
                    t0 = time.perf_counter()
                    fn()
                    ts.append(time.perf_counter() - t0)
                return ts

            hres = {}
            for mode in ("pinned", "pageable"):
                if mode == "pageable":
                    os.environ["ITKG_HOST_STAGING"] = "pageable"
                else:
                    os.environ.pop("ITKG_HOST_STAGING", None)
                ts = host_times(host_call, 5)
                hres[mode] = {"ms_median": round(float(np.median(ts)) * 1e3, 2),
                              "ms_min": round(float(np.min(ts)) * 1e3, 2),
                              "pcie_gb_s_both_ways": round(2 * 4.0 * n / float(np.median(ts))
                                                           / 1e9, 1),
                              "staged_pinned": itk.host_staging_pinned()}
            os.environ.pop("ITKG_HOST_STAGING", None)
            hres["bit_equal_to_device"] = bool(np.array_equal(
                hy, itk.gaussian(x, sigma, sp, arith=arith).cpu().numpy()))
            extras["host_to_host_512_f32"] = hres
            del hx, hy
            u8 = (torch.rand((512, 512, 512), device=dev) * 255).to(torch.uint8).cpu().numpy()
            ts = host_times(lambda: itk.cast_gaussian_cast_host(u8, sigma, np.uint8, arith=arith), 5)
            extras["host_to_host_512_u8_chain"] = {
                "ms_median": round(float(np.median(ts)) * 1e3, 2),
                "ms_min": round(float(np.min(ts)) * 1e3, 2),
                "gvox_s": round(n / float(np.median(ts)) / 1e9, 3)}
            del u8
            # the casts of that chain on device-resident 512^3 images
            # (CastFilter, itkg_convert_pixels): 1 + 4 bytes per voxel
            reg3 = itk.Region([0, 0, 0], [512] * 3)
            xu = torch.randint(0, 256, (512, 512, 512), dtype=torch.uint8, device=dev)
            yf = torch.empty((512, 512, 512), dtype=torch.float32, device=dev)
            iu = itk.Image(reg3, (1.0,) * 3, (0.0,) * 3, xu)
            of = itk.Image(reg3, (1.0,) * 3, (0.0,) * 3, yf)
            to_f = itk.CastFilter(iu, torch.float32)
            to_u = itk.CastFilter(of, torch.uint8, itk.CONVERT_ROUND)
            _, msf = timed_loop(lambda: to_f.generate_into(iu, of), 20, 3, stream)
            _, msu = timed_loop(lambda: to_u.generate_into(of, iu), 20, 3, stream)
            extras["cast_device_512"] = {
                "u8_to_f32_ms": round(float(np.mean(msf)), 4),
                "f32_to_u8_round_ms": round(float(np.mean(msu)), 4),
                "hbm_frac_5B": round(5.0 * n / (float(np.mean(msf)) * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
            del xu, yf, iu, of
            # the reference's CLI case (examples/blur_png.rs): one 512 x 512 u8
            # image, host to host, one piece (no helper threads)
            png = (torch.rand((512, 512), device=dev) * 255).to(torch.uint8).cpu().numpy()
            ts = host_times(lambda: itk.cast_gaussian_cast_host(png, sigma, np.uint8, arith=arith),
                            50)
            extras["host_to_host_512sq_u8_chain"] = {
                "us_median": round(float(np.median(ts)) * 1e6, 1),
                "us_min": round(float(np.min(ts)) * 1e6, 1)}
            itk.trim_scratch(0)

        if KEEP_GC:
            extras["gc_pauses_ms"] = {"count": len(GC_PAUSES_MS),
                                      "max": round(max(GC_PAUSES_MS, default=0.0), 3),
                                      "total": round(sum(GC_PAUSES_MS), 3)}
        result["extras"] = extras

    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        host = x.cpu().numpy()
        threads = host_threads()
        times = []
        for _ in range(3):
            t0 = time.perf_counter()
            oracle.ref_faithful(host, [0, 0, 0], sp, sigma, [0, 0, 0], [512, 512, 512], 1,
                                threads)
            times.append(time.perf_counter() - t0)
        tbest = float(min(times))
        cpu_g = n / tbest / 1e9
        result["cpu_baseline"] = {
            "value": round(cpu_g, 5), "unit": "Gvoxel/s", "cores": threads, "kind": "port",
            "sample": "the whole 512^3 sigma-2 volume through the oracle's threaded "
                      "restatement of the reference loop structure (leaf copy, materialised "
                      "index list, per-voxel parallel map, EXACT arithmetic); best of 3 runs: "
                      + ", ".join(f"{t:.2f} s" for t in times),
            "gpu_over_cpu": round(gvox / cpu_g, 1)}
        # beside it, the same arithmetic in a cache-friendly loop order (the
        # comparison a straightforward optimised CPU port gives)
        oout = np.empty_like(host)
        owork = np.empty(2 * host.size, dtype=np.float32)
        otimes = []
        for _ in range(4):  # the first call touches the output and work pages
            t0 = time.perf_counter()
            oracle.separable_opt(host, sp, sigma, threads, oout, owork)
            otimes.append(time.perf_counter() - t0)
        obest = float(min(otimes[1:]))
        opt_g = n / obest / 1e9
        result["cpu_baseline"]["optimised"] = {
            "value": round(opt_g, 4), "unit": "Gvoxel/s", "cores": threads,
            "sample": "the same volume and EXACT arithmetic (bit-identical), rows along x in "
                      "every pass, planes over the host threads, output and intermediates "
                      "preallocated; best of 3 runs after one untimed: "
                      + ", ".join(f"{t:.3f} s" for t in otimes[1:]),
            "gpu_over_cpu": round(gvox / opt_g, 1)}
        del oout, owork
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
