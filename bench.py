#!/usr/bin/env python3
"""bench.py -- inpainted MPix/s of the 1024x1024 nlpoisson workload on one MI355X.

One step = complete multiscale inpaintings (ImageInpainting::process: 11 scales,
every inpainting iteration until the stopping rule, each a PatchMatch NNF search
plus the NL-Poisson update with its CG solve) of a batch of 1024x1024 RGB float32
images (distinct seeded textures, 15 % union-of-ellipses holes); lambda 0.05,
patch 9, <= 200 iterations per scale, CG tolerance 1e-4 with <= 100 iterations,
FP32 CG with FP64 reductions and FP64 residual replacement.  The batch runs on
concurrent device-resident drivers (nli_inpaint_batch_dev); inputs are resident in
device memory before the timed region.  Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import contextlib
import ctypes
import importlib
import json
import os
import sys
import time

import numpy as np

# kernel arguments in device memory (the HIP runtime's default on this hardware, made
# explicit): the CG's short dependent kernels pay for every argument fetch otherwise
os.environ.setdefault("HIP_FORCE_DEV_KERNARG", "1")

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "inpainted MPix/s at 1024² nlpoisson (batched FP32+FP64-refine vs plain, same GPU)"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3


TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "k_pm_sweep_rgb1024_nlpoisson_traffic.md")


def measured_traffic():
    """HBM bytes per sweep from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary
    (copy-calibrated), or None if it is absent."""
    try:
        for line in open(TRAFFIC_SUMMARY):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def psnr(a, b):
    mse = float(np.mean((a.astype(np.float64) - b.astype(np.float64)) ** 2))
    return float("inf") if mse == 0 else 10 * np.log10(1.0 / mse)


def workload_params(nli, precision, iters=200, tol=0.1):
    return nli.Inpainting(method="nlpoisson", patch=9, sigma=10000.0, lam=0.05, iters=iters, tol=tol, scales=11,
                          coarse=0.1, conf_t=5.0, conf_a=0.1, init="poisson", pm_iters=10, pm_shots=30,
                          pm_window=-1, cg_tol=1e-4, cg_limit=100, cg_precision=precision,
                          cg_refine_every=8 if precision == nli.CG_MIXED else 0, seed=1)


def host_cpu():
    """CPU model name and logical CPU count of this host."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, os.cpu_count()


REF_OMP = os.path.join(ROOT, "oracle", "_ref", "libnli_ref_omp.so")


@contextlib.contextmanager
def c_stdout_to_stderr():
    """The reference's process() prints its progress with printf: send the C library's
    stdout to stderr, line-buffered, while it runs, so that stdout keeps the one JSON
    line."""
    libc = ctypes.CDLL(None)
    c_stdout = ctypes.c_void_p.in_dll(libc, "stdout")
    sys.stdout.flush()
    libc.fflush(None)
    saved = os.dup(1)
    os.dup2(2, 1)
    libc.setvbuf(c_stdout, None, 1, 0)   # _IOLBF
    try:
        yield
    finally:
        libc.fflush(None)
        os.dup2(saved, 1)
        os.close(saved)


def omp_threads(n):
    """Set the OpenMP thread count of this process (libgomp is shared by the oracle and
    the reference library)."""
    ctypes.CDLL("libgomp.so.1").omp_set_num_threads(int(n))


def cpu_params(O, threads, **kw):
    p = dict(method="nlpoisson", patch=9, sigma=10000.0, lam=0.05, iters=200, tol=0.1, scales=11, coarse=0.1,
             init="poisson", cg_tol=1e-4, cg_limit=100, seed=1, pm_threads=threads)
    p.update(kw)
    return O.InpaintParams(**p)


def cpu_baseline(wl, gpu_iterations, threads):
    """The reference's own CPU path (its sources built in place with OpenMP,
    oracle/_ref/libnli_ref_omp.so: ImageInpainting::process, OpenMP PatchMatch) on a
    bounded sample of the same workload: at most two inpainting iterations per scale.
    The oracle runs the same sample with per-scale timing; the full-run estimate is the
    oracle's per-scale iteration time x the GPU run's iterations per scale, scaled by the
    reference/oracle time ratio on the sample.  Without the reference build, the oracle's
    own estimate is reported (kind "port")."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    omp_threads(threads)
    orc = O.Oracle()
    t0 = time.perf_counter()
    _, st = orc.process(wl.image, wl.mask, cpu_params(O, threads, max_iters_per_scale=2))
    port_s = time.perf_counter() - t0
    est = st.seconds_pyramid
    for L in range(11):
        est += st.seconds_propagate[L]
        if st.iterations[L] > 0:
            est += st.seconds_iterations[L] / st.iterations[L] * gpu_iterations[L]
    H, W = wl.mask.shape
    model, nproc = host_cpu()
    port = dict(value=W * H / 1e6 / est, sample_seconds=port_s, estimated_full_seconds=est)
    if not os.path.exists(REF_OMP):
        return dict(value=port["value"], unit="MPix/s", cores=threads, kind="port", cpu_model=model,
                    host_logical_cpus=nproc,
                    sample=f"oracle process() of the same 1024x1024 workload with 2 inpainting iterations per "
                           f"scale ({port_s:.1f} s, PatchMatch on {threads} OpenMP threads); estimated full run "
                           f"{est:.0f} s = per-scale iteration time x the GPU run's iterations per scale")
    ref = O.Reference(REF_OMP)
    with c_stdout_to_stderr():
        t0 = time.perf_counter()
        ref.process(wl.image, wl.mask, cpu_params(O, threads, iters=2))
        ref_s = time.perf_counter() - t0
    est_ref = est * ref_s / port_s
    return dict(value=W * H / 1e6 / est_ref, unit="MPix/s", cores=threads, kind="reference", cpu_model=model,
                host_logical_cpus=nproc, port=port,
                sample=f"the reference's ImageInpainting::process (its sources built with OpenMP) on the same "
                       f"1024x1024 workload with at most 2 inpainting iterations per scale ({ref_s:.1f} s on "
                       f"{threads} OpenMP threads; the oracle took {port_s:.1f} s on the same sample); estimated "
                       f"full run {est_ref:.0f} s = the oracle's per-scale iteration time x the GPU run's "
                       f"iterations per scale x {ref_s / port_s:.3f}")


def cpu_full(wls, gpu_outs, threads, n, start=0):
    """Full reference runs (ImageInpainting::process to its stopping rule, OpenMP
    PatchMatch) of n images from index start with their workload's parameters, timed, and their
    hole PSNR (colour channels) and, for RGB-D, hole depth RMSE beside the GPU's outputs
    of the same images."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    omp_threads(threads)
    ref = O.Reference(REF_OMP)
    runs = []
    for k in range(start, start + n):
        wl = wls[k]
        hole = wl.mask.astype(bool)
        p = dict(wl.params, seed=1 + k, pm_threads=threads)
        with c_stdout_to_stderr():
            t0 = time.perf_counter()
            out = ref.process(wl.image, wl.mask, O.InpaintParams(**p))
            sec = time.perf_counter() - t0
        g = gpu_outs[k]
        run = dict(image=k, seconds=sec, mpix_s=wl.mask.size / 1e6 / sec,
                   psnr_hole_db_reference=psnr(out[hole][:, :3], wl.truth[hole][:, :3]),
                   psnr_hole_db_gpu=psnr(g[hole][:, :3], wl.truth[hole][:, :3]))
        run["gpu_minus_reference_db"] = run["psnr_hole_db_gpu"] - run["psnr_hole_db_reference"]
        if wl.image.shape[2] == 4:
            rm = lambda o: float(np.sqrt(np.mean((o[hole][:, 3].astype(np.float64) - wl.truth[hole][:, 3]) ** 2)))
            run["depth_rmse_m_reference"], run["depth_rmse_m_gpu"] = rm(out), rm(g)
        runs.append(run)
        log(f"full reference run {k}: {runs[-1]}")
    d = np.array([r["gpu_minus_reference_db"] for r in runs])
    return dict(threads=threads, runs=runs, mean_gpu_minus_reference_db=float(d.mean()),
                standard_error_db=float(d.std(ddof=1) / np.sqrt(len(d))) if len(d) > 1 else None)


def bench_rgbd(args, nli, syn):
    """RGB-D 1920x1080 nldepth images (distinct seeds), one parameter record, inputs
    resident on the device; --images of them per step (default 8)."""
    wls = [syn.config_rgbd(seed=2 + k) for k in range(args.images)]
    ctx = nli.Context(0)
    q = dict(wls[0].params)
    q["cg_precision"] = nli.CG_MIXED if args.precision == "mixed" else nli.CG_FP64
    q["cg_refine_every"] = 8 if args.precision == "mixed" else 0
    params = nli.Inpainting(**q)
    H, W, C = wls[0].image.shape
    d_in = [ctx.to_device(w.image) for w in wls]
    d_mask = [ctx.to_device(w.mask) for w in wls]
    d_out = [ctx.dev_alloc(w.image.nbytes) for w in wls]
    for i in range(args.warmup):
        t1 = time.perf_counter()
        ctx.inpaint_batch_dev(d_in, d_mask, d_out, [W] * len(wls), [H] * len(wls), C, params, args.workers)
        log(f"warmup {i}: {time.perf_counter() - t1:.2f} s")
    ctx.profile(["pm_sweep"])
    ctx.lib.nli_synchronize(ctx.h)
    t0 = time.perf_counter()
    for i in range(args.steps):
        t1 = time.perf_counter()
        ctx.inpaint_batch_dev(d_in, d_mask, d_out, [W] * len(wls), [H] * len(wls), C, params, args.workers)
        log(f"step {i}: {time.perf_counter() - t1:.2f} s for {len(wls)} images")
    ctx.lib.nli_synchronize(ctx.h)
    elapsed = time.perf_counter() - t0
    sweep = ctx.profile_read("pm_sweep")
    ctx.profile(False)
    outs = [ctx.from_device(d, w.image.shape, np.float32) for d, w in zip(d_out, wls)]
    rmse = [float(np.sqrt(np.mean((o[w.mask.astype(bool)][:, 3] - w.truth[w.mask.astype(bool)][:, 3]) ** 2)))
            for o, w in zip(outs, wls)]
    ms_per_step = elapsed / args.steps * 1e3
    achieved = sweep["work"] / sweep["ms"] / 1e9
    line = {
        "metric": "inpainted MPix/s, RGB-D 1920x1080 nldepth batch",
        "value": len(wls) * W * H / 1e6 / (ms_per_step / 1e3), "unit": "MPix/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": wls[0].name, "images": len(wls), "width": W, "height": H, "channels": C,
                   "hole_fraction": float(np.mean([w.mask.mean() for w in wls])), "cg_precision": args.precision,
                   "workers": args.workers},
        "roofline": dict(kernel="k_pm_sweep (one PatchMatch sweep)", bound="valu", achieved=achieved,
                         peak=FP32_PEAK_TFLOPS, unit="TFLOP/s", frac=achieved / FP32_PEAK_TFLOPS, traffic=None,
                         launches=sweep["launches"], avg_launch_ms=sweep["ms"] / max(sweep["launches"], 1)),
        "cpu_baseline": None,
        "depth_rmse_m": float(np.mean(rmse)),
        "cpu_full_runs": None,
    }
    if args.cpu_full > 0:
        log(f"full reference runs of {args.cpu_full} images on {args.cpu_threads} threads ...")
        line["cpu_full_runs"] = cpu_full(wls, outs, args.cpu_threads, args.cpu_full, args.cpu_full_start)
    print(json.dumps(line), flush=True)
    for p in d_in + d_mask + d_out:
        ctx.dev_free(p)
    ctx.close()


def bench_uneven(args, nli, syn):
    """The uneven batch: 64 images, per-image parameters, inputs resident on the device."""
    t0 = time.perf_counter()
    wls = syn.config_uneven(64)
    log(f"generated {len(wls)} images in {time.perf_counter() - t0:.1f} s")
    ctx = nli.Context(0)
    params = [nli.Inpainting(**w.params) for w in wls]
    d_in = [ctx.to_device(w.image) for w in wls]
    d_mask = [ctx.to_device(w.mask) for w in wls]
    d_out = [ctx.dev_alloc(w.image.nbytes) for w in wls]
    Ws = [w.image.shape[1] for w in wls]
    Hs = [w.image.shape[0] for w in wls]
    mpix = sum(W * H for W, H in zip(Ws, Hs)) / 1e6
    for i in range(args.warmup):
        t1 = time.perf_counter()
        ctx.inpaint_batch_dev(d_in, d_mask, d_out, Ws, Hs, 3, params, args.workers)
        log(f"warmup {i}: {time.perf_counter() - t1:.2f} s")
    ctx.profile(["pm_sweep"])
    ctx.lib.nli_synchronize(ctx.h)
    t0 = time.perf_counter()
    for i in range(args.steps):
        t1 = time.perf_counter()
        ctx.inpaint_batch_dev(d_in, d_mask, d_out, Ws, Hs, 3, params, args.workers)
        log(f"step {i}: {time.perf_counter() - t1:.2f} s for {len(wls)} images ({mpix:.1f} MPix)")
    ctx.lib.nli_synchronize(ctx.h)
    elapsed = time.perf_counter() - t0
    sweep = ctx.profile_read("pm_sweep")
    ctx.profile(False)
    outs = [ctx.from_device(d, w.image.shape, np.float32) for d, w in zip(d_out, wls)]
    psnrs = [psnr(o[w.mask.astype(bool)], w.truth[w.mask.astype(bool)]) for o, w in zip(outs, wls)]
    ms_per_step = elapsed / args.steps * 1e3
    achieved = sweep["work"] / sweep["ms"] / 1e9
    line = {
        "metric": "inpainted MPix/s, uneven batch of 64 (nlmeans/nlmedians, patch 7/9, 8 scales)",
        "value": mpix / (ms_per_step / 1e3), "unit": "MPix/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "uneven64", "images": len(wls), "megapixels": mpix, "workers": args.workers},
        "roofline": dict(kernel="k_pm_sweep (one PatchMatch sweep)", bound="valu", achieved=achieved,
                         peak=FP32_PEAK_TFLOPS, unit="TFLOP/s", frac=achieved / FP32_PEAK_TFLOPS, traffic=None,
                         launches=sweep["launches"], avg_launch_ms=sweep["ms"] / max(sweep["launches"], 1)),
        "cpu_baseline": None,
        "psnr_hole_db": float(np.mean(psnrs)),
    }
    print(json.dumps(line), flush=True)
    for p in d_in + d_mask + d_out:
        ctx.dev_free(p)
    ctx.close()


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--images", type=int, default=8, help="1024x1024 images per step")
    ap.add_argument("--workers", type=int, default=8,
                    help="lockstep groups (or, with --batch-mode concurrent, per-image drivers), one stream each")
    ap.add_argument("--batch-mode", choices=["lockstep", "concurrent"], default="concurrent",
                    help="lockstep: each group steps its images' scales and iterations together, one launch per "
                         "kernel for all of them, stopping rule on the device; concurrent: one driver per image")
    ap.add_argument("--precision", choices=["mixed", "fp64"], default="mixed")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-compare", action="store_true", help="skip the single-image and FP64 runs")
    ap.add_argument("--no-breakdown", action="store_true", help="skip the all-kernel timing batch")
    ap.add_argument("--no-colour", action="store_true", help="skip the sRGB <-> Lab conversion timing")
    ap.add_argument("--sweep-sample", type=int, default=16, metavar="K",
                    help="time one PatchMatch call in K with HIP events (its sweeps as one span; "
                         "roofline); 1 = every call")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="first run 3 device-to-device copies of MIB MiB (the known-size copy that "
                         "calibrates hardware byte counters)")
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--cpu-full", type=int, default=0, metavar="N",
                    help="also run the reference's CPU path in full on the first N images (minutes each) and "
                         "compare hole PSNR with the GPU's outputs")
    ap.add_argument("--cpu-full-start", type=int, default=0, metavar="S",
                    help="with --cpu-full N: run images S .. S+N-1 instead (--images must cover them), so "
                         "several calls can cover different images")
    ap.add_argument("--fixed-iters", type=int, default=10, metavar="K",
                    help="also time the batch, one image and the plain drop-in with exactly K inpainting "
                         "iterations per scale (stopping tolerance 0), for like-for-like comparisons; 0 = skip")
    ap.add_argument("--workload", choices=["rgb1024", "uneven64", "rgbd1080"], default="rgb1024",
                    help="rgb1024: the headline batch of 1024x1024 nlpoisson images; uneven64: 64 images of "
                         "sides 128-1024, nlmeans/nlmedians, patch 7/9, 8 scales (per-image parameters); "
                         "rgbd1080: 1920x1080 RGB-D images, nldepth (lambda_rgb 1, lambda_d 0), patch 9, 11 scales")
    args = ap.parse_args()
    if args.cpu_full > 0 and args.cpu_full_start + args.cpu_full > args.images:
        print(f"--cpu-full-start {args.cpu_full_start} + --cpu-full {args.cpu_full} exceeds --images "
              f"{args.images}", file=sys.stderr)
        sys.exit(2)
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (images are independent: "
              f"one process per GPU would take its share of them)", file=sys.stderr)
        sys.exit(2)

    nli = importlib.import_module("non-local-inpainting_amd")
    syn = importlib.import_module("non-local-inpainting_amd.synthetic")
    nli.batch_mode(args.batch_mode)
    if args.workload == "uneven64":
        return bench_uneven(args, nli, syn)
    if args.workload == "rgbd1080":
        return bench_rgbd(args, nli, syn)
    wls = [syn.config_1024(seed=1 + k) for k in range(args.images)]
    H, W, C = wls[0].image.shape
    precision = nli.CG_MIXED if args.precision == "mixed" else nli.CG_FP64
    params = workload_params(nli, precision)
    ctx = nli.Context(0)
    d_in = [ctx.to_device(w.image) for w in wls]
    d_mask = [ctx.to_device(w.mask) for w in wls]
    d_out = [ctx.dev_alloc(w.image.nbytes) for w in wls]
    Ws, Hs = [W] * len(wls), [H] * len(wls)

    if args.copy_probe > 0:
        nbytes = args.copy_probe << 20
        a, b = ctx.dev_alloc(nbytes), ctx.dev_alloc(nbytes)
        for _ in range(3):
            ctx.copy_d2d(b, a, nbytes)
        ctx.dev_free(a)
        ctx.dev_free(b)
        log(f"copy probe: 3 device copies of {nbytes} bytes")

    for i in range(args.warmup):
        t0 = time.perf_counter()
        ctx.inpaint_batch_dev(d_in, d_mask, d_out, Ws, Hs, C, params, args.workers)
        log(f"warmup {i}: {time.perf_counter() - t0:.2f} s for {len(wls)} images")

    # timed steps: only the dominant kernel (the PatchMatch sweep) carries timing events:
    # the sweeps of one PatchMatch call in --sweep-sample, as one span (an event pair is a
    # marker in the stream)
    ctx.profile(["pm_sweep"], period=args.sweep_sample)
    ctx.lib.nli_synchronize(ctx.h)
    t0 = time.perf_counter()
    for i in range(args.steps):
        t1 = time.perf_counter()
        stats = ctx.inpaint_batch_dev(d_in, d_mask, d_out, Ws, Hs, C, params, args.workers)
        log(f"step {i}: {time.perf_counter() - t1:.2f} s for {len(wls)} images; iterations per scale of image 0 "
            f"{stats[0].iterations}")
    ctx.lib.nli_synchronize(ctx.h)
    elapsed = time.perf_counter() - t0
    sweep = ctx.profile_read("pm_sweep")
    ctx.profile(False)
    outs = [ctx.from_device(d, w.image.shape, np.float32) for d, w in zip(d_out, wls)]
    psnrs = [psnr(o[w.mask.astype(bool)], w.truth[w.mask.astype(bool)]) for o, w in zip(outs, wls)]
    ms_per_step = elapsed / args.steps * 1e3
    value = len(wls) * W * H / 1e6 / (ms_per_step / 1e3)

    # roofline of the dominant kernel.  In the timed steps the event pair of a sampled call
    # brackets its sweeps on their own stream, but 7 other drivers' kernels share the CUs
    # meanwhile, so the span is not the kernel's duration.  The kernel's own duration comes
    # from one image on one driver, every call's sweeps timed as one span (the
    # configuration a kernel trace of one driver reports; the span holds the back-to-back
    # sweeps and their dispatch, not a marker per sweep); the timed region's figure is kept
    # beside it.
    ctx.profile(["pm_sweep"], period=1)
    ctx.inpaint_dev(d_in[0], d_mask[0], d_out[0], W, H, C, params)
    sw1 = ctx.profile_read("pm_sweep")
    ctx.profile(False)
    achieved = sw1["work"] / sw1["ms"] / 1e9            # flops per ms -> TFLOP/s
    shared = sweep["work"] / sweep["ms"] / 1e9
    roof = dict(kernel="k_pm_sweep (one PatchMatch sweep)", bound="valu", achieved=achieved,
                peak=FP32_PEAK_TFLOPS, unit="TFLOP/s", frac=achieved / FP32_PEAK_TFLOPS, traffic=measured_traffic(),
                traffic_source="profiles/k_pm_sweep_rgb1024_nlpoisson_traffic.md (bytes per sweep)",
                launches=sw1["launches"], avg_launch_ms=sw1["ms"] / max(sw1["launches"], 1),
                run="image 0 of the batch on one driver after the timed steps, same parameters, the 10 "
                    "sweeps of every PatchMatch call event-timed as one span on their stream",
                timed_region=dict(achieved=shared, frac=shared / FP32_PEAK_TFLOPS, launches=sweep["launches"],
                                  avg_launch_ms=sweep["ms"] / max(sweep["launches"], 1),
                                  sampled=f"one PatchMatch call in {args.sweep_sample} (in issue order over "
                                          f"all drivers) carries an event pair around its sweeps; spans "
                                          f"include the other drivers' concurrent kernels"),
                note="bound by FP32/FP64 vector (VALU) work, no MFMA: patch distances are not a contraction; "
                     "peak = the FP32 vector peak; algorithmic flops T*(2+R)*P^2*(9C+4) per sweep")

    # per-kernel breakdown: one more batch with every kernel timed (not part of the timed steps;
    # concurrent streams make each kernel's event time include the others' work)
    kern = None
    if not args.no_breakdown:
        ctx.profile(True)
        ctx.inpaint_batch_dev(d_in, d_mask, d_out, Ws, Hs, C, params, args.workers)
        kern = {}
        for k in nli.KERNELS:
            r = ctx.profile_read(k)
            # with several drivers the event spans of different streams overlap: their sum
            # can exceed the step's wall time and is not kernel time
            kern[k] = dict(launches=r["launches"], work=r["work"], overlapping_span_ms=r["ms"])
        kern["note"] = ("per-kernel event spans summed over 8 concurrent streams (they overlap, so they are not "
                        "kernel time); the rocprofv3 traces in profiles/ give kernel times")
        ctx.profile(False)

    compare = None
    if not args.no_compare:
        # the same image alone on one stream (device-resident), and in FP64
        t1 = time.perf_counter()
        st1 = ctx.inpaint_dev(d_in[0], d_mask[0], d_out[0], W, H, C, params)
        single = W * H / 1e6 / (time.perf_counter() - t1)
        out1 = ctx.from_device(d_out[0], wls[0].image.shape, np.float32)
        t1 = time.perf_counter()
        ctx.inpaint_dev(d_in[0], d_mask[0], d_out[0], W, H, C, workload_params(nli, nli.CG_FP64))
        fp64 = W * H / 1e6 / (time.perf_counter() - t1)
        out64 = ctx.from_device(d_out[0], wls[0].image.shape, np.float32)
        # the plain drop-in: the C++ plugin classes in the reference's composition, one
        # C-ABI call per reference call on host buffers
        t1 = time.perf_counter()
        out_plain, _ = ctx.inpaint_plain(wls[0].image, wls[0].mask, params)
        plain = W * H / 1e6 / (time.perf_counter() - t1)
        hole0 = wls[0].mask.astype(bool)
        compare = dict(plain_dropin_mpix_s=plain, batch_speedup_vs_plain=value / plain,
                       plain_equals_device_driver=bool(np.array_equal(out_plain, out1)),
                       single_image_mpix_s=single, single_image_fp64_mpix_s=fp64,
                       batch_speedup_vs_single=value / single,
                       psnr_mixed_minus_fp64_db=psnr(out1[hole0], wls[0].truth[hole0]) -
                       psnr(out64[hole0], wls[0].truth[hole0]),
                       single_seconds=dict(patchmatch=st1.seconds_patchmatch, update=st1.seconds_update,
                                           total=st1.seconds_total),
                       iterations_per_scale=st1.iterations, cg_iterations_total=st1.cg_iterations_total,
                       scale_sizes=st1.sizes, hole_pixels_per_scale=st1.hole_pixels,
                       targets_per_scale=st1.target_pixels,
                       random_search_windows_per_scale=[int(max(w, h)).bit_length() for w, h in st1.sizes])
        log(f"single image: {compare}")
        if args.fixed_iters > 0:
            # fixed work: every scale runs exactly K iterations (the stopping tolerance is 0)
            pf = workload_params(nli, precision, iters=args.fixed_iters, tol=0.0)
            t1 = time.perf_counter()
            stf = ctx.inpaint_batch_dev(d_in, d_mask, d_out, Ws, Hs, C, pf, args.workers)
            fb = len(wls) * W * H / 1e6 / (time.perf_counter() - t1)
            t1 = time.perf_counter()
            ctx.inpaint_dev(d_in[0], d_mask[0], d_out[0], W, H, C, pf)
            fs = W * H / 1e6 / (time.perf_counter() - t1)
            t1 = time.perf_counter()
            ctx.inpaint_plain(wls[0].image, wls[0].mask, pf)
            fp = W * H / 1e6 / (time.perf_counter() - t1)
            compare["fixed_iterations"] = dict(iterations_per_scale=args.fixed_iters, batched_mpix_s=fb,
                                               single_image_mpix_s=fs, plain_dropin_mpix_s=fp,
                                               batch_speedup_vs_plain=fb / fp,
                                               iterations_run=stf[0].iterations)
            log(f"fixed iterations: {compare['fixed_iterations']}")

    colour = None
    if not args.no_colour:
        # the colour path of the reference's command-line program (0-255 RGB <-> L*a*b*
        # around process(), main.cpp:117, 307): both conversions over the batch's images
        # (scaled to 0-255), device-resident, event-timed; not part of the timed steps
        d_rgb = [ctx.to_device(np.rint(w.image * 255.0).astype(np.float32)) for w in wls]   # 8-bit samples
        d_lab = ctx.dev_alloc(W * H * C * 4)
        d_back = ctx.dev_alloc(W * H * C * 4)
        for d in d_rgb[:1]:
            ctx.rgb_to_lab_dev(d, d_lab, W, H)
            ctx.lab_to_rgb_dev(d_lab, d_back, W, H)
        res = {}
        for name, fn, src, dst in (("rgb_to_lab", ctx.rgb_to_lab_dev, None, d_lab),
                                   ("lab_to_rgb", ctx.lab_to_rgb_dev, d_lab, d_back)):
            ctx.profile(["color"])
            for rep in range(4):
                for d in d_rgb:
                    fn(src if src is not None else d, dst, W, H)
            k = ctx.profile_read("color")
            ctx.profile(False)
            us = k["ms"] / max(k["launches"], 1) * 1e3
            gbs = k["work"] / k["ms"] / 1e6
            res[name] = dict(launches=k["launches"], avg_launch_us=us, bytes_per_launch=k["work"] / k["launches"],
                             achieved_gb_s=gbs, frac_of_hbm_peak=gbs / HBM_PEAK_GBS)
        for p in d_rgb + [d_lab, d_back]:
            ctx.dev_free(p)
        colour = dict(image=f"{W}x{H}x3 float32 holding 8-bit sample values 0-255", kernels=res,
                      note="24 algorithmic bytes per pixel (3 floats in, 3 out); each of the reference's powf "
                           "calls (3 per pixel per direction) is evaluated to the correctly rounded float in FP64 "
                           "(Newton-refined roots), which bounds the kernels rather than HBM")
        log(f"colour path: {colour}")

    cpu = None
    if not args.no_cpu_baseline:
        log(f"CPU baseline sample on {args.cpu_threads} threads ...")
        cpu = cpu_baseline(wls[0], stats[0].iterations, args.cpu_threads)
        log(f"CPU baseline: {cpu}")

    full = None
    if args.cpu_full > 0:
        log(f"full reference runs of {args.cpu_full} images on {args.cpu_threads} threads ...")
        full = cpu_full(wls, outs, args.cpu_threads, args.cpu_full, args.cpu_full_start)

    line = {
        "metric": METRIC, "value": value, "unit": "MPix/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": wls[0].name, "images": len(wls), "width": W, "height": H, "channels": C,
                   "hole_fraction": float(np.mean([w.mask.mean() for w in wls])), "method": "nlpoisson",
                   "lambda": 0.05, "patch": 9, "scales": 11, "max_iters": 200, "cg": "tol 1e-4, <= 100 iterations",
                   "cg_precision": args.precision, "workers": args.workers},
        "roofline": roof,
        "cpu_baseline": cpu,
        "cpu_full_runs": full,
        "psnr_hole_db": float(np.mean(psnrs)),
        "kernels": kern,
        "compare": compare,
        "colour_path": colour,
    }
    print(json.dumps(line), flush=True)
    for p in d_in + d_mask + d_out:
        ctx.dev_free(p)
    ctx.close()


if __name__ == "__main__":
    main()
