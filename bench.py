"""Benchmark of the direct (point-mode) SFB decomposition on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
                    [--workload clustered1m|lognormal2p25|uniform10k|batched64|healpix4m|shotnoise2p25|ingest2p25]
                    [--precision mixed|fp64] [--cpu-sample N | --no-cpu-baseline]

The default workload is lognormal2p25: 2^25 points of a lognormal density field
cut to a ball of rmax 1000, lmax 256, k_max 0.5 (n_l 159 -> 52), decomposed in
mixed precision and compared with the plain FP64 GPU path on the same points.
A step is one full sfbhip_direct call (k_ln lookup, point preparation, the
contraction kernels, the FP64 reduction and C_l assembly) on points already
resident in HBM.  The metric is point-mode pairs/s = N * sum_l (2l+1) n_l /
step time.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "SFB point-mode pairs/s (N·Σ_l(2l+1)n_l); mixed vs FP64 GPU at ≤1e-6 rel C_l"

# peaks (MI355X datasheet, dense)
PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "fp16": 2516.6}
PEAK_HBM_GBS = 8000.0

WORKLOADS = {
    # 1,000,000 clustered points (20,000 parents x Poisson(50), sigma 5, clipped), lmax 64, k_max 0.2, rmax 1000
    "clustered1m": dict(lmax=64, nmax=64, rmax=1000.0, k_max=0.2),
    # 2^25 lognormal points, lmax 256, k_max 0.5, rmax 1000
    "lognormal2p25": dict(lmax=256, nmax=160, rmax=1000.0, k_max=0.5),
    # 10,000 uniform points, lmax 16, k_max 0.1, rmax 500
    "uniform10k": dict(lmax=16, nmax=16, rmax=500.0, k_max=0.1),
    # 64 catalogues: N log-uniform in [1e4, 2e6], rmax U[300, 1500], lmax U{32..128}, k_max 0.15,
    # each uniform in its ball, decomposed in one batched call
    "batched64": dict(lmax=128, nmax=0, rmax=0.0, k_max=0.15, batched=True),
    # HEALPix path: 4,000,000 points, p(r) ~ r^2 exp(-(r/400)^2) on [0, 1000], nside 256, lmax 128,
    # k_max 0.15 (FP64; the reference's public wrapper semantics: linear Bessel table for N > 1000)
    "healpix4m": dict(lmax=128, nmax=48, rmax=1000.0, k_max=0.15, nside=256, healpix=True),
    # the shot-noise step after the path: (V/N)(2/pi) k^2 int r^2 w(r) j_l(k r)^2 dr for every kept
    # mode of the 2^25-point configuration (lmax 256, k_max 0.5, rmax 1000: 26,102 integrals), w = tophat
    "shotnoise2p25": dict(lmax=256, nmax=160, rmax=1000.0, k_max=0.5, shotnoise=True),
    # the ingestion step before the path: 2^25 float32 cartesian records (a Gadget-like box of side
    # 2000 centred by shift -1000) through tophat(1000) + angular mask (nside 64) + Gaussian(400)
    "ingest2p25": dict(n=2 ** 25, rmax=1000.0, nside=64, r0=400.0, ingest=True),
}


def make_batch(seed=0, ncat=64):
    """The 64-catalogue batch: per catalogue (points, lmax, nmax, rmax, k_max)."""
    from catana_amd import workloads as W
    rng = np.random.default_rng(seed)
    cats = []
    for i in range(ncat):
        n = int(round(10 ** rng.uniform(4.0, np.log10(2e6))))
        rmax = float(rng.uniform(300.0, 1500.0))
        lmax = int(rng.integers(32, 129))
        nmax = int(np.ceil(0.15 * rmax / np.pi)) + 1  # > the number of zeros below k_max rmax for every l
        cats.append(dict(np=W.uniform_ball(n, rmax, seed=1000 + i), lmax=lmax, nmax=nmax, rmax=rmax, k_max=0.15))
    return cats


# HBM bytes per call of the dominant kernels (all launches of one call), measured with
# rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE and scaled by what the counters read for a 512 MiB
# device-to-device copy (FETCH_SIZE x 2.0, WRITE_SIZE x 1.0).  Counters cannot be read from
# inside the timed run, so the figure is carried here per workload: (workload, precision,
# operands, fp64_below, chunks) for the point path, (workload, precision) for the others.
PROFILED_TRAFFIC_BYTES = {("clustered1m", "mixed", "split16", 2, 4): 7.59e8,   # split16 classes
                          ("lognormal2p25", "mixed", "split16", 2, 4): 9.84e11,  # split16 classes
                          ("healpix4m", "mixed"): 2.000e10,                    # maps + ring contraction
                          ("ingest2p25",): 1.337e9}                            #  fp64_below, chunks)


def make_points(name):
    from catana_amd import workloads as W
    if name == "clustered1m":
        return W.clustered_ball(20000, 1000.0, mean_children=50, sigma=5.0, seed=2)
    if name == "lognormal2p25":
        return W.lognormal_ball(2 ** 25, rmax=1000.0, ngrid=256, k0=0.1, seed=5)
    if name == "uniform10k":
        return W.uniform_ball(10000, 500.0, seed=1)
    raise ValueError(name)


# default CPU-baseline subsample per workload (points): each timed run is a few seconds on 16 threads
CPU_SAMPLE = {"clustered1m": 1 << 17, "lognormal2p25": 4096, "uniform10k": 10000, "batched64": 1 << 15}
CPU_RUNS = 3
KERNELS = {"auto": 0, "256": 1, "512": 2, "grouped": 3}


def contraction_kernel(prec_fp64, operands, kernel):
    """Name of the kernel that runs the point-mode contraction for these options."""
    if prec_fp64:
        return "sfb_direct_f64_kernel"
    if operands == "split16" and kernel in ("auto", "grouped"):
        return "sfb_direct_split16_kernel"
    if operands == "split16" and kernel == "512":
        return "sfb_direct_mixed2_kernel"
    return "sfb_direct_mixed_kernel"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or "unknown"


def cpu_baseline(points, cfg, nl, sample, threads, runs=CPU_RUNS):
    """The oracle (C/OpenMP restatement of sfb_raw.cpp:13-108 with the public
    wrapper's interpolated = N > 1000, sfb_decomposition.cpp:17) timed on a
    seeded subsample of the same catalogue, median of `runs` runs.  The
    per-call interpolation-table build (1000 nmax j_l values per l,
    SBesselInterpolator.cpp:10-17) does not depend on N: it is timed on its own
    (a 1001-point call) and the full-catalogue time is estimated as
    table + (N / n) (median - table)."""
    from oracle import oracle as O
    rng = np.random.default_rng(123)
    N = points.shape[0]
    n = min(sample, N)
    sub = points[np.sort(rng.choice(N, n, replace=False))]
    times = []
    for _ in range(runs):
        t0 = time.perf_counter()
        O.sfb_raw(sub, cfg["lmax"], cfg["nmax"], cfg["rmax"], nl=nl, interpolated=n > 1000, nthreads=threads)
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    t_tab = 0.0
    if n > 1000:
        t0 = time.perf_counter()
        O.sfb_raw(sub[:1001], cfg["lmax"], cfg["nmax"], cfg["rmax"], nl=nl, interpolated=True, nthreads=threads)
        t_tab = min(time.perf_counter() - t0, med)
    t_full = t_tab + (N / n) * (med - t_tab)
    l = np.arange(cfg["lmax"])
    pairs = N * int(((2 * l + 1) * nl).sum())
    return {"value": pairs / t_full, "unit": "pairs/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(), "nproc": os.cpu_count(),
            "runs_s": [round(t, 3) for t in times], "table_build_s": round(t_tab, 3),
            "sample": f"{n} of {N} points (seeded subset), same lmax/k_max, reference-default linear Bessel "
                      f"table; median of {runs} runs {med:.2f} s (spread {min(times):.2f}-{max(times):.2f} s), "
                      f"table build {t_tab:.2f} s; full catalogue estimated as table + (N/n)(median - table) "
                      f"= {t_full:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="lognormal2p25", choices=sorted(WORKLOADS))
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--operands", default="split16", choices=["fp32", "split16", "split3"],
                    help="matrix-core operand format of the mixed path")
    ap.add_argument("--chunks", type=int, default=4, choices=[1, 2, 4, 8],
                    help="mixed path: m chunks accumulated per work item")
    ap.add_argument("--fp64-below", type=int, default=2, help="mixed path: l below this run in FP64")
    ap.add_argument("--kernel", default="auto", choices=sorted(KERNELS),
                    help="mixed path, split-FP16 operands: the contraction kernel")
    ap.add_argument("--batch-mode", default="graph", choices=["graph", "plan", "call"],
                    help="batched64: replay a captured batch plan (graph), enqueue the plan (plan), or one "
                         "sfbhip_direct_batched call per step (call)")
    ap.add_argument("--autotune", action="store_true", help="batched64: run the tile-table autotuner first")
    ap.add_argument("--cpu-sample", type=int, default=0,
                    help="points of the CPU baseline's seeded subsample (0: the workload's default, "
                         "a few seconds per run on 16 threads, three runs)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="run one device-to-device copy of MIB MiB before the workload (a kernel of known "
                         "traffic for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: runs on one GPU only (across GPUs the SFB path would shard the points "
              "and sum f_lmn with one all-reduce); --gpus must be 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from catana_amd import capi

    if args.calibrate_copy > 0:
        n = args.calibrate_copy * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy: {args.calibrate_copy} MiB read + {args.calibrate_copy} MiB written")
        del src, dst

    cfg = WORKLOADS[args.workload]
    if cfg.get("batched"):
        run_batched(args)
        return
    if cfg.get("healpix"):
        run_healpix(args, cfg)
        return
    if cfg.get("shotnoise"):
        run_shotnoise(args, cfg)
        return
    if cfg.get("ingest"):
        run_ingest(args, cfg)
        return
    log(f"generating workload {args.workload}")
    pts = make_points(args.workload)
    log(f"{pts.shape[0]} points")
    N = pts.shape[0]
    dev = torch.device("cuda:0")
    d_pts = torch.from_numpy(np.ascontiguousarray(pts)).to(dev)
    h = capi.Handle(0)
    prec = capi.PREC_MIXED if args.precision == "mixed" else capi.PREC_FP64
    ops = {"fp32": capi.OPERANDS_FP32, "split16": capi.OPERANDS_SPLIT_FP16, "split3": capi.OPERANDS_SPLIT3_FP16}
    kw = dict(k_max=cfg["k_max"], precision=prec, bessel=capi.BESSEL_EXACT, mixed_fp64_below=args.fp64_below,
              mixed_operands=ops[args.operands], mixed_chunks=args.chunks, mixed_kernel=KERNELS[args.kernel])
    p = capi.make_params(cfg["lmax"], cfg["nmax"], cfg["rmax"], **kw)
    _, nl = h.kln(p)
    l = np.arange(cfg["lmax"])
    pairs = N * int(((2 * l + 1) * nl).sum())
    flops = 4.0 * N * int(((l + 1) * nl).sum())
    # the dominant kernel: the mixed contraction (l >= fp64_below) or, in FP64, the FP64 contraction
    lo = min(args.fp64_below, cfg["lmax"]) if prec == capi.PREC_MIXED else 0
    flops_dom = 4.0 * N * int(((l[lo:] + 1) * nl[lo:]).sum())

    stream = torch.cuda.current_stream(dev)
    log("warmup")
    for _ in range(args.warmup):
        capi.sfb_direct_torch(h, d_pts, cfg["lmax"], cfg["nmax"], cfg["rmax"], **kw)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    kern_ms, phases = [], []
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        k, c, _ = capi.sfb_direct_torch(h, d_pts, cfg["lmax"], cfg["nmax"], cfg["rmax"], **kw)
        kern_ms.append(h.last_kernel_ms())
        phases.append(h.last_phase_ms())
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    log(f"timed {args.steps} steps: {wall:.3f} s")
    ms = e0.elapsed_time(e1) / args.steps
    kms = float(np.mean(kern_ms))
    ph = np.mean(np.array(phases), axis=0)

    if prec == capi.PREC_FP64:
        dtype, peak = "f64", PEAK_TFLOPS["fp64"]
    elif args.operands == "split16":
        dtype, peak = "f16 2-piece split, 3 products (f32 accumulate, f64 tile sums)", PEAK_TFLOPS["fp16"] / 3.0
    elif args.operands == "split3":
        dtype, peak = "f16 3-piece split, 6 products (f32 accumulate, f64 tile sums)", PEAK_TFLOPS["fp16"] / 6.0
    else:
        dtype, peak = "f32 (f32 operands and accumulate, f64 tile sums)", PEAK_TFLOPS["fp32"]
    achieved = flops_dom / (kms * 1e-3) / 1e12
    out = {
        "metric": METRIC, "value": pairs / (ms * 1e-3), "unit": "pairs/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": dtype,
        "data": "synthetic, seeded (catana_amd/workloads.py), inputs resident in HBM",
        "config": {"workload": args.workload, "N": int(N), "lmax": cfg["lmax"], "k_max": cfg["k_max"],
                   "rmax": cfg["rmax"], "precision": args.precision, "operands": args.operands,
                   "chunks": args.chunks, "kernel": args.kernel,
                   "fp64_below": args.fp64_below, "pairs_per_step": pairs,
                   "wall_s_timed_region": wall},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak,
                     "traffic": PROFILED_TRAFFIC_BYTES.get((args.workload, args.precision, args.operands, args.fp64_below,
                                                              args.chunks)),
                     "kernel": contraction_kernel(prec == capi.PREC_FP64, args.operands, args.kernel),
                     "kernel_ms": kms, "algorithmic_flops_per_launch": flops_dom,
                     "l_range": [lo, cfg["lmax"]],
                     "phases_ms": {"fp64_contraction": float(ph[0]), "mixed_contraction": float(ph[1]),
                                   "reduce_and_cln": float(ph[2])}},
    }
    if prec != capi.PREC_FP64 and not args.no_accuracy:
        log("FP64 reference run for the accuracy figures")
        _, cm, fm = capi.sfb_direct_torch(h, d_pts, cfg["lmax"], cfg["nmax"], cfg["rmax"], store_flmn=True, **kw)
        torch.cuda.synchronize()
        t64 = time.perf_counter()
        k64, c64, f64 = capi.sfb_direct_torch(h, d_pts, cfg["lmax"], cfg["nmax"], cfg["rmax"],
                                              k_max=cfg["k_max"], precision=capi.PREC_FP64,
                                              bessel=capi.BESSEL_EXACT, store_flmn=True)
        torch.cuda.synchronize()
        t64 = time.perf_counter() - t64
        # the plain FP64 GPU path on the same points, one call (it includes the
        # f_lmn store, which the timed steps skip; its cost is negligible)
        out["fp64_gpu"] = {"ms_per_step": t64 * 1e3, "value": pairs / t64, "unit": "pairs/s",
                           "speedup_of_value": out["value"] / (pairs / t64)}
        a, b = cm.cpu().numpy(), c64.cpu().numpy()
        assert np.array_equal(a, c.cpu().numpy()), "store_flmn changed c_ln"
        mask = b != 0
        rel = np.abs(a[mask] / b[mask] - 1)
        worst = np.unravel_index(np.argmax(np.where(mask, np.abs(a / np.where(mask, b, 1) - 1), 0)), a.shape)
        fa = capi.split_flmn(fm.cpu().numpy(), cfg["lmax"], cfg["nmax"])
        fb = capi.split_flmn(f64.cpu().numpy(), cfg["lmax"], cfg["nmax"])
        ferr = [float(np.linalg.norm(x - y) / np.linalg.norm(y)) for x, y in zip(fa, fb) if np.linalg.norm(y) > 0]
        out["accuracy"] = {"max_rel_cln_vs_fp64": float(rel.max()), "median_rel_cln_vs_fp64": float(np.median(rel)),
                           "worst_cln_l_n": [int(worst[0]), int(worst[1])],
                           "max_l_rel_norm_flmn_vs_fp64": max(ferr)}
    if not args.no_cpu_baseline:
        log("CPU baseline")
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        out["cpu_baseline"] = cpu_baseline(pts, cfg, nl, args.cpu_sample or CPU_SAMPLE.get(args.workload, 65536),
                                           threads)
    h.close()
    print(json.dumps(out))


def run_healpix(args, cfg):
    """The HEALPix workload: a step is one binning (sfbhip_pixelize) plus one
    pixelised decomposition (sfbhip_healpix) on points resident in HBM.  With
    --precision mixed (default) the ring path: FP32 j_l-weighted maps, ring
    Legendre table, split-FP16 ring contraction; HBM roofline on the algorithmic
    bytes sum_l [4 N + 8 npix n_l] (radii read once per l, each FP32 map written
    and read once).  With --precision fp64 the FP64 pixel contraction with the
    reference's linear j_l table (the drop-in)."""
    import torch
    from catana_amd import capi, workloads as W
    log("generating 4,000,000 points with radial selection")
    pts = W.radial_selection(4_000_000, cfg["rmax"], r0=400.0, seed=3)
    dev = torch.device("cuda:0")
    d_pts = torch.from_numpy(np.ascontiguousarray(pts)).to(dev)
    h = capi.Handle(0)
    L, NM, R, ns = cfg["lmax"], cfg["nmax"], cfg["rmax"], cfg["nside"]
    _, nl = h.kln(capi.make_params(L, NM, R, k_max=cfg["k_max"]))
    l = np.arange(L)
    N = pts.shape[0]
    pairs = N * int(((2 * l + 1) * nl).sum())
    npix = 12 * ns * ns
    flops = 4.0 * npix * int(((l + 1) * nl).sum())
    hbm_bytes = float(sum(4.0 * N + 8.0 * npix * int(nl[x]) for x in range(L)))
    ring = args.precision == "mixed"
    prec = capi.PREC_MIXED if ring else capi.PREC_FP64

    def step(**kw):
        off, rad = capi.pixelize_torch(h, d_pts, ns)
        return capi.sfb_healpix_torch(h, off, rad, ns, L, NM, R, k_max=cfg["k_max"], **kw)

    log("warmup")
    for _ in range(args.warmup):
        step(precision=prec)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream(dev)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    kern, phases = [], []
    e0.record(stream)
    for _ in range(args.steps):
        k, c, _ = step(precision=prec)
        kern.append(h.last_kernel_ms())
        phases.append(h.last_phase_ms())
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    kms = float(np.mean(kern))
    ph = np.mean(np.array(phases), axis=0)
    log(f"timed {args.steps} steps: {ms:.2f} ms per step")
    if ring:
        achieved = hbm_bytes / (kms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": achieved / PEAK_HBM_GBS,
                "traffic": PROFILED_TRAFFIC_BYTES.get(("healpix4m", "mixed")),
                "kernel": "ring_maps_kernel + healpix_ring_contract_kernel", "kernel_ms": kms,
                "algorithmic_bytes_per_launch": hbm_bytes,
                "phases_ms": {"maps": float(ph[0]), "ring_contraction": float(ph[1]),
                              "reduce_and_cln": float(ph[2])},
                "matrix_core_tflops": flops / (float(ph[1]) * 1e-3) / 1e12 if ph[1] > 0 else None}
        dtype = "f32 maps; f16 2-piece split, 3 products (f32 accumulate, f64 tile sums)"
        bessel = "cubic j_l table (h = 1/16) of the ring path"
    else:
        achieved = flops / (kms * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": achieved, "peak": PEAK_TFLOPS["fp64"], "unit": "TFLOP/s",
                "frac": achieved / PEAK_TFLOPS["fp64"], "traffic": None,
                "kernel": "sfb_direct_f64_kernel (pixel mode)", "kernel_ms": kms,
                "algorithmic_flops_per_launch": flops}
        dtype = "f64"
        bessel = "reference linear table (N > 1000)"
    out = {
        "metric": METRIC, "value": pairs / (ms * 1e-3), "unit": "pairs/s (point-mode equivalent)", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": dtype, "data": "synthetic, seeded (catana_amd/workloads.py), inputs resident in HBM",
        "config": {"workload": "healpix4m", "N": int(N), "nside": ns, "lmax": L, "k_max": cfg["k_max"], "rmax": R,
                   "pairs_per_step": pairs, "precision": args.precision, "bessel": bessel},
        "roofline": roof,
    }
    if ring and not args.no_accuracy:
        log("FP64 HEALPix path (exact j_l, then the reference table) for the accuracy figures")
        _, cr, _ = step(precision=prec, store_flmn=True)
        t64 = time.perf_counter()
        _, ce, _ = step(precision=capi.PREC_FP64, bessel=capi.BESSEL_EXACT)
        torch.cuda.synchronize()
        t64 = time.perf_counter() - t64
        _, ct, _ = step(precision=capi.PREC_FP64)
        torch.cuda.synchronize()
        a, b, t = cr.cpu().numpy(), ce.cpu().numpy(), ct.cpu().numpy()
        m = b != 0
        rel = np.abs(a[m] / b[m] - 1)
        relt = np.abs(a[m] / t[m] - 1)
        out["accuracy"] = {"max_rel_cln_vs_fp64_exact": float(rel.max()),
                           "median_rel_cln_vs_fp64_exact": float(np.median(rel)),
                           "max_rel_cln_vs_fp64_reference_table": float(relt.max()),
                           "median_rel_cln_vs_fp64_reference_table": float(np.median(relt))}
        out["fp64_gpu_exact_ms"] = t64 * 1e3
    if args.no_cpu_baseline:
        out["cpu_baseline"] = None
    else:
        # the oracle's restatement of _sfb_reverse_fft (every needed ring-DFT bin
        # evaluated directly) over all points and pixels, for every 16th multipole
        # from l = 8; pairs/s counts the sampled multipoles' pairs
        log("CPU baseline")
        sys.path.insert(0, ROOT)
        from oracle import oracle as O
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        off_h, rad_h = O.pixelize(pts, ns)
        sample = list(range(8, L, 16))
        nls = np.zeros_like(nl)
        nls[sample] = nl[sample]
        t0 = time.perf_counter()
        O.sfb_reverse_fft(off_h, rad_h, ns, L, NM, R, nl=nls, interpolated=N > 1000, nthreads=threads)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": N * int(((2 * l + 1) * nls).sum()) / dt,
                               "unit": "pairs/s (point-mode equivalent)", "cores": threads, "kind": "port",
                               "sample": f"all {N} points, nside {ns}, multipoles l = {sample} only "
                                         f"(n_l = {[int(nl[x]) for x in sample]}), reference linear table, "
                                         f"{dt:.2f} s"}
    h.close()
    print(json.dumps(out))


# FP64 VALU operations per node of sbessel_integral_kernel (ADD + MUL + 2 FMA + TRANS, x64 lanes),
# from rocprofv3 --pmc SQ_INSTS_VALU_*_F64 on the shotnoise2p25 workload (profiles/)
PROFILED_F64_OPS_PER_NODE = {"shotnoise2p25": 1186.0}


def run_shotnoise(args, cfg):
    """The shot-noise workload: a step is one sfbhip_sbessel_integrals call over
    every kept (l, n) mode, w = 1 on [0, rmax] as a two-sample table, r^2 on the
    device."""
    import torch
    from catana_amd import capi
    h = capi.Handle(0)
    L, NM, R = cfg["lmax"], cfg["nmax"], cfg["rmax"]
    k, nl = h.kln(capi.make_params(L, NM, R, k_max=cfg["k_max"]))
    ls = np.concatenate([np.full(nl[l], l, np.int32) for l in range(L)])
    ks = np.concatenate([k[l, :nl[l]] for l in range(L)])
    n = ls.size
    tab = (np.ones(2), 0.0, R * 1.001)
    nodes = capi.sbessel_nodes(h, ls, ks, ks, R).size
    for _ in range(args.warmup):
        capi.sbessel_integrals(h, ls, ks, ks, R, table=tab, r_power=2)
    torch.cuda.synchronize()
    kern = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        I = capi.sbessel_integrals(h, ls, ks, ks, R, table=tab, r_power=2)
        kern.append(h.last_kernel_ms())
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    kms = float(np.mean(kern))
    # closed form of the tophat case: int_0^R r^2 j_l(kR)^2 = R^3/2 [j_l^2 - j_{l-1} j_{l+1}](kR)
    from scipy.special import spherical_jn as J
    x = ks * R
    jm = np.where(ls == 0, np.cos(x) / x, J(np.maximum(ls - 1, 0), x))
    ref = R ** 3 / 2 * (J(ls, x) ** 2 - jm * J(ls + 1, x))
    rel = np.abs(I / ref - 1)
    ops = PROFILED_F64_OPS_PER_NODE.get("shotnoise2p25")
    roof = {"bound": "fp64-valu", "achieved": None, "peak": PEAK_TFLOPS["fp64"], "unit": "TFLOP/s", "frac": None,
            "traffic": None, "kernel": "sbessel_integral_kernel", "kernel_ms": kms, "nodes_per_launch": int(nodes)}
    if ops:
        a = ops * nodes / (kms * 1e-3) / 1e12
        roof.update(achieved=a, frac=a / PEAK_TFLOPS["fp64"], f64_ops_per_node=ops)
    out = {
        "metric": "double spherical-Bessel integrals/s (shot-noise term of every kept mode)",
        "value": n / (ms * 1e-3), "unit": "integrals/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: k_ln of the configuration, w = tophat",
        "config": {"workload": "shotnoise2p25", "lmax": L, "k_max": cfg["k_max"], "rmax": R, "integrals": int(n),
                   "quadrature_nodes": int(nodes)},
        "roofline": roof,
        "accuracy": {"vs_closed_form_max_rel": float(rel.max()), "median_rel": float(np.median(rel))},
    }
    if not args.no_cpu_baseline:
        sys.path.insert(0, ROOT)
        from oracle import oracle as O
        idx = np.arange(0, n, 8)
        t0 = time.perf_counter()
        for i in idx:
            O.sbessel_integral(int(ls[i]), R, ks[i], ks[i], power=2)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": idx.size / dt, "unit": "integrals/s", "cores": 1, "kind": "port",
                               "sample": f"{idx.size} of the {n} integrals (every 8th), oracle "
                                         "restatement of double_sbessel_integrator, adaptive Gauss-Legendre"}
    h.close()
    print(json.dumps(out))


def run_ingest(args, cfg):
    """The ingestion workload: a step is one sfbhip_filter_points call over all
    records already resident in HBM (conversion, three filters, compaction)."""
    import torch
    from catana_amd import capi
    h = capi.Handle(0)
    n, R, ns = cfg["n"], cfg["rmax"], cfg["nside"]
    rng = np.random.default_rng(21)
    rec = rng.uniform(0.0, 2 * R, (n, 3)).astype(np.float32)
    mask = (rng.uniform(size=12 * ns * ns) < 0.7).astype(np.float32)
    dev = torch.device("cuda:0")
    d_rec = torch.from_numpy(rec).to(dev)
    d_mask = torch.from_numpy(mask).to(dev)
    filters = [dict(kind=capi.FILTER_TOPHAT, a=R), dict(kind=capi.FILTER_ANGULAR_MASK, nside=ns, values=d_mask),
               dict(kind=capi.FILTER_GAUSSIAN, a=cfg["r0"])]

    def step(i):
        return capi.filter_points_torch(h, d_rec, capi.RECORD_CARTESIAN_F32, filters, seed=9, index_base=0,
                                        shift=-R)

    for i in range(args.warmup):
        out = step(i)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        out = step(i)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    kept = int(out.shape[0])
    algo = 12.0 * n + 24.0 * kept  # records read once, kept points written once
    achieved = algo / (ms * 1e-3) / 1e9
    res = {
        "metric": "records filtered/s (conversion + tophat + angular mask + Gaussian + compaction)",
        "value": n / (ms * 1e-3), "unit": "records/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32 in, f64 out",
        "data": "synthetic: uniform float32 cartesian records in a box of side 2000, seeded",
        "config": {"workload": "ingest2p25", "records": n, "kept": kept, "nside": ns, "R0_gauss": cfg["r0"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": 8000.0, "unit": "GB/s",
                     "frac": achieved / 8000.0, "traffic": PROFILED_TRAFFIC_BYTES.get(("ingest2p25",)),
                     "kernel": "ingest_flag_kernel + scan + compact_convert_kernel (whole call)",
                     "algorithmic_bytes_per_call": algo},
    }
    # parity on a sample: the device's kept set of the first 2^16 records equals the oracle's
    sys.path.insert(0, ROOT)
    from oracle import ingest as I
    m = 1 << 16
    sub = capi.filter_points_torch(h, d_rec[:m], capi.RECORD_CARTESIAN_F32, filters, seed=9, index_base=0,
                                   shift=-R).cpu().numpy()
    fo = [("tophat", R), ("mask", mask), ("gaussian", cfg["r0"])]
    exp, _ = I.apply_filters(I.cartesian_to_points(rec[:m].astype(np.float64), -R), fo, seed=9)
    res["accuracy"] = {"sample_records": m, "kept_equal": bool(sub.shape == exp.shape),
                       "max_abs_diff": float(np.abs(sub - exp).max()) if sub.shape == exp.shape and sub.size else None}
    if not args.no_cpu_baseline:
        t0 = time.perf_counter()
        I.apply_filters(I.cartesian_to_points(rec[:m].astype(np.float64), -R), fo, seed=9)
        dt = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": m / dt, "unit": "records/s", "cores": 1, "kind": "port",
                               "sample": f"{m} of {n} records, numpy restatement (oracle/ingest.py)"}
    h.close()
    print(json.dumps(res))


def run_batched(args):
    """The 64-catalogue workload.  A step is one replay of the captured batch
    plan (--batch-mode graph, the default), one eager enqueue of the plan
    (plan), or one sfbhip_direct_batched call (call)."""
    import torch
    from catana_amd import capi
    log("generating 64 catalogues")
    cats = make_batch()
    dev = torch.device("cuda:0")
    for c in cats:
        c["points"] = torch.from_numpy(np.ascontiguousarray(c["np"])).to(dev)
    h = capi.Handle(0)
    prec = capi.PREC_MIXED if args.precision == "mixed" else capi.PREC_FP64
    ops = {"fp32": capi.OPERANDS_FP32, "split16": capi.OPERANDS_SPLIT_FP16, "split3": capi.OPERANDS_SPLIT3_FP16}
    kw = dict(precision=prec, mixed_fp64_below=args.fp64_below, mixed_operands=ops[args.operands],
              mixed_chunks=args.chunks, mixed_kernel=KERNELS[args.kernel])
    pairs, flops_dom = 0, 0.0
    lo = min(args.fp64_below, 10 ** 9) if prec == capi.PREC_MIXED else 0
    for c in cats:
        _, nl = h.kln(capi.make_params(c["lmax"], c["nmax"], c["rmax"], k_max=c["k_max"]))
        l = np.arange(c["lmax"])
        pairs += c["np"].shape[0] * int(((2 * l + 1) * nl).sum())
        flops_dom += 4.0 * c["np"].shape[0] * int(((l[lo:] + 1) * nl[lo:]).sum())
    ntot = sum(c["np"].shape[0] for c in cats)
    log(f"{ntot} points in {len(cats)} catalogues")
    mode = args.batch_mode
    if mode != "call" and (prec != capi.PREC_MIXED or args.operands != "split16" or args.kernel == "512"):
        log("batch plans run the mixed split-FP16 256-thread path: falling back to --batch-mode call")
        mode = "call"
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    extra = {}
    if mode == "call":
        for _ in range(args.warmup):
            capi.sfb_direct_batched_torch(h, cats, **kw)
        torch.cuda.synchronize()
        stream = torch.cuda.current_stream(dev)
        kern, phases = [], []
        e0.record(stream)
        for _ in range(args.steps):
            res = capi.sfb_direct_batched_torch(h, cats, **kw)
            kern.append(h.last_kernel_ms())
            phases.append(h.last_phase_ms())
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / args.steps
        kms = float(np.mean(kern))
        ph = np.mean(np.array(phases), axis=0)
        kname = contraction_kernel(prec == capi.PREC_FP64, args.operands, args.kernel)
    else:
        pkw = dict(mixed_fp64_below=args.fp64_below, zeros=kw.get("zeros", capi.ZEROS_REFERENCE))
        if args.autotune:
            log("autotuning the tile table")
            t0 = time.perf_counter()
            extra["tile_table"] = capi.batch_autotune(h, cats, reps=2, **pkw)
            extra["autotune_s"] = time.perf_counter() - t0
        else:
            extra["tile_table"] = capi.get_tile_table(h)
        s = torch.cuda.Stream(device=dev)
        sp = s.cuda_stream
        t0 = time.perf_counter()
        plan = capi.BatchPlan(h, cats, **pkw)
        extra["plan_create_ms"] = (time.perf_counter() - t0) * 1e3
        run = plan.enqueue
        if mode == "graph":
            plan.capture(sp)
            run = plan.launch
        torch.cuda.synchronize()
        for _ in range(args.warmup):
            run(sp)
        plan.sync(sp)
        e0.record(s)
        for _ in range(args.steps):
            run(sp)
        e1.record(s)
        plan.sync(sp)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / args.steps
        # the grouped kernel's duration from HIP events of eager enqueues of the same plan
        phs = []
        for _ in range(max(3, min(args.steps, 10))):
            plan.enqueue(sp)
            phs.append(plan.sync(sp))
        ph = np.mean(np.array(phs), axis=0)
        kms = float(ph[1])
        res = plan.outputs
        kname = "sfb_direct_split16_grouped_kernel"
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        capi.sfb_direct_batched_torch(h, cats, **kw)
        torch.cuda.synchronize()
        extra["direct_batched_call_ms"] = (time.perf_counter() - t0) * 1e3
        extra["phases_ms_eager_plan"] = {"fp64_low_l": float(ph[0]), "grouped_contraction": float(ph[1]),
                                         "reduce_and_outputs": float(ph[2]), "whole_enqueue": float(ph[3])}
    # the same catalogues one call each (what a caller looping over sfb_decomposition does)
    t0 = time.perf_counter()
    for c in cats:
        capi.sfb_direct_torch(h, c["points"], c["lmax"], c["nmax"], c["rmax"], k_max=c["k_max"],
                              bessel=capi.BESSEL_EXACT, **kw)
    torch.cuda.synchronize()
    seq_ms = (time.perf_counter() - t0) * 1e3
    if prec == capi.PREC_FP64:
        dtype, peak = "f64", PEAK_TFLOPS["fp64"]
    elif args.operands == "split16":
        dtype, peak = "f16 2-piece split, 3 products (f32 accumulate, f64 tile sums)", PEAK_TFLOPS["fp16"] / 3.0
    elif args.operands == "split3":
        dtype, peak = "f16 3-piece split, 6 products (f32 accumulate, f64 tile sums)", PEAK_TFLOPS["fp16"] / 6.0
    else:
        dtype, peak = "f32 (f32 operands and accumulate, f64 tile sums)", PEAK_TFLOPS["fp32"]
    achieved = flops_dom / (kms * 1e-3) / 1e12
    out = {
        "metric": METRIC, "value": pairs / (ms * 1e-3), "unit": "pairs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": dtype, "data": "synthetic, seeded (bench.make_batch), inputs resident in HBM",
        "config": {"workload": "batched64", "catalogues": len(cats), "N_total": int(ntot),
                   "lmax_range": [min(c["lmax"] for c in cats), max(c["lmax"] for c in cats)], "k_max": 0.15,
                   "precision": args.precision, "operands": args.operands, "batch_mode": mode,
                   "pairs_per_step": pairs, "sequential_single_calls_ms": seq_ms, **extra},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
                     "traffic": None, "kernel": kname, "kernel_ms": kms, "algorithmic_flops_per_launch": flops_dom},
    }
    if mode == "call":
        out["roofline"]["phases_ms"] = {"fp64_contraction": float(ph[0]), "mixed_contraction": float(ph[1]),
                                        "reduce_and_cln": float(ph[2])}
    if prec != capi.PREC_FP64 and not args.no_accuracy:
        log("FP64 batched run for the accuracy figures")
        torch.cuda.synchronize()
        t64 = time.perf_counter()
        r64 = capi.sfb_direct_batched_torch(h, cats, precision=capi.PREC_FP64)
        torch.cuda.synchronize()
        t64 = time.perf_counter() - t64
        worst, meds = 0.0, []
        for (_, cm, _), (_, c64, _) in zip(res, r64):
            a, b = cm.cpu().numpy(), c64.cpu().numpy()
            m = b != 0
            rel = np.abs(a[m] / b[m] - 1)
            worst = max(worst, float(rel.max()))
            meds.append(float(np.median(rel)))
        out["fp64_gpu"] = {"ms_per_step": t64 * 1e3, "value": pairs / t64, "unit": "pairs/s",
                           "speedup_of_value": out["value"] / (pairs / t64)}
        out["accuracy"] = {"max_rel_cln_vs_fp64": worst, "median_rel_cln_vs_fp64": float(np.median(meds))}
    if not args.no_cpu_baseline:
        log("CPU baseline")
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        c = max(cats, key=lambda c: c["np"].shape[0])
        _, nl = h.kln(capi.make_params(c["lmax"], c["nmax"], c["rmax"], k_max=c["k_max"]))
        out["cpu_baseline"] = cpu_baseline(c["np"], dict(lmax=c["lmax"], nmax=c["nmax"], rmax=c["rmax"]), nl,
                                           args.cpu_sample or CPU_SAMPLE["batched64"], threads)
        out["cpu_baseline"]["sample"] += " (largest catalogue of the batch)"
    if mode != "call":
        plan.close()
    h.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
