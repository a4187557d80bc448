#!/usr/bin/env python3
"""Throughput of a_lm -> HEALPix synthesis on one MI355X.

Workloads (--workload):
  mixed  (default) Nside = 2048, lmax = 4096, nmaps = 64: the FP64 path and
         each mixed-precision contraction scheme (FP32 MFMA, FP16x3, BF16x3,
         BF16x6) on the same a_lm; one step = one synthesis of all 64 maps.
         Reports each scheme's maps/s and its rms / max error against the
         FP64 map for C_l = (l+1)^-2 and for flat C_l = 1, and the FP64
         path's maps/s at nmaps = 64 and at nmaps = 1.  `value` is the
         maps/s of the fastest scheme whose rms error meets 1e-6 for both
         spectra (FP32 MFMA if none does); `roofline` prices its contraction
         kernel.
  fp64   one step = one FP64 synthesis (Legendre sweep + ring
         stage) of an a_lm set already resident in HBM into a map in HBM,
         through the library's device-pointer execute.  Nside = 2048,
         lmax = mmax = 4096, nmaps = 1, seeded power-law a_lm
         (C_l = (l+1)^-2).
  grad   one step = one forward synthesis and one adjoint (the gradient of
         the weighted least-squares loss of the differentiable-synthesis
         configuration: Nside 512, lmax = mmax 1024, nmaps 3, weights 0 for
         |z| < 0.2 else uniform in [0.5, 2]), FP64, device buffers.
  batch  one step = one execute of a batch of 96 uneven requests (Nside
         uniform over {128 .. 2048}, lmax = mmax = 2 Nside, nmaps over
         {1, 3, 8}, seeded), FP64, replayed from the batch's HIP graph.

Prints one JSON line.  `roofline` prices the dominant kernel: for fp64 and
batch the Legendre sweep against the FP64 peak with the reference's flop
convention (5 + 4 nmaps flops per evaluated (l, m, ring-pair) triple); for
mixed the headline scheme's contraction kernel: its algorithmic flops
(4 nmaps per triple) against the raw dense peak of the MFMA type it runs on
(FP32 157.3, FP16 / BF16 2500 TFLOP/s); the rate of the split products it
executes (x1 FP32, x3 FP16x3 / BF16x3, x6 BF16x6) is a separate field, and
`ring_stage` prices the ring FFT stage against HBM bandwidth.  Kernel
durations are HIP events recorded on the execute stream
around the kernel.  `cpu_baseline` times a vectorised C restatement of the
reference's CPU path (oracle/sht_fast.c, OpenMP; the oracle's plan) on a
bounded sample of the same workload, min over 10 repetitions.
--calibrate-copy MIB runs a device copy of that size first (untimed), as
the known-size kernel against which HBM byte counters are scaled.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=["fp64", "batch", "mixed", "grad"], default="mixed")
    ap.add_argument("--mixed-nmaps", type=int, default=64)
    ap.add_argument("--schemes", default="f32,f16x3,bf16x3,bf16x6")
    ap.add_argument("--fp64-steps", type=int, default=3,
                    help="mixed workload: timed steps of the FP64 reference path")
    ap.add_argument("--nside", type=int, default=2048)
    ap.add_argument("--lmax", type=int, default=None)
    ap.add_argument("--nmaps", type=int, default=1)
    ap.add_argument("--nreq", type=int, default=96, help="batch workload: requests")
    ap.add_argument("--measure", action="store_true",
                    help="batch workload: autotune sweep shapes (WAVEMOTH_MEASURE)")
    ap.add_argument("--fly", action="store_true",
                    help="generate the Lambda seeds inside every execute")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-m-stride", type=int, default=None,
                    help="CPU baseline: Legendre stage on every k-th m (default 16 for the "
                         "mixed workload, 1 = all m for fp64)")
    ap.add_argument("--cpu-ring-stride", type=int, default=None,
                    help="CPU baseline: ring stage on every k-th ring pair (default 16 for "
                         "the mixed workload, 1 for fp64)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB")
    ap.add_argument("--seed", type=int, default=20482048)
    return ap.parse_args()


def host_info():
    """The host the CPU baseline ran on: CPU model, logical CPUs, physical
    cores (distinct (package, core) pairs), NUMA nodes and the CPUs this
    process may run on."""
    info = {"cpu": None, "logical_cpus": os.cpu_count(), "physical_cores": None,
            "numa_nodes": None, "affinity_cpus": None}
    try:
        cores, pkg, core = set(), None, None
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name" and info["cpu"] is None:
                    info["cpu"] = v
                elif k == "physical id":
                    pkg = v
                elif k == "core id":
                    core = v
                elif not k and pkg is not None:
                    cores.add((pkg, core))
                    pkg = core = None
        if pkg is not None:
            cores.add((pkg, core))
        info["physical_cores"] = len(cores) or None
    except OSError:
        pass
    try:
        nodes = [d for d in os.listdir("/sys/devices/system/node") if d.startswith("node")]
        info["numa_nodes"] = len(nodes)
    except OSError:
        pass
    try:
        info["affinity_cpus"] = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        pass
    return info


def min_time(fn, reps):
    """The reference's timing rule (bench/benchutils.h:113-137): one untimed
    warm-up call, then `reps` timed calls; (min, mean) seconds."""
    fn()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return min(ts), sum(ts) / len(ts)


def cpu_baseline(N, L, K, alm, m_stride, ring_stride, reps=10):
    """CPU baseline on a bounded sample of the nmaps = K workload, built from
    the oracle's plan (its precompute stands for the reference's resource
    files and is not timed, as the reference's plan is not):
      - Legendre stage: the vectorised restatement of the reference's SIMD
        kernel (oracle/sht_fast.c) on m = 0, m_stride, ... at nmaps = K,
        scaled to all m by the dense (l, m) count;
      - ring stage: fold + per-ring-pair c2r of all K maps (pocketfft for
        FFTW) on ring pairs 0, ring_stride, ..., scaled by the pair count.
    Each leg: one warm-up, then the minimum over `reps` repetitions."""
    import numpy as np
    from oracle import oracle as O
    M = L
    t0 = time.perf_counter()
    op = O.OraclePlan(N, L, M, K, m_start=0, m_stop=M + 1, m_step=m_stride)
    t_plan = time.perf_counter() - t0
    ms = range(0, M + 1, m_stride)
    dense_s = sum(L - m + 1 for m in ms)
    dense_all = sum(L - m + 1 for m in range(M + 1))
    threads = O.num_threads()
    work = np.zeros((len(ms), 2, op.R, K), np.complex128)
    width = [0]

    def leg():
        width[0] = op.legendre_stage_fast(alm, compact=True, work=work)[1]

    t_leg = min_time(leg, reps)
    T_s = op.trimmed_triples()
    del work
    rs = range(0, op.R, ring_stride)
    work_s = np.full((M + 1, 2, len(rs), K), 0.5 + 0.25j)
    t_ring = min_time(lambda: op.ring_stage_sample(work_s, ring_stride, workers=threads), reps)
    del work_s
    leg_full = t_leg[0] * dense_all / dense_s
    ring_full = t_ring[0] * op.R / len(rs)
    isa = {8: "AVX-512F", 4: "AVX2+FMA"}[width[0]]
    return {
        "value": K / (leg_full + ring_full),
        "unit": "maps/s",
        "cores": threads,
        "kind": "port",
        "sample": ("oracle/sht_fast.c (%s, %d OpenMP threads): the reference's SIMD Legendre "
                   "kernel restated (strips of rings sharing one recurrence, row blocks in "
                   "L1, register-blocked product) at nmaps=%d on %d of %d m values (every "
                   "%d-th), min %.3f s over %d repetitions after a warm-up, scaled by the "
                   "dense (l, m) count %d/%d; ring stage (fold, phase, pocketfft c2r of all "
                   "%d maps per ring, %d threads) on %d of %d ring pairs (every %d-th), min "
                   "%.3f s, scaled by %d/%d; plan precompute (%.1f s) not timed, as the "
                   "reference's resource precompute is not" %
                   (isa, threads, K, len(ms), M + 1, m_stride, t_leg[0], reps, dense_all,
                    dense_s, K, threads, len(rs), op.R, ring_stride, t_ring[0], op.R, len(rs),
                    t_plan)),
        "legendre_s": leg_full,
        "ring_s": ring_full,
        "legendre_sample_s": {"min": t_leg[0], "mean": t_leg[1]},
        "ring_sample_s": {"min": t_ring[0], "mean": t_ring[1]},
        "legendre_gflops": T_s * (5 + 4 * K) / t_leg[0] / 1e9,
        "flop_convention": "(5 + 4*nmaps) per (l, m, ring-pair) triple at or after first l",
        "isa": isa,
        "host": host_info(),
    }


def calibrate_copy(mib, dev):
    import torch
    n = mib * (1 << 20) // 8
    src = torch.ones(n, dtype=torch.float64, device=dev)
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    return {"bytes_read": n * 8, "bytes_written": n * 8, "launches": 3}


def measured_traffic(key):
    """HBM bytes per launch of a kernel from the committed rocprofv3 PMC summary
    (FETCH_SIZE + WRITE_SIZE, copy-calibrated), or None."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            return json.load(f)[key]["bytes"]
    except (OSError, KeyError, ValueError):
        return None


def roofline(flops, kernel_ms, kernel, extra):
    achieved = flops / (kernel_ms / 1e3) / 1e12
    r = {
        "bound": "mfma",
        "pipe": "fp64 VALU (MI355X FP64 vector and matrix peaks are both 78.6 TFLOP/s)",
        "achieved": achieved,
        "peak": FP64_PEAK_TFLOPS,
        "unit": "TFLOP/s",
        "frac": achieved / FP64_PEAK_TFLOPS,
        "traffic": None,
        "kernel": kernel,
        "flops_per_launch": flops,
        "flop_convention": "(5 + 4*nmaps) per evaluated (l, m, ring-pair) triple",
        "kernel_ms": kernel_ms,
    }
    r.update(extra)
    return r


def run_fp64(args, dev, res):
    import torch

    from wavemoth_amd import WAVEMOTH_SEEDS_ON_THE_FLY
    from wavemoth_amd.alm import power_law_alm
    from wavemoth_amd.torch_sht import DevicePlan

    N = args.nside
    L = args.lmax if args.lmax is not None else 2 * N
    K = args.nmaps
    alm_h = power_law_alm(L, L, K, args.seed)
    alm = torch.from_numpy(alm_h).to(dev)
    npix = 12 * N * N
    out = torch.empty((npix, K), dtype=torch.float64, device=dev)
    flags = WAVEMOTH_SEEDS_ON_THE_FLY if args.fly else 0
    t0 = time.perf_counter()
    plan = DevicePlan(N, L, L, K, flags)
    t_plan = time.perf_counter() - t0
    info = plan.info()

    for _ in range(args.warmup):
        plan.synthesize_into(alm, out)
    torch.cuda.synchronize()
    plan.set_profiling(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.synthesize_into(alm, out)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    stage_ms, nrec = plan.stage_times()
    plan.set_profiling(False)

    T_eps = info["triples_eps"]
    T_dense = info["triples_dense"]
    flops = (5 + 4 * K) * T_eps
    ring_bytes = 2 * (L + 1) * 2 * N * 2 * K * 16 + npix * K * 8
    res.update({
        "metric": "maps/s synthesis Nside=%d lmax=%d: FP64" % (N, L),
        "value": K * args.steps / (t1 - t0),
        "ms_per_step": (t1 - t0) * 1e3 / args.steps,
        "dtype": "f64",
        "data": "synthetic (seeded power-law a_lm, C_l=(l+1)^-2)",
        "config": {"workload": "a_lm->HEALPix synthesis, FP64, seeds %s" %
                               ("per execute" if args.fly else "cached in plan"),
                   "nside": N, "lmax": L, "mmax": L, "nmaps": K, "npix": npix},
        "roofline": roofline(flops, stage_ms[1], "k_legendre", {
            "traffic": measured_traffic("k_legendre nside=%d lmax=%d nmaps=%d" % (N, L, K)),
            "traffic_source": "profiles/traffic.json (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE; "
                              "x2.0 / x1.0 from a 1 GiB copy)",
            "triples_eps": T_eps, "triples_dense": T_dense,
            "triples_executed": info["triples_executed"],
            "achieved_dense_equiv": (5 + 4 * K) * T_dense / (stage_ms[1] / 1e3) / 1e12}),
        "stages_ms": {"prescale": stage_ms[0], "legendre": stage_ms[1], "rings": stage_ms[2],
                      "events": nrec},
        "ring_stage": {"bytes": ring_bytes,
                       "GB_per_s": ring_bytes / (stage_ms[2] / 1e3) / 1e9 if stage_ms[2] else None,
                       "frac_hbm": (ring_bytes / (stage_ms[2] / 1e3) / 1e9 / HBM_PEAK_GBS
                                    if stage_ms[2] else None)},
        "plan_s": t_plan,
        "output_finite": bool(torch.isfinite(out).all().item()),
    })
    if not args.no_cpu_baseline:
        res["cpu_baseline"] = cpu_baseline(N, L, K, alm_h, args.cpu_m_stride or 1,
                                           args.cpu_ring_stride or 1)


def run_grad(args, dev, res):
    """Forward + adjoint of the differentiable-synthesis configuration.  The
    adjoint's Legendre kernel does the forward sweep's flops, so it is priced
    with the same (5 + 4 nmaps) convention."""
    import numpy as np
    import torch

    from wavemoth_amd.alm import power_law_alm
    from wavemoth_amd.healpix import get_ring_offsets
    from wavemoth_amd.torch_sht import DevicePlan

    N, L, K = 512, 1024, 3
    npix = 12 * N * N
    alm = torch.from_numpy(power_law_alm(L, L, K, args.seed)).to(dev)
    d = torch.from_numpy(power_law_alm(L, L, K, args.seed + 1)).to(dev)
    plan = DevicePlan(N, L, L, K, 0)
    info = plan.info()
    data = torch.empty((npix, K), dtype=torch.float64, device=dev)
    plan.synthesize_into(d, data)
    g = torch.Generator(device="cpu").manual_seed(args.seed)
    data += torch.randn((npix, K), generator=g, dtype=torch.float64).to(dev)
    # weights: 0 for |z| < 0.2, else uniform in [0.5, 2] (per pixel)
    offs = get_ring_offsets(N)
    z = np.empty(npix)
    for i in range(4 * N - 1):
        ii = i + 1
        zi = 1 - ii * ii / (3.0 * N * N) if ii < N else (
            4.0 / 3 - 2.0 * ii / (3 * N) if ii <= 3 * N else -(1 - (4 * N - ii) ** 2 / (3.0 * N * N)))
        z[offs[i]:offs[i + 1]] = zi
    w = torch.from_numpy(np.where(np.abs(z) < 0.2, 0.0,
                                  0.5 + 1.5 * np.random.default_rng(args.seed).random(npix)))
    w = w.to(dev).unsqueeze(1)
    out = torch.empty((npix, K), dtype=torch.float64, device=dev)
    grad = torch.empty_like(alm)

    def step():
        plan.synthesize_into(alm, out)
        r = w * (out - data)
        plan.adjoint_into(r, grad)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    fwd = adj = 0.0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ev[0].record()
        plan.synthesize_into(alm, out)
        ev[1].record()
        r = w * (out - data)
        plan.adjoint_into(r, grad)
        ev[2].record()
        ev[2].synchronize()
        fwd += ev[0].elapsed_time(ev[1])
        adj += ev[1].elapsed_time(ev[2])
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    T_eps = info["triples_eps"]
    adj_ms = adj / args.steps
    res.update({
        "metric": "maps/s differentiable synthesis (forward + adjoint) Nside=%d lmax=%d "
                  "nmaps=%d: FP64" % (N, L, K),
        "value": K * args.steps / (t1 - t0),
        "ms_per_step": (t1 - t0) * 1e3 / args.steps,
        "dtype": "f64",
        "data": "synthetic (seeded power-law a_lm and data, N(0,1) noise, the configuration's "
                "weights)",
        "config": {"workload": "gradient of the weighted least-squares loss: synthesis + "
                               "adjoint (unweighted analysis)",
                   "nside": N, "lmax": L, "mmax": L, "nmaps": K, "npix": npix},
        "roofline": roofline((5 + 4 * K) * T_eps, adj_ms, "adjoint (k_ring_adjoint + "
                             "k_legendre_adj + k_adj_finish)", {
            "note": "whole adjoint execute timed with events (its kernels together)"}),
        "stages_ms": {"forward": fwd / args.steps, "residual_and_adjoint": adj_ms},
        "grad_finite": bool(torch.isfinite(grad).all().item()),
    })


def run_batch(args, dev, res):
    import torch

    from wavemoth_amd import WAVEMOTH_MEASURE, BatchPlan, random_requests
    from wavemoth_amd.torch_sht import random_alm

    reqs = random_requests(args.nreq, seed=args.seed % (1 << 31))
    t0 = time.perf_counter()
    b = BatchPlan(reqs, WAVEMOTH_MEASURE if args.measure else 0)
    t_plan = time.perf_counter() - t0
    info = b.info()
    alms = [random_alm(L, M, K, args.seed + i, dev) for i, (N, L, M, K) in enumerate(reqs)]
    outs = [torch.empty((12 * N * N, K), dtype=torch.float64, device=dev)
            for (N, L, M, K) in reqs]
    for _ in range(args.warmup):
        b.synthesize_into(alms, outs)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.synthesize_into(alms, outs)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    # stage split from direct-launch executes (profiling bypasses the graph)
    b.set_profiling(True)
    for _ in range(2):
        b.synthesize_into(alms, outs)
    stage_ms, nrec = b.stage_times()
    b.set_profiling(False)
    maps = info["maps"]
    shapes = {}
    for r in range(len(reqs)):
        ri = b.request_info(r)
        shapes["%d/%d" % (ri["Nside"], ri["nmaps"])] = [ri["rings_per_lane"],
                                                       ri["maps_per_pass"],
                                                       ri["wrap_register"]]
    res.update({
        "metric": "maps/s batched uneven synthesis (%d requests): FP64" % len(reqs),
        "value": maps * args.steps / (t1 - t0),
        "ms_per_step": (t1 - t0) * 1e3 / args.steps,
        "dtype": "f64",
        "data": "synthetic (seeded power-law a_lm, C_l=(l+1)^-2, drawn on the device)",
        "config": {"workload": "batched uneven a_lm->HEALPix synthesis, FP64, hipGraph replay",
                   "requests": len(reqs), "maps": maps,
                   "nside_choices": [128, 256, 512, 1024, 2048], "nmaps_choices": [1, 3, 8],
                   "lmax": "2*Nside", "autotuned": bool(args.measure)},
        "roofline": roofline(info["flops"], stage_ms[1], "k_batch_legendre", {
            "triples_eps": info["triples_eps"], "triples_dense": info["triples_dense"],
            "slots_executed": info["slots_executed"],
            "sweep_items": info["sweep_items"], "sweep_workgroups": info["sweep_workgroups"]}),
        "stages_ms": {"prescale": stage_ms[0], "legendre": stage_ms[1], "rings": stage_ms[2],
                      "events": nrec, "note": "direct launches; value uses graph replays"},
        "sweep_shapes": {"Nside/nmaps -> [rings_per_lane, maps_per_pass, wrap_register]":
                         dict(sorted(shapes.items()))},
        "plan_s": t_plan,
        "device_bytes": info["device_bytes"],
        "output_finite": all(bool(torch.isfinite(o).all().item()) for o in outs),
    })


SCHEMES = {"f32": (0x1000, "FP32 MFMA", 157.3, 1),
           "bf16x3": (0x2000, "BF16x3", 2500.0, 3),
           "bf16x6": (0x4000, "BF16x6", 2500.0, 6),
           "f16x3": (0x8000, "FP16x3", 2500.0, 3)}


def _time_plan(plan, alm, out, warmup, steps):
    import torch
    for _ in range(warmup):
        plan.synthesize_into(alm, out)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.synthesize_into(alm, out)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    plan.set_profiling(True)
    for _ in range(2):
        plan.synthesize_into(alm, out)
    stage_ms, _ = plan.stage_times()
    plan.set_profiling(False)
    return dt, stage_ms


def _err(x, ref):
    import torch
    d = x - ref
    return {"rms": float(torch.linalg.norm(d) / torch.linalg.norm(ref)),
            "max": float(d.abs().max() / torch.sqrt(torch.mean(ref * ref)))}


def run_mixed(args, dev, res):
    import torch

    from wavemoth_amd.torch_sht import DevicePlan, random_alm

    N = args.nside
    L = args.lmax if args.lmax is not None else 2 * N
    K = args.mixed_nmaps
    npix = 12 * N * N
    alms = {"powerlaw": random_alm(L, L, K, args.seed, dev),
            "flat": random_alm(L, L, K, args.seed + 1, dev, flat=True)}
    out = torch.empty((npix, K), dtype=torch.float64, device=dev)

    # FP64 reference maps and FP64 throughput at the same nmaps
    p64 = DevicePlan(N, L, L, K)
    info = p64.info()
    T_eps = info["triples_eps"]
    dt64, st64 = _time_plan(p64, alms["powerlaw"], out, 1, args.fp64_steps)
    refs = {}
    for name, a in alms.items():
        p64.synthesize_into(a, out)
        torch.cuda.synchronize()
        refs[name] = out.clone()
    del p64
    # FP64 at nmaps = 1 (the FP64 headline shape)
    p1 = DevicePlan(N, L, L, 1)
    a1 = alms["powerlaw"][:, :1].contiguous()
    o1 = torch.empty((npix, 1), dtype=torch.float64, device=dev)
    dt1, st1 = _time_plan(p1, a1, o1, 2, 10)
    del p1, a1, o1

    schemes = {}
    for key in [k.strip() for k in args.schemes.split(",") if k.strip()]:
        flag, label, peak, nprod = SCHEMES[key]
        p = DevicePlan(N, L, L, K, flag)
        dt, st = _time_plan(p, alms["powerlaw"], out, args.warmup, args.steps)
        errs = {}
        for name, a in alms.items():
            p.synthesize_into(a, out)
            torch.cuda.synchronize()
            errs[name] = _err(out, refs[name])
        del p
        flops = 4.0 * K * T_eps
        # FP32 q read by the ring stage + the FP64 map it writes
        ring_bytes = (2 * N) * 2 * (L + 1) * K * 8 + npix * K * 8
        schemes[key] = {
            "label": label, "maps_per_s": K / dt, "ms_per_step": dt * 1e3,
            "stages_ms": {"prescale": st[0], "contraction": st[1], "rings": st[2]},
            "contraction_tflops": flops / (st[1] / 1e3) / 1e12,
            "mfma_peak_tflops": peak,
            "frac": flops / (st[1] / 1e3) / 1e12 / peak,
            "executed_product_tflops": nprod * flops / (st[1] / 1e3) / 1e12,
            "executed_product_frac": nprod * flops / (st[1] / 1e3) / 1e12 / peak,
            "ring_stage": {"bytes": ring_bytes,
                           "GB_per_s": ring_bytes / (st[2] / 1e3) / 1e9,
                           "frac_hbm": ring_bytes / (st[2] / 1e3) / 1e9 / HBM_PEAK_GBS},
            "errors_vs_fp64": errs,
            "meets_1e-6": all(e["rms"] <= 1e-6 for e in errs.values()),
        }
    ok = [k for k in schemes if schemes[k]["meets_1e-6"]]
    if ok:
        main_key = max(ok, key=lambda k: schemes[k]["maps_per_s"])
    else:
        main_key = "f32" if "f32" in schemes else next(iter(schemes))
    m = schemes[main_key]
    nprod = SCHEMES[main_key][3]
    kidx = ["f32", "bf16x3", "bf16x6", "f16x3"].index(main_key)
    ncp = (2 * K + 31) // 32 * 32
    # the contraction kernel the plan launches (FP16x3: Lambda built in
    # registers, k_mixr<CT>; the others: producer / consumer k_mix<S, CB>)
    kname = ("k_mixr<%d>" % (1 if ncp <= 64 else 2) if main_key == "f16x3"
             else "k_mix<%d, %d>" % (kidx, min(ncp // 32, 4)))
    res.update({
        "metric": "maps/s synthesis Nside=%d lmax=%d: FP64 vs batched FP32-MFMA @<=1e-6 err "
                  "(value: %s contraction, the fastest scheme within 1e-6)" % (N, L, m["label"]),
        "value": m["maps_per_s"],
        "ms_per_step": m["ms_per_step"],
        "dtype": {"f32": "f32 MFMA", "bf16x3": "bf16x3 split products on bf16 MFMA",
                  "bf16x6": "bf16x6 split products on bf16 MFMA",
                  "f16x3": "f16x3 split products on f16 MFMA"}[main_key]
                 + " contraction, FP32 accumulation, FP64 l-chunk reduction; FP64 "
                   "recurrence; FP32 q and ring FFTs; FP64 map",
        "scheme": main_key,
        "data": "synthetic (seeded a_lm drawn on the device; C_l=(l+1)^-2 timed, flat C_l=1 "
                "also checked)",
        "config": {"workload": "a_lm->HEALPix synthesis of %d maps, mixed-precision "
                               "contraction vs FP64" % K,
                   "nside": N, "lmax": L, "mmax": L, "nmaps": K, "npix": npix},
        "roofline": {
            "bound": "mfma", "achieved": m["contraction_tflops"],
            "peak": SCHEMES[main_key][2], "unit": "TFLOP/s",
            "frac": m["frac"],
            "executed_product_tflops": m["executed_product_tflops"],
            "executed_product_frac": m["executed_product_frac"],
            "traffic": measured_traffic("%s nside=%d lmax=%d nmaps=%d" % (kname, N, L, K)),
            "traffic_source": "profiles/traffic.json (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE; "
                              "x2.0 / x1.0 from a 1 GiB copy)",
            "kernel": kname,
            "flops_per_launch": 4.0 * K * T_eps,
            "flop_convention": "4*nmaps per (l, m, ring-pair) triple at or after first l (the "
                               "contraction's algorithmic flops, priced against the raw dense "
                               "peak of the MFMA type; the FP64 recurrence is not counted; "
                               "executed_product_* count the %d split products per MAC)" % nprod,
            "kernel_ms": m["stages_ms"]["contraction"]},
        "ring_stage": m["ring_stage"],
        "fp64": {"nmaps": K, "maps_per_s": K / dt64, "ms_per_step": dt64 * 1e3,
                 "stages_ms": {"prescale": st64[0], "legendre": st64[1], "rings": st64[2]},
                 "nmaps1_maps_per_s": 1.0 / dt1, "nmaps1_stages_ms": st1},
        "speedup_vs_fp64_same_nmaps": m["maps_per_s"] / (K / dt64),
        "speedup_vs_fp64_nmaps1": m["maps_per_s"] * dt1,
        "schemes": schemes,
        "triples_eps": T_eps,
    })
    if not args.no_cpu_baseline:
        a_h = alms["powerlaw"].cpu().numpy()
        res["cpu_baseline"] = cpu_baseline(N, L, K, a_h, args.cpu_m_stride or 16,
                                           args.cpu_ring_stride or 16)


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        sys.stderr.write("bench.py runs on exactly one GPU: wavemoth_amd has no multi-GPU "
                         "path (--gpus %d, WORLD_SIZE %d requested)\n" % (args.gpus, world))
        return 2
    import torch

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    res = {"metric": None, "value": None, "unit": "maps/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": None, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None}
    if args.calibrate_copy:
        res["copy_calibration"] = calibrate_copy(args.calibrate_copy, dev)
    if args.workload == "fp64":
        run_fp64(args, dev, res)
    elif args.workload == "mixed":
        run_mixed(args, dev, res)
    elif args.workload == "grad":
        run_grad(args, dev, res)
    else:
        run_batch(args, dev, res)
    res.setdefault("cpu_baseline", None)
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
