#!/usr/bin/env python3
"""Benchmark of the DWF Wilson hopping term (Dhop) on one MI355X.

Workload (the headline): 24^4 lattice x Ls=16, FP64 storage and arithmetic,
one right-hand side, seeded SU(3) links and Gaussian spinor, periodic.  One
"step" = one application of Dhop to the whole 5D lattice
(5,308,416 5D sites x 1320 flop, bench.cc:145).  Inputs are resident in HBM
before the timed region; the timed region is K applications bracketed by
stream synchronisations.

Also reported (same lattice, same seeds; `variants`): FP32 1-RHS, FP32 and
FP16 storage with 16 right-hand sides, DhopDag, each with its throughput per
right-hand side relative to FP64 1-RHS and its error against the FP64 result.

Prints ONE JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

DIMS = (24, 24, 24, 24)
LS = 16
GAUGE_SEED, SPINOR_SEED = 201, 202  # seeds of the 24^4 x 16 configuration
HBM_PEAK_GBPS = 8000.0  # MI355X HBM3E, AMD-published
FP64_VECTOR_TFLOPS = 78.6


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def traffic_from_profiles():
    """HBM bytes per launch of the FP64 kernel, from the committed
    rocprofv3 PMC summary (FETCH_SIZE + WRITE_SIZE, copy-calibrated)."""
    path = os.path.join(ROOT, "profiles", "dhop_f64_24x24x24x24_ls16_pmc.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def cpu_baseline(dbl, psi, threads, budget_s=12.0):
    """The oracle (plain-C restatement of the reference CPU Hand kernel) with
    OpenMP over 4D sites on this host, on the same 24^4 x 16 FP64 inputs."""
    from oracle import oracle as orc
    V = int(np.prod(DIMS))
    nbr, prm = orc.tables(DIMS, 1, LS)
    U = orc.pack_gauge(DIMS, 1, dbl)
    inp = orc.pack_spinor(DIMS, 1, LS, psi)
    orc.hand(U, inp, nbr, prm, V, LS, 1, nthreads=threads)  # rep 0, untimed (reference convention)
    t0 = time.perf_counter()
    n = 0
    while True:
        orc.hand(U, inp, nbr, prm, V, LS, 1, nthreads=threads)
        n += 1
        if time.perf_counter() - t0 >= budget_s or n >= 50:
            break
    dt = (time.perf_counter() - t0) / n
    gflops = 1320.0 * V * LS / dt / 1e9
    return {"value": round(gflops, 3), "unit": "GFLOP/s", "cores": threads, "kind": "port",
            "sample": f"24^4 x Ls16 FP64 1-RHS, {n} full applications after 1 untimed (oracle, OpenMP "
                      f"{threads} threads, -O3 -march=x86-64-v3)", "ms_per_application": round(dt * 1e3, 2)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-variants", action="store_true", help="only the FP64 1-RHS headline")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-tune", action="store_true",
                    help="use the default launch configuration (profiling runs: no tuner launches)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (the hopping term on one "
              "GPU's lattice is the whole measured path; the multi-GPU halo-exchange design is described in "
              "DESIGN.md but not built).", file=sys.stderr)
        sys.exit(2)

    import gridbench_amd as gbx
    if gbx.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(3)

    V = int(np.prod(DIMS))
    sites5 = V * LS
    log("generating inputs")
    dbl = gbx.gen_doubled_gauge(DIMS, GAUGE_SEED)
    psi = gbx.gen_spinor(DIMS, LS, 1, SPINOR_SEED)[0]

    geom = gbx.Geometry(DIMS)
    U64 = gbx.GaugeField(geom, "f64").upload(dbl)
    x64 = gbx.SpinorField(geom, LS, 1, "f64").upload_rhs(0, psi)
    y64 = gbx.SpinorField(geom, LS, 1, "f64")

    if args.no_tune:
        best = gbx.dhop_get_cfg(U64, x64)
    else:
        log("tuning FP64 1-RHS")
        best, best_us = gbx.tune(U64, y64, x64)
        log(f"tuned: {best.as_dict()} {best_us:.1f} us")

    # ------------------------------------------------ timed headline region
    for _ in range(args.warmup):
        gbx.dhop(U64, y64, x64)
    geom.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        gbx.dhop(U64, y64, x64)
    geom.sync()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / args.steps
    gflops = 1320.0 * sites5 / (ms * 1e-3) / 1e9
    # kernel duration from HIP events on the kernel's own stream
    us_ev = gbx.dhop_time(U64, y64, x64, nrep=args.steps)
    bytes_alg = gbx.bytes_per_site("f64", "f64", LS, 1) * sites5
    achieved = bytes_alg / (us_ev * 1e-6) / 1e9
    log(f"FP64 1-RHS: {ms:.3f} ms/step wall, {us_ev:.1f} us/launch events, {gflops:.1f} GFLOP/s, {achieved:.0f} GB/s")
    ref64 = y64.download_rhs(0)

    stream_gbps = gbx.stream_copy_bw(4 << 30, 20)
    log(f"STREAM copy: {stream_gbps:.0f} GB/s")

    variants = []
    if not args.no_variants:
        def variant(name, prec, nrhs, dag=False, comp12=False):
            U = U64 if prec == "f64" else gbx.GaugeField(geom, prec, comp12).upload(dbl)
            x = gbx.SpinorField(geom, LS, nrhs, prec)
            for r in range(nrhs):  # right-hand side r = stream r of the spinor seed (gbx_gen_spinor)
                x.upload_rhs(r, psi if r == 0 else gbx.gen_gaussian(SPINOR_SEED, r, 0, V * LS * 24))
            y = gbx.SpinorField(geom, LS, nrhs, prec)
            cfg, tus = gbx.tune(U, y, x, dag=dag)
            us = gbx.dhop_time(U, y, x, dag=dag, nrep=20)
            gbx.dhop(U, y, x, dag=dag)
            got = y.download_rhs(0)
            if dag:
                want = ref64_dag
            else:
                want = ref64
            err = float(np.linalg.norm((got - want).ravel()) / np.linalg.norm(want.ravel()))
            per_rhs_us = us / nrhs
            b = gbx.bytes_per_site(prec, prec, LS, nrhs, comp12) * sites5 * nrhs
            v = {"name": name, "prec": prec, "nrhs": nrhs, "dag": dag, "comp12": comp12,
                 "us_per_launch": round(us, 2), "gflops": round(1320.0 * sites5 * nrhs / (us * 1e-6) / 1e9, 1),
                 "hbm_gbps_alg": round(b / (us * 1e-6) / 1e9, 1),
                 "per_rhs_speedup_vs_f64_1rhs": round(us_ev / per_rhs_us, 3),
                 "rel_l2_vs_f64": float(f"{err:.3e}"), "cfg": cfg.as_dict()}
            log(f"variant {v}")
            variants.append(v)
            for f in (x, y):
                f.close()

        gbx.dhop(U64, y64, x64, dag=True)
        ref64_dag = y64.download_rhs(0)
        variant("f64_1rhs_dag", "f64", 1, dag=True)
        variant("f32_1rhs", "f32", 1)
        variant("f32_16rhs", "f32", 16)
        variant("f16_16rhs", "f16", 16)
        variant("bf16_16rhs", "bf16", 16)
        variant("f16_16rhs_comp12", "f16", 16, comp12=True)

    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        log(f"CPU baseline with {threads} threads")
        cpu = cpu_baseline(dbl, psi, threads)
        log(f"CPU baseline: {cpu}")

    result = {
        "metric": "DWF Dhop GFLOP/s + HBM GB/s @24^4xLs16: FP64 1-RHS vs FP32/FP16 16-RHS",
        "value": round(gflops, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded SU(3) links, Gaussian spinor; gauge seed 201, spinor seed 202)",
        "config": {"workload": "DWF Dhop (Hand DhopSite, non-dagger) 24^4 x Ls16 FP64 1-RHS", "dims": list(DIMS),
                   "Ls": LS, "nrhs": 1, "flop_per_5d_site": 1320, "kernel_cfg": best.as_dict()},
        "hbm_gbps_alg": round(achieved, 1),
        "achievable_hbm_gbps_stream_copy": round(stream_gbps, 1),
        "frac_of_achievable_hbm": round(achieved / stream_gbps, 3),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": traffic_from_profiles(),
                     "algorithmic_bytes_per_launch": bytes_alg, "us_per_launch_events": round(us_ev, 2)},
        "cpu_baseline": cpu,
        "variants": variants,
    }
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
