#!/usr/bin/env python3
"""Benchmark of the DWF Wilson hopping term (Dhop) on one MI355X.

Workload (the headline): 24^4 lattice x Ls=16, FP64 storage and arithmetic,
one right-hand side, seeded SU(3) links and Gaussian spinor, periodic.  One
"step" = one application of Dhop to the whole 5D lattice
(5,308,416 5D sites x 1320 flop, bench.cc:145).  Inputs are resident in HBM
before the timed region; the timed region is K applications bracketed by
stream synchronisations.

Also reported, each with its own seeds:
  variants   24^4 x Ls16: DhopDag FP64; FP32 1-RHS; FP32/FP16/BF16 storage
             with 16 right-hand sides (+ FP16 with 12-real links); 24^4 x
             Ls12 x 16 RHS in FP32 and BF16 (gauge seed 401, spinors 402).
             Throughput per right-hand side relative to FP64 1-RHS (per 5D
             site: the Ls12 forms are scaled by 12/16), error against the
             FP64 kernel on the same input; for the multi-RHS rows each
             kernel form (VALU, MFMA, LDS-tiled) is also autotuned and timed
             on its own ("forms").
  mixed      32^4 x Ls16: Dhop error of FP32/BF16/FP16/FP16+12-real storage
             against FP64 on the same input.
  cg         32^4 x Ls16: defect-correction CG on M^dag M x = b,
             M = 1 - 0.12 Dhop, b Gaussian, to 1e-10 relative residual, inner
             solves in FP32/BF16/FP16 (and FP64 only, for comparison); and
             M x = b by the even-odd Schur complement on a red-black pair
             against CG on the normal equations over the whole lattice.
  dwf        24^4 x Ls16 domain-wall operator (mass 0.1, M5 1.0): D x = b
             by the even-odd Schur complement, FP32 / FP16 inner solves.
  batch      40 lattices L^4 x Ls x Nrhs (L in {8,12,16,20,24}, Ls in
             {8,12,16}, Nrhs in 1..8, drawn with seed 505), FP32, one
             persistent graph-captured launch; aggregate GFLOP/s and the
             per-lattice error against FP64.
  halo       24^4 x Ls16 FP64 cut by process grids 1x1x1x2, 1x1x1x4 and
             1x2x2x2 (domain parts with ghost slots): the assembled result
             against the whole lattice's, and for one block the Dhop time on
             its share, on its interior and on its faces, the face pack /
             unpack times, the face bytes, and one rank's whole schedule with
             the faces moved on this GPU (serial, and overlapped with the
             interior on a second stream); the xGMI transfer time is
             modelled (bytes / link rate), not measured.
  dropin     the reference entry dslash_kernel (gbx_dslash_kernel_f64 /
             _f32) on host arrays in the reference layout with
             TableGenerate's tables at 24^4 x Ls16, as bench.cc calls it:
             its own device time (applications 1..nrep-1) and the whole
             call's wall time.

Launch configurations come from the autotuner; its results for this machine
are cached in gridbench_amd/tuned_mi355x.jsonl (re-validated on every run).

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
# The CPU baseline's OpenMP threads, one per core, packed (read when the
# OpenMP runtime starts, i.e. before the libraries below are loaded).
os.environ.setdefault("OMP_PLACES", "cores")
os.environ.setdefault("OMP_PROC_BIND", "close")

DIMS = (24, 24, 24, 24)
LS = 16
GAUGE_SEED, SPINOR_SEED = 201, 202  # seeds of the 24^4 x 16 configuration
# Traversal block of the FP64 1-RHS rows (the headline).  Longer along t
# than the library's default 6^4: with the TDIAG basis's half-spinor t legs,
# a sweep over block shapes found 4x6x6x12 fastest for this problem (FP32
# 16-RHS keeps the default).
FP64_BLOCK = (4, 6, 6, 12)
# Traversal blocks of the 16-bit 16-RHS rows (12 KB per site at Ls16): a
# smaller 4D block keeps each XCD's sites in flight more compact; measured
# 4-4.5 % faster than 6^4 for FP16 / BF16 at Ls16 and Ls12.
MULTI16_BLOCK = (2, 3, 3, 3)
C4_GAUGE_SEED, C4_SPINOR_SEED = 401, 402  # 24^4 x Ls12 x 16 right-hand sides
HBM_PEAK_GBPS = 8000.0  # MI355X HBM3E, AMD-published
KAPPA = 0.12
TUNE_CACHE = os.path.join(ROOT, "gridbench_amd", "tuned_mi355x.jsonl")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# Headline check: the TDIAG result, set back to the chiral basis, against the
# chiral kernel (itself pinned to the oracle at <= 1e-14 by the tests).
HEADLINE_CHECK_TOL = 1e-13
# rocprofv3 PMC summaries (TCC read/write requests, copy-calibrated) of the
# kernels behind the headline and the 16-RHS rows, with the launch
# configuration each one measured: a row's traffic figure is used only when
# the row ran that configuration.
PROFILE_OF = {"f64_1rhs": ("dhop_f64_24x24x24x24_ls16_pmc.md",
                           {"kind": 0, "spw": 4, "threads": 64, "mc": 0, "pf": 2, "tile": 0}),
              "f32_16rhs": ("dhop_tile_f32_24x24x24x24_ls16_n16_stats_pmc.md",
                            {"kind": 2, "threads": 768, "mc": 64, "pf": 0, "tile": 0x2223}),
              "f16_16rhs": ("dhop_mfma_f16_bf16_24x24x24x24_n16_stats_pmc.md",
                            {"kind": 1, "spw": 4, "threads": 256, "mc": 0, "pf": 0, "tile": 0})}


def traffic_from_profiles(row, cfg):
    """HBM bytes per launch of a row's kernel, from its committed rocprofv3
    PMC summary (the `traffic_bytes_per_launch:` line); None when the row ran
    another configuration than the profiled one, or the summary is missing."""
    name, profiled = PROFILE_OF[row]
    if any(cfg.get(k) != v for k, v in profiled.items()):
        return None
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def cpu_baseline(dbl, psi, threads, budget_s=8.0):
    """The CPU Hand kernel on this host: the oracle's C restatement of the
    reference kernel (WilsonKernelsHandCpu.h:385-451 semantics) built for
    timing (-O3, AVX2 or AVX-512 as the host allows, multiply-adds fused),
    OpenMP over 4D sites, Nsimd 1, on the same 24^4 x 16 FP64 inputs, rep 0
    untimed as the reference times.  Also: a thread sweep at 24^4 x 16 and the
    reference's CPU configuration 8^4 x Ls8 at its nrep = 1000 (bench.cc:146),
    one thread and all threads.  About 25 s of CPU time in all."""
    from oracle import oracle as orc
    V = int(np.prod(DIMS))
    nbr, prm = orc.tables(DIMS, 1, LS)
    U = orc.pack_gauge(DIMS, 1, dbl)
    inp = orc.pack_spinor(DIMS, 1, LS, psi)
    out = orc.aligned_empty(inp.size, np.float64)

    def per_app(nthr, budget, nmax):
        orc.port_hand(U, inp, nbr, prm, V, LS, nthr, out)  # rep 0, untimed
        t0 = time.perf_counter()
        n = 0
        while True:
            orc.port_hand(U, inp, nbr, prm, V, LS, nthr, out)
            n += 1
            if time.perf_counter() - t0 >= budget or n >= nmax:
                break
        return (time.perf_counter() - t0) / n, n

    flop = 1320.0 * V * LS
    dt, n = per_app(threads, budget_s, 50)
    # the timing build against the oracle (fused multiply-adds round differently)
    want = orc.hand(U, inp, nbr, prm, V, LS, 1, nthreads=threads)
    port_err = float(np.linalg.norm(out - want) / np.linalg.norm(want))
    del want
    sweep = []
    for t in (1, 2, 4, 8, 16, 32, 64):
        if t > threads:
            break
        d, k = (dt, n) if t == threads else per_app(t, 2.0, 20)
        sweep.append({"threads": t, "gflops": round(flop / d / 1e9, 2), "applications": k})
    if threads not in [r["threads"] for r in sweep]:
        sweep.append({"threads": threads, "gflops": round(flop / dt / 1e9, 2), "applications": n})
    # 8^4 x Ls8 (the reference's CPU configuration), nrep = 1000 with rep 0
    # untimed, like dslash_kernel_cpu's own timer (WilsonKernelsHandCpu.h:407-450)
    d1, Ls1 = (8, 8, 8, 8), 8
    V1 = int(np.prod(d1))
    nbr1, prm1 = orc.tables(d1, 1, Ls1)
    import gridbench_amd as gbx
    U1 = orc.pack_gauge(d1, 1, gbx.gen_doubled_gauge(d1, 101))
    in1 = orc.pack_spinor(d1, 1, Ls1, gbx.gen_spinor(d1, Ls1, 1, 102)[0])
    out1 = orc.aligned_empty(in1.size, np.float64)
    c1 = []
    for t in sorted({1, threads}):
        orc.port_hand(U1, in1, nbr1, prm1, V1, Ls1, t, out1)
        t0 = time.perf_counter()
        for _ in range(999):
            orc.port_hand(U1, in1, nbr1, prm1, V1, Ls1, t, out1)
        d = (time.perf_counter() - t0) / 999
        c1.append({"threads": t, "gflops": round(1320.0 * V1 * Ls1 / d / 1e9, 2), "us_per_application": round(d * 1e6, 1)})
    build = orc.port_lib()[1]
    return {"value": round(flop / dt / 1e9, 3), "unit": "GFLOP/s", "cores": threads, "kind": "port",
            "sample": f"24^4 x Ls16 FP64 1-RHS, {n} full applications after 1 untimed (the oracle's C restatement "
                      f"built for timing: {build}, -O3, fused multiply-adds; OpenMP {threads} threads)",
            "ms_per_application": round(dt * 1e3, 2), "rel_l2_vs_oracle": float(f"{port_err:.3e}"),
            "thread_sweep_24x24x24x24_ls16": sweep,
            "c1_8x8x8x8_ls8_nrep1000": c1,
            "host": host_info()}


def cpu_reference_build(dbl, psi, budget_s=10.0):
    """The reference's own CPU Hand kernel (dslash_kernel_cpu compiled from its
    sources into oracle/_ref, generic backend, Nsimd 1, one thread: the build
    has no OpenMP) on the same inputs, timed by its own clock; None when that
    library was not built in this tree."""
    from oracle import oracle as orc
    V = int(np.prod(DIMS))
    nbr, prm = orc.tables(DIMS, 1, LS)
    U = orc.pack_gauge(DIMS, 1, dbl)
    inp = orc.pack_spinor(DIMS, 1, LS, psi)
    t0 = time.perf_counter()
    us1 = orc.ref_time(U, inp, nbr, prm, V, LS, 1, 2)  # rep 0 untimed, rep 1 timed
    if us1 is None:
        return None
    per = max(time.perf_counter() - t0, 1e-3) / 2
    nrep = max(2, min(20, int(budget_s / per)))
    us = orc.ref_time(U, inp, nbr, prm, V, LS, 1, nrep)
    dt = us * 1e-6 / (nrep - 1)
    return {"value": round(1320.0 * V * LS / dt / 1e9, 3), "unit": "GFLOP/s", "cores": 1, "kind": "reference",
            "sample": f"24^4 x Ls16 FP64 1-RHS, the reference's dslash_kernel_cpu (GEN backend, Nsimd 1, "
                      f"-O2 -ffp-contract=off), nrep={nrep}, its timer over reps 1..{nrep - 1}",
            "ms_per_application": round(dt * 1e3, 2)}


def host_info():
    """The host the CPU baseline ran on."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "logical_cpus": os.cpu_count(),
            "omp": {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OMP_PLACES", "OMP_PROC_BIND")}}


def rel_err_device(gbx, geom, got, want64):
    """|got - want| / |want| over every right-hand side, on the device (got in
    any precision, want64 FP64 of the same shape)."""
    d = gbx.SpinorField(geom, want64.Ls, want64.nrhs, "f64").copy_from(got)
    d.axpy(-1.0, want64)
    e = float(np.sqrt(d.norm2() / want64.norm2()))
    d.close()
    return e


class Tuner:
    def __init__(self, gbx, cache, retune):
        self.gbx, self.cache = gbx, cache
        if retune and cache and os.path.exists(cache):
            os.remove(cache)

    def __call__(self, U, y, x, dag=False, kinds=None):
        return self.gbx.tune(U, y, x, dag=dag, cache_path=self.cache, kinds=kinds)


def variants_section(gbx, tune, geom, dbl, U64, psi, ref64, ref64_dag, us_f64, only=None, compare_forms=True,
                     geom_h=None, U64h=None, geom16=None):
    V = int(np.prod(DIMS))
    out = []
    c4 = {}  # the 24^4 x Ls12 x 16-RHS configuration: its own gauge field (seed 401), spinors seed 402

    def variant(name, prec, nrhs, Ls=LS, dag=False, comp12=False, compare=False, tdiag=False):
        if only and name not in only:
            return
        # FP64 1-RHS rows on the headline's geometry, 16-bit 16-RHS rows on
        # MULTI16_BLOCK, the others on the default one
        fp64_row = prec == "f64" and nrhs == 1 and Ls == LS and geom_h is not None
        multi16_row = prec in ("f16", "bf16") and nrhs == 16 and geom16 is not None
        g = geom_h if fp64_row else geom16 if multi16_row else geom
        block = FP64_BLOCK if fp64_row else MULTI16_BLOCK if multi16_row else None
        sites5 = V * Ls
        if Ls == LS:
            gauge, Uref = dbl, (U64h if fp64_row else U64)
        else:
            if "dbl" not in c4:
                c4["dbl"] = gbx.gen_doubled_gauge(DIMS, C4_GAUGE_SEED)
            if id(g) not in c4:  # the FP64 reference field on this row's geometry
                c4[id(g)] = gbx.GaugeField(g, "f64").upload(c4["dbl"])
            gauge, Uref = c4["dbl"], c4[id(g)]
        U = Uref if (prec == "f64" and not comp12) else gbx.GaugeField(g, prec, comp12).upload(gauge)
        x = gbx.SpinorField(g, Ls, nrhs, prec)
        if Ls == LS:
            x.fill_gaussian(SPINOR_SEED + 1000)  # right-hand sides 1.. ; 0 is the headline spinor
            x.upload_rhs(0, psi)
            want = ref64_dag if dag else ref64
        else:
            x.fill_gaussian(C4_SPINOR_SEED)
            x64 = gbx.SpinorField(g, Ls, nrhs, "f64").copy_from(x)
            w = gbx.SpinorField(g, Ls, nrhs, "f64")
            gbx.dhop(Uref, w, x64, dag=dag)
        if tdiag:  # the spin basis with diagonal gamma_t, set before any timing
            x.set_basis(gbx.BASIS_TDIAG)
        y = gbx.SpinorField(g, Ls, nrhs, prec)
        cfg, _ = tune(U, y, x, dag=dag)
        us = gbx.dhop_time(U, y, x, dag=dag, nrep=100)
        gbx.dhop(U, y, x, dag=dag)
        if tdiag:
            y.set_basis(gbx.BASIS_CHIRAL)
        if Ls == LS:
            got = y.download_rhs(0)
            err = float(np.linalg.norm((got - want).ravel()) / np.linalg.norm(want.ravel()))
            err_what = "rhs 0 vs FP64 kernel on the headline spinor"
        else:
            err = rel_err_device(gbx, g, y, w)
            err_what = "all rhs vs FP64 kernel on the same (rounded) input"
            for f in (x64, w):
                f.close()
        forms = None
        if compare and compare_forms:  # each kernel form (VALU, matrix-core, LDS-tiled), autotuned on its own
            forms = {}
            for kind in ("valu", "mfma", "tile"):
                try:
                    kc, _ = tune(U, y, x, dag=dag, kinds=kind)
                    forms[kind] = {"us_per_launch": round(gbx.dhop_time(U, y, x, dag=dag, cfg=kc, nrep=100), 2),
                                   "cfg": kc.as_dict()}
                except gbx.GbxError as e:
                    forms[kind] = {"unsupported": str(e)}
        b = gbx.bytes_per_site(prec, prec, Ls, nrhs, comp12) * sites5 * nrhs
        v = {"name": name, "prec": prec, "nrhs": nrhs, "Ls": Ls, "dag": dag, "comp12": comp12,
             "spin_basis": "tdiag" if tdiag else "chiral",
             "us_per_launch": round(us, 2), "gflops": round(1320.0 * sites5 * nrhs / (us * 1e-6) / 1e9, 1),
             "hbm_gbps_alg": round(b / (us * 1e-6) / 1e9, 1),
             "per_rhs_speedup_vs_f64_1rhs": round(us_f64 * (Ls / LS) / (us / nrhs), 3),
             "rel_l2_vs_f64": float(f"{err:.3e}"), "error_of": err_what, "cfg": cfg.as_dict(),
             "traversal_block": list(block) if block else "default (6^4)"}
        if forms:
            v["forms"] = forms
        log(f"variant {v}")
        out.append(v)
        for f in (x, y):
            f.close()
        if U is not Uref:
            U.close()

    variant("f64_1rhs_dag", "f64", 1, dag=True, tdiag=True)
    variant("f64_1rhs_comp12", "f64", 1, comp12=True, tdiag=True)
    variant("f32_1rhs", "f32", 1, tdiag=True)
    variant("f32_1rhs_chiral", "f32", 1)
    variant("f32_16rhs", "f32", 16, compare=True, tdiag=True)
    variant("f32_16rhs_chiral", "f32", 16)
    variant("f16_16rhs", "f16", 16, compare=True)
    variant("bf16_16rhs", "bf16", 16, compare=True)
    variant("f16_16rhs_comp12", "f16", 16, comp12=True)
    variant("f32_16rhs_ls12", "f32", 16, Ls=12, compare=True, tdiag=True)
    variant("bf16_16rhs_ls12", "bf16", 16, Ls=12, compare=True)
    return out


def mixed_section(gbx, tune):
    """32^4 x Ls16: Dhop error of each storage format against FP64."""
    dims, Ls = (32, 32, 32, 32), 16
    geom = gbx.Geometry(dims)
    dbl = gbx.gen_doubled_gauge(dims, 301)
    U64 = gbx.GaugeField(geom, "f64").upload(dbl)
    x64 = gbx.SpinorField(geom, Ls, 1, "f64").fill_gaussian(302)
    y64 = gbx.SpinorField(geom, Ls, 1, "f64")
    tune(U64, y64, x64)
    us64 = gbx.dhop_time(U64, y64, x64, nrep=20)
    gbx.dhop(U64, y64, x64)
    rows = [{"prec": "f64", "comp12": False, "us_per_launch": round(us64, 2), "rel_l2_vs_f64": 0.0}]
    for prec, comp12 in (("f32", False), ("bf16", False), ("f16", False), ("f16", True), ("f32", True)):
        U = gbx.GaugeField(geom, prec, comp12).upload(dbl)
        x = gbx.SpinorField(geom, Ls, 1, prec).copy_from(x64)
        y = gbx.SpinorField(geom, Ls, 1, prec)
        tune(U, y, x)
        us = gbx.dhop_time(U, y, x, nrep=20)
        gbx.dhop(U, y, x)
        e_store = rel_err_device(gbx, geom, x, x64)
        e = rel_err_device(gbx, geom, y, y64)
        r = {"prec": prec, "comp12": comp12, "us_per_launch": round(us, 2), "rel_l2_vs_f64": float(f"{e:.3e}"),
             "input_rounding_rel_l2": float(f"{e_store:.3e}")}
        log(f"mixed {r}")
        rows.append(r)
        for f in (U, x, y):
            f.close()
    return {"dims": list(dims), "Ls": Ls, "gauge_seed": 301, "spinor_seed_device": 302, "rows": rows}


def cg_section(gbx, tune):
    """32^4 x Ls16 defect-correction CG to 1e-10."""
    dims, Ls = (32, 32, 32, 32), 16
    geom = gbx.Geometry(dims)
    dbl = gbx.gen_doubled_gauge(dims, 301)
    U64 = gbx.GaugeField(geom, "f64").upload(dbl)
    b = gbx.SpinorField(geom, Ls, 1, "f64").fill_gaussian(303)
    x = gbx.SpinorField(geom, Ls, 1, "f64")
    y = gbx.SpinorField(geom, Ls, 1, "f64")
    tune(U64, y, b)
    rows = []
    for lo, comp12 in (("f32", False), ("bf16", False), ("f16", False), ("f16", True), (None, False)):
        Ulo = None
        if lo is not None:
            Ulo = gbx.GaugeField(geom, lo, comp12).upload(dbl)
            xs = gbx.SpinorField(geom, Ls, 1, lo).copy_from(b)
            ys = gbx.SpinorField(geom, Ls, 1, lo)
            tune(Ulo, ys, xs)
            xs.close()
            ys.close()
        x.zero()
        st = gbx.cg_solve(U64, Ulo, x, b, KAPPA, tol=1e-10, inner_tol=1e-3,
                          max_outer=60, max_inner=5000)
        # the residual once more, independently of the solver: r = b - M^dag M x
        t = gbx.SpinorField(geom, Ls, 1, "f64")
        gbx.dhop_axpby(U64, t, x, 1.0, -KAPPA)
        gbx.dhop_axpby(U64, y, t, 1.0, -KAPPA, dag=True)
        y.axpy(-1.0, b)
        rel = float(np.sqrt(y.norm2() / b.norm2()))
        t.close()
        r = {"inner": lo or "f64 (no defect correction)", "comp12": comp12, **st,
             "recomputed_rel_resid": float(f"{rel:.3e}"), "converged": rel <= 1e-10}
        log(f"cg {r}")
        rows.append(r)
        if Ulo is not None:
            Ulo.close()
    return {"dims": list(dims), "Ls": Ls, "kappa": KAPPA, "tol": 1e-10, "gauge_seed": 301,
            "rhs_seed_device": 303, "rows": rows}


def cg_eo_section(gbx, tune):
    """32^4 x Ls16: M x = b by the even-odd Schur complement (defect-correction
    CGNR on the even sites) against CG on the normal equations M^dag M x =
    M^dag b over the whole lattice, same b, FP32 inner solves for the latter."""
    dims, Ls = (32, 32, 32, 32), 16
    dbl = gbx.gen_doubled_gauge(dims, 301)
    b_lex = gbx.gen_spinor(dims, Ls, 1, 304)
    be, bo = gbx.rb_split(dims, b_lex)
    del b_lex
    ge, go = gbx.Geometry.rb_pair(dims)
    Ue = gbx.GaugeField(ge, "f64").upload(dbl)
    Uo = gbx.GaugeField(go, "f64").upload(dbl)
    b = (gbx.SpinorField(ge, Ls, 1, "f64").upload(be), gbx.SpinorField(go, Ls, 1, "f64").upload(bo))
    x = (gbx.SpinorField(ge, Ls, 1, "f64"), gbx.SpinorField(go, Ls, 1, "f64"))
    for U, xin, yout in ((Ue, x[1], x[0]), (Uo, x[0], x[1])):
        tune(U, yout, xin)
    rows = []
    for lo in ("f32", "f16", "bf16", None):
        low = None
        if lo is not None:
            low = (gbx.GaugeField(ge, lo).upload(dbl), gbx.GaugeField(go, lo).upload(dbl))
            ts = (gbx.SpinorField(ge, Ls, 1, lo), gbx.SpinorField(go, Ls, 1, lo))
            tune(low[0], ts[0], ts[1])
            tune(low[1], ts[1], ts[0])
        for f in x:
            f.zero()
        st = gbx.cg_solve_eo((Ue, Uo), low, x, b, KAPPA, tol=1e-10, inner_tol=1e-3, max_outer=60, max_inner=5000)
        r = {"inner": lo or "f64", **st, "converged": st["true_resid"] <= 1e-10}
        log(f"cg_eo {r}")
        rows.append(r)
    for f in (Ue, Uo, x[0], x[1], b[0], b[1]):
        f.close()
    # the same system on the whole lattice through the normal equations
    g = gbx.Geometry(dims)
    U = gbx.GaugeField(g, "f64").upload(dbl)
    U32 = gbx.GaugeField(g, "f32").upload(dbl)
    bf = gbx.SpinorField(g, Ls, 1, "f64").upload(gbx.rb_merge(dims, be, bo))
    del be, bo
    mdb = gbx.SpinorField(g, Ls, 1, "f64")
    xf = gbx.SpinorField(g, Ls, 1, "f64").zero()
    tune(U, mdb, bf)
    gbx.dhop_axpby(U, mdb, bf, 1.0, -KAPPA, dag=True)
    st = gbx.cg_solve(U, U32, xf, mdb, KAPPA, tol=1e-10, inner_tol=1e-3, max_outer=60, max_inner=5000)
    # |b - M x| / |b| of that solution, like the even-odd solver reports
    t = gbx.SpinorField(g, Ls, 1, "f64")
    gbx.dhop_axpby(U, t, xf, 1.0, -KAPPA)
    t.axpy(-1.0, bf)
    full = {"inner": "f32", **st, "m_resid": float(f"{np.sqrt(t.norm2() / bf.norm2()):.3e}")}
    log(f"cg normal equations {full}")
    eo32 = rows[0]
    return {"dims": list(dims), "Ls": Ls, "kappa": KAPPA, "tol": 1e-10, "gauge_seed": 301, "b_seed": 304,
            "system": "M x = b, M = 1 - kappa Dhop", "even_odd": rows, "normal_equations_whole_lattice": full,
            "speedup_even_odd_f32_vs_normal_f32": round(full["seconds"] / eo32["seconds"], 2)}


def dwf_section(gbx, tune):
    """24^4 x Ls16 domain-wall operator D = Mooee - Dhop / 2 (the reference's
    DomainWallFermion parameters mass 0.1, M5 1.0): D x = b by the even-odd
    Schur complement, FP64 outer, FP32 / FP16 inner."""
    Ls = LS
    dbl = gbx.gen_doubled_gauge(DIMS, GAUGE_SEED)
    b_lex = gbx.gen_spinor(DIMS, Ls, 1, 501)
    be, bo = gbx.rb_split(DIMS, b_lex)
    del b_lex
    ge, go = gbx.Geometry.rb_pair(DIMS)
    U = (gbx.GaugeField(ge, "f64").upload(dbl), gbx.GaugeField(go, "f64").upload(dbl))
    b = (gbx.SpinorField(ge, Ls, 1, "f64").upload(be), gbx.SpinorField(go, Ls, 1, "f64").upload(bo))
    x = (gbx.SpinorField(ge, Ls, 1, "f64"), gbx.SpinorField(go, Ls, 1, "f64"))
    tune(U[0], x[0], x[1])
    tune(U[1], x[1], x[0])
    rows = []
    for lo in ("f32", "f16"):
        low = (gbx.GaugeField(ge, lo).upload(dbl), gbx.GaugeField(go, lo).upload(dbl))
        ts = (gbx.SpinorField(ge, Ls, 1, lo), gbx.SpinorField(go, Ls, 1, lo))
        tune(low[0], ts[0], ts[1])
        tune(low[1], ts[1], ts[0])
        for f in x:
            f.zero()
        st = gbx.dwf_solve_eo(U, low, x, b, tol=1e-10, max_outer=60)
        r = {"inner": lo, **st, "converged": st["true_resid"] <= 1e-10}
        log(f"dwf_eo {r}")
        rows.append(r)
    return {"dims": list(DIMS), "Ls": Ls, "M5": gbx.DWF_M5, "mass": gbx.DWF_MASS, "gauge_seed": GAUGE_SEED,
            "b_seed": 501, "system": "D x = b, D = Mooee - Dhop / 2 (Shamir)", "even_odd": rows}


def batch_section(gbx, nrep, shape=None):
    """40 uneven lattices in one persistent graph-captured launch, FP32, on
    fields in the TDIAG spin basis (errors are basis-independent norms).
    shape = (threads, pf, wgpc, points) fixes the launch shape instead of
    tuning it (for profiling)."""
    rng = np.random.default_rng(505)
    specs = [(int(rng.choice([8, 12, 16, 20, 24])), int(rng.choice([8, 12, 16])), int(rng.integers(1, 9)))
             for _ in range(40)]
    lats = []
    for i, (L, Ls, nrhs) in enumerate(specs):
        dims = (L, L, L, L)
        g = gbx.Geometry(dims)
        dbl = gbx.gen_doubled_gauge(dims, 5050 + i)
        U = gbx.GaugeField(g, "f32").upload(dbl)
        x = gbx.SpinorField(g, Ls, nrhs, "f32").fill_gaussian(6060 + i).set_basis(gbx.BASIS_TDIAG)
        y = gbx.SpinorField(g, Ls, nrhs, "f32")
        lats.append((g, dbl, U, x, y))
    b = gbx.Batch([l[2] for l in lats], [l[4] for l in lats], [l[3] for l in lats])
    us_default = b.time(3)
    if shape:
        b.set_cfg(gbx.kcfg(shape[3], shape[0], 0, 1, pf=shape[1], wgpc=shape[2]))
        bcfg = b.get_cfg()
    else:
        bcfg, _ = b.tune(nrep=3)  # threads, gather depth, workgroups per CU, points per item
    items, wgs = b.info()
    b.run()
    us = b.time(nrep)
    flop = sum(1320.0 * L ** 4 * Ls * nrhs for L, Ls, nrhs in specs)
    byts = sum(gbx.bytes_per_site("f32", "f32", Ls, nrhs) * L ** 4 * Ls * nrhs for L, Ls, nrhs in specs)
    errs = []
    for (g, dbl, U, x, y), (L, Ls, nrhs) in zip(lats, specs):
        U64 = gbx.GaugeField(g, "f64").upload(dbl)
        x64 = gbx.SpinorField(g, Ls, nrhs, "f64").copy_from(x)
        y64 = gbx.SpinorField(g, Ls, nrhs, "f64")
        gbx.dhop(U64, y64, x64)
        errs.append(float(f"{rel_err_device(gbx, g, y, y64):.3e}"))
        for f in (U64, x64, y64):
            f.close()
    b.close()
    r = {"lattices": [{"L": L, "Ls": Ls, "nrhs": n} for L, Ls, n in specs], "spin_basis": "tdiag", "work_items": items,
         "persistent_workgroups": wgs, "launch_shape": {k: v for k, v in bcfg.as_dict().items()
                                                         if k in ("threads", "pf", "wgpc", "nt", "spw")},
         "us_per_launch_untuned_shape": round(us_default, 2),
         "us_per_launch": round(us, 2), "gflops": round(flop / (us * 1e-6) / 1e9, 1),
         "hbm_gbps_alg": round(byts / (us * 1e-6) / 1e9, 1), "rel_l2_vs_f64_per_lattice": errs,
         "max_rel_l2_vs_f64": max(errs)}
    log(f"batch: {us:.1f} us, {r['gflops']} GFLOP/s, max err {max(errs):.2e}")
    return r


def dropin_section(gbx, dbl, psi, ref64, nrep):
    """The reference entry itself: dslash_kernel (gbx_dslash_kernel_f64 /
    _f32, WilsonKernelsHand.h:17-29) called as bench.cc calls it, on host
    arrays in the reference layout (Nsimd 1, where it is the lexicographic
    layout) with TableGenerate's tables, 24^4 x Ls16.  Reported: the
    entry's own device time of applications 1..nrep-1 (the reference's
    timing convention) with the reference's rounding, and the wall time of
    the whole call (table check, both PCIe copies, repacking)."""
    V = int(np.prod(DIMS))
    nbr, prm = gbx.table_generate(DIMS, 1, LS)
    rows = []
    for dt, name in ((np.float64, "f64"), (np.float32, "f32")):
        U = np.ascontiguousarray(dbl, dtype=dt).reshape(-1)
        inp = np.ascontiguousarray(psi, dtype=dt).reshape(-1)
        out = np.zeros_like(inp)
        gbx.dslash_kernel(2, U, out, inp, nbr, V, LS, prm, 1)  # device buffers allocated once
        t0 = time.perf_counter()
        us = gbx.dslash_kernel(nrep, U, out, inp, nbr, V, LS, prm, 1)
        wall = time.perf_counter() - t0
        path = gbx.dslash_last_path()
        per = us / (nrep - 1)
        err = float(np.linalg.norm(out.astype(np.float64) - ref64.reshape(-1)) / np.linalg.norm(ref64))
        r = {"prec": name, "nsimd": 1, "nrep": nrep, "path": "SoA kernel" if path == 1 else "table-driven kernel",
             "us_per_application_device": round(per, 2),
             "gflops_device": round(1320.0 * V * LS / (per * 1e-6) / 1e9, 1),
             "wall_s_whole_call": round(wall, 3),
             "rel_l2_vs_device_resident_f64": float(f"{err:.3e}")}
        log(f"dropin {r}")
        rows.append(r)
        del U, inp, out
    return {"dims": list(DIMS), "Ls": LS, "rows": rows,
            "note": "device time counts applications 1..nrep-1 of the reference-rounding kernel; the wall time "
                    "includes the PCIe copies of U, in, the tables and out"}


XGMI_GBPS = 153.6  # one MI355X xGMI link, AMD-published per-link figure (direction split not stated)


def host_timed(fn, geom, nrep):
    """Microseconds per call from the host clock around nrep enqueued calls
    (every call's device work is far longer than its launch)."""
    fn()
    geom.sync()
    t0 = time.perf_counter()
    for _ in range(nrep):
        fn()
    geom.sync()
    return (time.perf_counter() - t0) * 1e6 / nrep


def halo_section(gbx, tune, dbl, psi, ref64, nrep):
    """One GPU's share of a domain-decomposed 24^4 x Ls16 FP64 Dhop, for
    process grids 1x1x1x2 and 1x1x1x4 (slabs along T) and 1x2x2x2 (one
    8-GPU node).  Every block is built and its faces exchanged on this GPU
    (checks the assembled result against the whole-lattice application);
    block 0 is timed: Dhop on its whole share, on the interior slots alone,
    on the face slots alone, the face pack / unpack kernels, and one rank's
    whole schedule with the faces moved on this GPU.  The xGMI transfer
    itself needs several GPUs: its time is the face bytes at one link's
    published rate (a model, labelled as such)."""
    rows = []
    for grid in ((1, 1, 1, 2), (1, 1, 1, 4), (1, 2, 2, 2)):
        coords = gbx.grid_coords(grid)
        geoms = {c: gbx.Geometry.part_of(DIMS, grid, c) for c in coords}
        loc = gbx.part_split(DIMS, grid, psi[None])
        U = {c: gbx.GaugeField(g, "f64").upload(dbl) for c, g in geoms.items()}
        x = {c: gbx.SpinorField(g, LS, 1, "f64").upload(loc[c]) for c, g in geoms.items()}
        y = {c: gbx.SpinorField(g, LS, 1, "f64") for c, g in geoms.items()}
        del loc
        c0 = coords[0]
        g0, u0, x0, y0 = geoms[c0], U[c0], x[c0], y[c0]
        best, _ = tune(u0, y0, x0)
        bufs = gbx.halo_buffers(x)
        # projected faces packed straight into the neighbours' half-slots,
        # interior slots, then the face slots reading the halves as they are
        gbx.halo_exchange_in_place(x)
        for c in coords:
            gbx.dhop_range(U[c], y[c], x[c], 0, geoms[c].ninterior)
            gbx.dhop_faces_projected(U[c], y[c], x[c])
        got = gbx.part_merge(DIMS, grid, {c: y[c].download() for c in coords})[0]
        err = float(np.linalg.norm(got - ref64) / np.linalg.norm(ref64))
        del got
        cut = [d for d in range(4) if grid[d] > 1]
        us_part = gbx.dhop_time(u0, y0, x0, nrep=nrep)
        nint, vol = g0.ninterior, g0.volume
        us_int = host_timed(lambda: gbx.dhop_range(u0, y0, x0, 0, nint), g0, nrep)
        us_bnd = host_timed(lambda: gbx.dhop_range(u0, y0, x0, nint, vol), g0, nrep)
        us_bnd_proj = host_timed(lambda: gbx.dhop_faces_projected(u0, y0, x0), g0, nrep)

        # every face in one launch; whole faces are packed into y0's own
        # ghosts (a scratch target of the right size)
        scratch = {(d, f): y0.halo_ghost(d, f) for d in cut for f in (0, 1)}
        sendb = {(d, f): bufs[(c0, d)][f].ptr for d in cut for f in (0, 1)}
        recvb = {(d, s): bufs[(c0, d)][1 - s].ptr for d in cut for s in (0, 1)}
        us_pack = host_timed(lambda: x0.halo_pack_all(scratch), g0, nrep)
        us_pack_proj = host_timed(lambda: x0.halo_pack_all(sendb, proj=True), g0, nrep)
        us_unpack_proj = host_timed(lambda: x0.halo_unpack_all(recvb, proj=True), g0, nrep)
        # one rank's schedule on this GPU (faces moved by a device copy)
        us_ov, us_se = gbx.halo_overlap_time(u0, y0, x0, proj=True, nrep=nrep)
        us_ov_w, us_se_w = gbx.halo_overlap_time(u0, y0, x0, proj=False, nrep=nrep)
        # Per application a block sends two faces per cut dimension.  With two
        # blocks along a dimension both go to one neighbour (one link), with
        # more to two neighbours; different neighbours use different links,
        # which run at once.
        xfer_us = max((2 if grid[d] == 2 else 1) * x0.halo_bytes(d, True) for d in cut) / (XGMI_GBPS * 1e9) * 1e6
        # The pack kernel needs CUs, which the interior kernel holds while it
        # runs (the measured single-GPU schedule shows no overlap of kernels
        # on two streams): only the transfer, on copy engines, overlaps the
        # interior.  The faces land where the face launch reads them.
        overlapped = us_pack_proj + max(us_int, xfer_us) + us_bnd_proj
        r = {"grid": list(grid), "local_dims": list(g0.dims), "local_sites": vol, "interior_sites": nint,
             "face_sites": {"xyzt"[d]: g0.face[d] for d in cut}, "kernel_cfg": best.as_dict(),
             "rel_l2_assembled_vs_whole_lattice": float(f"{err:.3e}"),
             "us_dhop_local_share_events": round(us_part, 2), "us_dhop_interior_host": round(us_int, 2),
             "us_dhop_faces_host": round(us_bnd, 2), "us_dhop_faces_projected_ghosts_host": round(us_bnd_proj, 2),
             "face_bytes_full": {"xyzt"[d]: x0.halo_bytes(d) for d in cut},
             "face_bytes_projected": {"xyzt"[d]: x0.halo_bytes(d, True) for d in cut},
             "us_pack_all_faces_full": round(us_pack, 2), "us_pack_all_faces_projected": round(us_pack_proj, 2),
             "us_unpack_all_faces_projected": round(us_unpack_proj, 2),
             "us_application_serial_local_move": round(us_se, 2),
             "us_application_overlapped_local_move": round(us_ov, 2),
             "us_application_serial_local_move_whole_faces": round(us_se_w, 2),
             "us_application_overlapped_local_move_whole_faces": round(us_ov_w, 2),
             "model_xgmi_gbps_per_link": XGMI_GBPS, "model_us_xgmi": round(xfer_us, 1),
             "model_us_per_application": round(overlapped, 1),
             "model": "pack + max(interior, xGMI) + face slots from projected ghosts",
             "rel_l2_check": "assembled from interior ranges and projected-ghost face launches",
             "gflops_local_share": round(1320.0 * vol * LS / (us_part * 1e-6) / 1e9, 1)}
        log(f"halo {r}")
        rows.append(r)
        for pair in bufs.values():
            pair[0].close()
            pair[1].close()
        for f in (*y.values(), *x.values(), *U.values()):
            f.close()
    return {"dims": list(DIMS), "Ls": LS, "prec": "f64",
            "note": "one GPU holds every block; the xGMI time is the model bytes / link rate, not a measurement",
            "grids": rows}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000, help="timed applications (the reference's nrep)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-variants", action="store_true", help="only the FP64 1-RHS headline")
    ap.add_argument("--no-mixed", action="store_true", help="skip the 32^4 mixed-precision errors")
    ap.add_argument("--no-cg", action="store_true", help="skip the defect-correction CG")
    ap.add_argument("--no-batch", action="store_true", help="skip the uneven batch")
    ap.add_argument("--no-halo", action="store_true", help="skip the domain-part halo section")
    ap.add_argument("--no-dropin", action="store_true", help="skip the reference-entry (dslash_kernel) section")
    ap.add_argument("--headline-only", action="store_true",
                    help="= --no-variants --no-mixed --no-cg --no-batch --no-halo --no-dropin")
    ap.add_argument("--only", default="", help="comma-separated variant names to run (profiling)")
    ap.add_argument("--no-compare", action="store_true", help="skip the MFMA-vs-VALU form comparison (profiling)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--headline-kinds", default="",
                    help="comma-separated kernel kinds the headline's tuner may use (valu, tile, march); profiling")
    ap.add_argument("--batch-shape", default="",
                    help="threads,pf,wgpc,points: the batch's launch shape, not tuned (profiling)")
    ap.add_argument("--tune-cache", default=TUNE_CACHE, help="autotuner cache (JSON lines); '' = tune afresh")
    ap.add_argument("--retune", action="store_true", help="delete the cache file first")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (the hopping term on one "
              "GPU's lattice is the whole measured path; one GPU's slab of a domain decomposition, with its "
              "face pack / unpack, is the 'halo' section; transfers between GPUs are not run).", file=sys.stderr)
        sys.exit(2)
    if args.headline_only:
        args.no_variants = args.no_mixed = args.no_cg = args.no_batch = args.no_halo = args.no_dropin = True

    import gridbench_amd as gbx
    if gbx.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(3)
    tune = Tuner(gbx, args.tune_cache or None, args.retune)

    V = int(np.prod(DIMS))
    sites5 = V * LS
    log("generating inputs")
    dbl = gbx.gen_doubled_gauge(DIMS, GAUGE_SEED)
    psi = gbx.gen_spinor(DIMS, LS, 1, SPINOR_SEED)[0]

    geom = gbx.Geometry(DIMS)
    U64 = gbx.GaugeField(geom, "f64").upload(dbl)
    x64 = gbx.SpinorField(geom, LS, 1, "f64").upload_rhs(0, psi)
    y64 = gbx.SpinorField(geom, LS, 1, "f64")
    # The headline field: the same spinor on the FP64_BLOCK traversal order,
    # stored in the spin basis with diagonal gamma_t (converted once, before
    # the timed region), so that the kernel's +-t legs gather half a spinor.
    # Its result, set back to the chiral basis, is checked against the chiral
    # kernel's on the default geometry below.
    geom_h = gbx.Geometry(DIMS, block=FP64_BLOCK)
    U64h = gbx.GaugeField(geom_h, "f64").upload(dbl)
    xh = gbx.SpinorField(geom_h, LS, 1, "f64").upload_rhs(0, psi).set_basis(gbx.BASIS_TDIAG)
    yh = gbx.SpinorField(geom_h, LS, 1, "f64")

    if args.headline_kinds:  # a restricted search is not adopted by gbx_dhop: pass its result explicitly
        best, best_us = tune(U64h, yh, xh, kinds=args.headline_kinds.split(","))
        hcfg = best
    else:
        best, best_us = tune(U64h, yh, xh)
        hcfg = None
    log(f"FP64 1-RHS config: {best.as_dict()} {best_us:.1f} us")

    # ------------------------------------------------ timed headline region
    for _ in range(args.warmup):
        gbx.dhop(U64h, yh, xh, cfg=hcfg)
    geom_h.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        gbx.dhop(U64h, yh, xh, cfg=hcfg)
    geom_h.sync()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / args.steps
    gflops = 1320.0 * sites5 / (ms * 1e-3) / 1e9
    # kernel duration from HIP events on the kernel's own stream
    us_ev = gbx.dhop_time(U64h, yh, xh, cfg=hcfg, nrep=args.steps)
    bytes_alg = gbx.bytes_per_site("f64", "f64", LS, 1) * sites5
    achieved = bytes_alg / (us_ev * 1e-6) / 1e9
    log(f"FP64 1-RHS: {ms:.3f} ms/step wall, {us_ev:.1f} us/launch events, {gflops:.1f} GFLOP/s, {achieved:.0f} GB/s")
    gbx.dhop(U64h, yh, xh, cfg=hcfg)
    yh.set_basis(gbx.BASIS_CHIRAL)
    got_h = yh.download_rhs(0)
    # the same application on the chiral-basis field of the same geometry
    xh.set_basis(gbx.BASIS_CHIRAL)
    best_chiral, _ = tune(U64h, yh, xh)
    us_chiral = gbx.dhop_time(U64h, yh, xh, nrep=min(args.steps, 200))
    for f in (xh, yh):
        f.close()
    # the chiral kernel on the default geometry: reference for every later section
    gbx.dhop(U64, y64, x64)
    ref64 = y64.download_rhs(0)
    err_h = float(np.linalg.norm((got_h - ref64).ravel()) / np.linalg.norm(ref64.ravel()))
    log(f"FP64 1-RHS chiral-basis field: {us_chiral:.1f} us/launch; headline result vs it: rel L2 {err_h:.2e}")

    stream_gbps = gbx.stream_copy_bw(4 << 30, 20)
    read3_gbps = gbx.stream_bw(6 << 30, 3, 20)
    read0_gbps = gbx.stream_bw(4 << 30, 0, 20)
    gather_gbps = gbx.gather_bw(2 << 30, 10)  # 256-byte runs at random places: the gathers' shape, from HBM
    gather_l2_gbps = gbx.gather_bw(2 << 20, 10)  # the same from a buffer every XCD's L2 holds (not HBM)
    achievable = max(stream_gbps, read3_gbps, read0_gbps, gather_gbps)
    log(f"HBM probes: copy {stream_gbps:.0f}, 3:1 read:write {read3_gbps:.0f}, read-only {read0_gbps:.0f}, "
        f"gathers of 256 B runs {gather_gbps:.0f} GB/s (from L2: {gather_l2_gbps:.0f} GB/s)")

    extra = {}
    if not args.no_variants:
        gbx.dhop(U64, y64, x64, dag=True)
        ref64_dag = y64.download_rhs(0)
        extra["variants"] = variants_section(gbx, tune, geom, dbl, U64, psi, ref64, ref64_dag, us_ev,
                                               only=set(filter(None, args.only.split(","))),
                                               compare_forms=not args.no_compare, geom_h=geom_h, U64h=U64h,
                                               geom16=gbx.Geometry(DIMS, block=MULTI16_BLOCK))
    if not args.no_mixed:
        extra["mixed_precision"] = mixed_section(gbx, tune)
    if not args.no_cg:
        extra["cg"] = cg_section(gbx, tune)
        extra["cg_even_odd"] = cg_eo_section(gbx, tune)
        extra["dwf"] = dwf_section(gbx, tune)
    if not args.no_batch:
        shape = tuple(int(v) for v in args.batch_shape.split(",")) if args.batch_shape else None
        extra["batch"] = batch_section(gbx, 20, shape)
    if not args.no_halo:
        extra["halo"] = halo_section(gbx, tune, dbl, psi, ref64, 100)
    if not args.no_dropin:
        extra["dropin"] = dropin_section(gbx, dbl, psi, ref64, 20)

    cpu = ref_cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        log(f"CPU baseline with {threads} threads")
        cpu = cpu_baseline(dbl, psi, threads)
        log(f"CPU baseline: {cpu}")
        ref_cpu = cpu_reference_build(dbl, psi)
        log(f"reference build on the CPU: {ref_cpu}")

    # the multi-RHS rows of the metric, compact (the full rows are in "variants")
    rows16 = {}
    for v in extra.get("variants", []):
        if v["name"] in ("f32_16rhs", "f16_16rhs"):
            t = traffic_from_profiles(v["name"], v["cfg"])
            alg = gbx.bytes_per_site(v["prec"], v["prec"], v["Ls"], v["nrhs"], v["comp12"]) * sites5 * v["nrhs"]
            rows16[v["name"]] = {"us": v["us_per_launch"], "gflops": v["gflops"], "hbm_gbps_alg": v["hbm_gbps_alg"],
                                 "frac": round(v["hbm_gbps_alg"] / HBM_PEAK_GBPS, 4),
                                 "per_rhs_speedup": v["per_rhs_speedup_vs_f64_1rhs"], "rel_l2": v["rel_l2_vs_f64"],
                                 "traffic_ratio": round(t / alg, 3) if t else None,
                                 "kind": {0: "valu", 1: "mfma", 2: "tile", 3: "march"}.get(v["cfg"]["kind"], "?"),
                                 "traversal_block": v["traversal_block"]}
    dropin_us = None
    for r in extra.get("dropin", {}).get("rows", []):
        if r["prec"] == "f64":
            dropin_us = r["us_per_application_device"]
    cfg_extra = {"chiral_basis_us_per_launch": round(us_chiral, 2),
                 "chiral_basis_gflops": round(1320.0 * sites5 / (us_chiral * 1e-6) / 1e9, 1),
                 "dropin_f64_us_per_application": dropin_us}
    roof_extra = {}
    for name, r in rows16.items():
        for k in ("us", "gflops", "per_rhs_speedup", "rel_l2", "kind"):
            cfg_extra[f"{name}_{k}"] = r[k]
        roof_extra[f"{name}_frac"] = r["frac"]
        roof_extra[f"{name}_traffic_ratio"] = r["traffic_ratio"]
    t64 = traffic_from_profiles("f64_1rhs", best.as_dict())
    if t64:  # the memory-side rate the counters imply for this run's launch time, against the largest probe
        roof_extra["traffic_gbps"] = round(t64 / (us_ev * 1e-6) / 1e9, 1)
        roof_extra["traffic_frac_of_achievable"] = round(t64 / (us_ev * 1e-6) / 1e9 / achievable, 3)

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
                   "Ls": LS, "nrhs": 1, "flop_per_5d_site": 1320, "kernel_cfg": best.as_dict(),
                   "spin_basis": "tdiag (gamma_t diagonal; converted before the timed region)",
                   "traversal_block": list(FP64_BLOCK), **cfg_extra},
        "chiral_basis": {"us_per_launch_events": round(us_chiral, 2), "kernel_cfg": best_chiral.as_dict(),
                         "gflops": round(1320.0 * sites5 / (us_chiral * 1e-6) / 1e9, 1),
                         "headline_rel_l2_vs_chiral_kernel": float(f"{err_h:.3e}")},
        "hbm_gbps_alg": round(achieved, 1),
        "achievable_hbm_gbps": {"copy": round(stream_gbps, 1), "read3_write1": round(read3_gbps, 1),
                                "read_only": round(read0_gbps, 1), "gather_256B_runs": round(gather_gbps, 1),
                                "max": round(achievable, 1)},
        "l2_gather_256B_runs_gbps": round(gather_l2_gbps, 1),
        "frac_of_achievable_hbm": round(achieved / achievable, 3),
        "headline_check": {"rel_l2_vs_chiral_kernel": float(f"{err_h:.3e}"), "tol": HEADLINE_CHECK_TOL,
                           "passed": err_h <= HEADLINE_CHECK_TOL},
        "multi_rhs_16": rows16,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": t64,
                     "traffic_ratio": round(t64 / bytes_alg, 3) if t64 else None,
                     "algorithmic_bytes_per_launch": bytes_alg, "us_per_launch_events": round(us_ev, 2),
                     "frac_of_achievable": round(achieved / achievable, 3), **roof_extra},
        "cpu_baseline": cpu,
        "cpu_reference_build": ref_cpu,
        **extra,
    }
    print(json.dumps(result), flush=True)
    if err_h > HEADLINE_CHECK_TOL:
        print(f"bench.py: the headline result differs from the chiral kernel's by {err_h:.3e} (rel L2) "
              f"> {HEADLINE_CHECK_TOL:g}", file=sys.stderr)
        sys.exit(4)


if __name__ == "__main__":
    main()
