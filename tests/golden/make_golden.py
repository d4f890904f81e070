#!/usr/bin/env python3
"""Writes tests/golden/hand_fixtures.json from the REFERENCE kernel.

The outputs come from the reference's own dslash_kernel_cpu
(WilsonKernelsHandCpu.h:385-451) and scalar dslash_simple.h kernel, compiled
from the reference sources by `make -C oracle ref` into oracle/_ref/ (this
script refuses to run without that build).  Inputs are the library's seeded
generator (gbx_gen_su3_links / gbx_gen_spinor) in the reference host layout
(bench.cc:105-138) with tables in the reference's form.  Recorded per case,
mirroring the reference bench's own check (bench.cc:180-199): SHA-256 of the
tables, of the input arrays and of the output bytes, the output's sum of
squares and its first 64 values (as hex floats); for the smallest case the
full output.  The dagger has no reference kernel; its fixtures are
gamma5 * reference(gamma5 psi).

Usage:  python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import gridbench_amd as gbx  # noqa: E402
from oracle import oracle as orc  # noqa: E402

CASES = [
    # dims, Ls, nsimd, dtype, full_output
    ((2, 2, 2, 2), 2, 1, "f64", True),
    ((4, 4, 4, 4), 4, 1, "f64", False),
    ((4, 4, 4, 4), 4, 2, "f64", False),
    ((4, 4, 4, 4), 4, 4, "f64", False),
    ((4, 4, 4, 4), 4, 8, "f64", False),
    ((4, 4, 4, 4), 4, 2, "f32", False),
    ((4, 4, 4, 4), 4, 4, "f32", False),
    ((4, 4, 4, 4), 4, 8, "f32", False),
    ((4, 4, 4, 4), 4, 16, "f32", False),
    ((8, 8, 8, 8), 8, 1, "f64", False),  # the reference's CPU configuration (C1)
    ((8, 8, 8, 8), 8, 2, "f64", False),
    ((4, 6, 2, 8), 3, 1, "f64", False),
]
GAUGE_SEED, SPINOR_SEED = 101, 102


def hexs(a):
    return [float(x).hex() for x in np.asarray(a, dtype=np.float64).reshape(-1)]


def make_inputs(dims, Ls, nsimd, dtype):
    dbl = gbx.double_links(dims, gbx.gen_su3_links(dims, GAUGE_SEED))
    psi = gbx.gen_spinor(dims, Ls, 1, SPINOR_SEED)[0]
    U = orc.aligned_copy(orc.pack_gauge(dims, nsimd, dbl).astype(dtype))
    inp = orc.aligned_copy(orc.pack_spinor(dims, nsimd, Ls, psi).astype(dtype))
    nbr, prm = orc.tables(dims, nsimd, Ls)
    return U, inp, nbr, prm


def main():
    for w in (16, 32, 64, 128):
        if orc.ref_lib(w) is None:
            sys.exit(f"oracle/_ref/libref_w{w}.so missing: run `make -C oracle ref` (needs the reference sources)")
    out = {"generator": {}, "cases": []}
    # generator pins: raw Gaussian stream values and one SU(3) link field
    out["generator"]["gaussian_seed7_stream3_first16"] = hexs(gbx.gen_gaussian(7, 3, 0, 16))
    out["generator"]["gaussian_seed7_stream3_from1000000_8"] = hexs(gbx.gen_gaussian(7, 3, 1000000, 8))
    links = gbx.gen_su3_links((2, 2, 2, 2), GAUGE_SEED)
    out["generator"]["su3_links_2222_seed101_sha256"] = orc.sha256(links)
    out["generator"]["su3_links_2222_seed101_first18"] = hexs(links.reshape(-1)[:18])
    out["generator"]["spinor_4444_ls4_seed102_sha256"] = orc.sha256(gbx.gen_spinor((4, 4, 4, 4), 4, 1, SPINOR_SEED))
    for dims, Ls, nsimd, prec, full in CASES:
        dtype = np.float64 if prec == "f64" else np.float32
        U, inp, nbr, prm = make_inputs(dims, Ls, nsimd, dtype)
        nsite = int(np.prod(dims)) // nsimd
        ref = orc.ref_hand(U, inp, nbr, prm, nsite, Ls, nsimd)
        refdag = orc.gamma5(orc.ref_hand(U, orc.gamma5(inp, nsimd).astype(dtype), nbr, prm, nsite, Ls, nsimd),
                            nsimd).astype(dtype)
        case = {
            "dims": list(dims), "Ls": Ls, "nsimd": nsimd, "prec": prec,
            "gauge_seed": GAUGE_SEED, "spinor_seed": SPINOR_SEED,
            "sha256_nbr": orc.sha256(nbr), "sha256_prm": orc.sha256(prm),
            "sha256_U": orc.sha256(U), "sha256_in": orc.sha256(inp),
            "sha256_out": orc.sha256(ref), "sha256_out_dag": orc.sha256(refdag),
            "norm2_out": float(np.sum(ref.astype(np.float64) ** 2)).hex(),
            "first64_out": hexs(ref[:64]), "first64_out_dag": hexs(refdag[:64]),
        }
        if nsimd == 1 and prec == "f64":
            simple = orc.ref_simple(U, inp, nbr, prm, nsite, Ls)
            case["sha256_out_simple"] = orc.sha256(simple)
        if full:
            case["out"] = hexs(ref)
            case["out_dag"] = hexs(refdag)
        out["cases"].append(case)
        print(dims, Ls, nsimd, prec, case["sha256_out"][:16], flush=True)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hand_fixtures.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
