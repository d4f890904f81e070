"""The benchmark workloads at their full shapes, each against the oracle (the
CPU restatement of the reference Hand kernel, WilsonKernelsHandCpu.h:385-451).
The reference checks every benchmark run against its CPU result
(bench.cc:180-199; bench_simple.cc:114-119 asserts it); these are those
checks for the shapes the bench times, so that an index-width or launch-shape
mistake that only shows at full size cannot pass.

* 24^4 x Ls16 (5.3 M 5D sites): the timed FP64 headline instance itself --
  TDIAG spin basis, 4x6x6x12 traversal blocks, the cached tuned launch
  shape -- Dhop and DhopDag <= 1e-14; FP32 1-RHS (chiral and TDIAG) <= 1e-6.
* 32^4 x Ls16 (16.8 M): FP32, BF16, FP16 and FP16 with 12-real links, each on
  its tuned kernel, against the FP64 oracle of the exact input (FP32 <= 1e-6,
  16-bit <= 2.5 x the input's storage rounding); defect-correction CG on
  M^dag M x = b (M = 1 - 0.12 Dhop) to 1e-10 with each inner precision, the
  residual recomputed by the oracle on the host.
* 24^4 x Ls12 x 16 RHS (63.7 M; 6.1 GB in FP32, past 2^32 bytes): FP32 on the
  VALU and LDS-tiled kernels, BF16 on the VALU and MFMA kernels; right-hand
  sides 0, 7 and 15 against the oracle and every right-hand side against the
  FP64 kernel on the same input (that kernel itself against the oracle).
* The uneven batch of 40 lattices (shapes drawn with seed 505, as the bench
  draws them), FP32 on TDIAG fields, one graph launch: every lattice and
  right-hand side against the oracle.

Seeds are the bench's: gauge 201 / spinor 202 (24^4 x 16), gauge 301 /
device-filled spinor 302 and CG right-hand side 303 (32^4), gauge 401 /
device-filled spinors 402 (24^4 x 12 x 16), batch gauge 5050 + i / spinors
6060 + i.
"""
import json
import os
import shutil

import numpy as np
import pytest

from conftest import ROOT, rel_l2, storage_tol

pytestmark = pytest.mark.gpu

NTHREADS = min(16, os.cpu_count() or 1)
TUNED = os.path.join(ROOT, "gridbench_amd", "tuned_mi355x.jsonl")
HEADLINE_BLOCK = (4, 6, 6, 12)
KAPPA = 0.12


class LatticeOracle:
    """The oracle's Dhop on one lattice and gauge field: Nsimd 1 tables and
    layout (the reference layout is then the lexicographic one)."""

    def __init__(self, orc, dims, Ls, dbl):
        self.orc, self.dims, self.Ls = orc, tuple(dims), Ls
        self.V = int(np.prod(dims))
        self.nbr, self.prm = orc.tables(self.dims, 1, Ls)
        self.U = orc.pack_gauge(self.dims, 1, dbl)

    def dhop(self, psi, dag=False):
        O = self.orc
        inp = O.pack_spinor(self.dims, 1, self.Ls, psi)
        out = O.hand(self.U, inp, self.nbr, self.prm, self.V, self.Ls, 1, dag=dag, nthreads=NTHREADS)
        del inp
        return O.unpack_spinor(self.dims, 1, self.Ls, out).reshape(self.V, self.Ls, 4, 3, 2)


def cached_cfg(gbx, sig):
    with open(TUNED) as f:
        for line in f:
            d = json.loads(line)
            if d["sig"] == sig:
                return gbx.kcfg(d["spw"], d["threads"], d["xcd"], d["nt"], kind=d["kind"], mc=d["mc"], pf=d["pf"],
                                wgpc=d["wgpc"], cblk=d["cblk"], tile=d["tile"])
    raise KeyError(sig)


@pytest.fixture
def tune_cache(tmp_path):
    """A scratch copy of the committed tuning cache (tuning appends to it)."""
    p = tmp_path / "tuned.jsonl"
    shutil.copy(TUNED, p)
    return str(p)


# ------------------------------------------------------------ 24^4 x Ls16
@pytest.fixture(scope="module")
def c2(gbx, orc):
    dims, Ls = (24, 24, 24, 24), 16
    dbl = gbx.gen_doubled_gauge(dims, 201)
    psi = gbx.gen_spinor(dims, Ls, 1, 202)
    o = LatticeOracle(orc, dims, Ls, dbl)
    return {"dims": dims, "Ls": Ls, "dbl": dbl, "psi": psi, "want": {dag: o.dhop(psi[0], dag) for dag in (False, True)}}


@pytest.mark.parametrize("dag", [False, True])
def test_headline_instance_24x24x24x24_ls16(gpu, gbx, c2, dag):
    """The exact timed instance: TDIAG field on 4x6x6x12 blocks with the
    cached tuned shape (spw 4, 64 threads, three-buffer gathers)."""
    cfg = cached_cfg(gbx, "p0_c0_Ls16_n1_24x24x24x24_td_b4x6x6x12")
    assert (cfg.kind, cfg.spw, cfg.threads, cfg.pf) == (0, 4, 64, 2)
    g = gbx.Geometry(c2["dims"], block=HEADLINE_BLOCK)
    U = gbx.GaugeField(g, "f64").upload(c2["dbl"])
    x = gbx.SpinorField(g, c2["Ls"], 1, "f64").upload(c2["psi"]).set_basis(gbx.BASIS_TDIAG)
    y = gbx.SpinorField(g, c2["Ls"], 1, "f64")
    gbx.dhop(U, y, x, dag=dag, cfg=cfg)
    got = y.set_basis(gbx.BASIS_CHIRAL).download()[0]
    err = rel_l2(got, c2["want"][dag])
    print(f"headline instance dag={dag}: rel L2 vs oracle {err:.3e}")
    assert err <= 1e-14
    # the same kernel shape on the chiral field of that geometry
    x.set_basis(gbx.BASIS_CHIRAL)
    gbx.dhop(U, y, x, dag=dag, cfg=cfg)
    assert rel_l2(y.download()[0], c2["want"][dag]) <= 1e-14


@pytest.mark.parametrize("tdiag", [False, True])
def test_f32_1rhs_24x24x24x24_ls16(gpu, gbx, c2, tune_cache, tdiag):
    g = gbx.Geometry(c2["dims"])
    U = gbx.GaugeField(g, "f32").upload(c2["dbl"])
    x = gbx.SpinorField(g, c2["Ls"], 1, "f32").upload(c2["psi"])
    if tdiag:
        x.set_basis(gbx.BASIS_TDIAG)
    y = gbx.SpinorField(g, c2["Ls"], 1, "f32")
    cfg, _ = gbx.tune(U, y, x, cache_path=tune_cache)
    for dag in (False, True):
        gbx.dhop(U, y, x, dag=dag)
        got = y.set_basis(gbx.BASIS_CHIRAL).download()[0]
        err = rel_l2(got, c2["want"][dag])
        print(f"FP32 tdiag={tdiag} dag={dag} cfg={cfg.as_dict()}: rel L2 vs FP64 oracle {err:.3e}")
        assert err <= 1e-6


# ------------------------------------------------------------ 32^4 x Ls16
@pytest.fixture(scope="module")
def c3(gbx, orc):
    dims, Ls = (32, 32, 32, 32), 16
    dbl = gbx.gen_doubled_gauge(dims, 301)
    g = gbx.Geometry(dims)
    U64 = gbx.GaugeField(g, "f64").upload(dbl)
    x64 = gbx.SpinorField(g, Ls, 1, "f64").fill_gaussian(302)
    psi = x64.download()[0]
    o = LatticeOracle(orc, dims, Ls, dbl)
    want = o.dhop(psi)
    yield {"dims": dims, "Ls": Ls, "dbl": dbl, "geom": g, "U64": U64, "x64": x64, "psi": psi, "want": want,
           "oracle": o}
    for f in (x64, U64):
        f.close()


@pytest.mark.parametrize("prec,comp12", [("f32", False), ("bf16", False), ("f16", False), ("f16", True)])
def test_storage_formats_32x32x32x32_ls16(gpu, gbx, c3, tune_cache, prec, comp12):
    g, Ls = c3["geom"], c3["Ls"]
    U = gbx.GaugeField(g, prec, comp12).upload(c3["dbl"])
    x = gbx.SpinorField(g, Ls, 1, prec).copy_from(c3["x64"])
    y = gbx.SpinorField(g, Ls, 1, prec)
    cfg, _ = gbx.tune(U, y, x, cache_path=tune_cache)
    gbx.dhop(U, y, x)
    err = rel_l2(y.download()[0], c3["want"])
    tol = storage_tol(prec, c3["psi"])
    print(f"{prec} comp12={comp12} cfg={cfg.as_dict()}: rel L2 vs FP64 oracle {err:.3e} (bound {tol:.3e})")
    assert err <= tol
    for f in (U, x, y):
        f.close()


def test_f64_32x32x32x32_ls16(gpu, gbx, c3):
    y = gbx.SpinorField(c3["geom"], c3["Ls"], 1, "f64")
    gbx.dhop(c3["U64"], y, c3["x64"])
    assert rel_l2(y.download()[0], c3["want"]) <= 1e-14
    y.close()


@pytest.mark.parametrize("lo,comp12", [("f32", False), ("bf16", False), ("f16", False), ("f16", True)])
def test_cg_32x32x32x32_ls16_reaches_1e10(gpu, gbx, c3, lo, comp12):
    """Defect-correction CG on M^dag M x = b to 1e-10; the residual of the
    returned x is recomputed on the host with the oracle's Dhop."""
    g, Ls = c3["geom"], c3["Ls"]
    b = gbx.SpinorField(g, Ls, 1, "f64").fill_gaussian(303)
    x = gbx.SpinorField(g, Ls, 1, "f64").zero()
    Ulo = gbx.GaugeField(g, lo, comp12).upload(c3["dbl"])
    st = gbx.cg_solve(c3["U64"], Ulo, x, b, KAPPA, tol=1e-10, inner_tol=1e-3, max_outer=60, max_inner=5000)
    xs, bs = x.download()[0], b.download()[0]
    for f in (b, x, Ulo):
        f.close()
    o = c3["oracle"]
    t = xs - KAPPA * o.dhop(xs)
    r = bs - (t - KAPPA * o.dhop(t, dag=True))
    del t
    rel = float(np.linalg.norm(r) / np.linalg.norm(bs))
    print(f"inner {lo} comp12={comp12}: {st}; oracle residual {rel:.3e}")
    assert st["true_resid"] <= 1e-10
    assert rel <= 2e-10
    assert st["outer_iters"] >= 2 and st["lo_applies"] > st["hi_applies"]


# ---------------------------------------------------- 24^4 x Ls12 x 16 RHS
C4_CHECK_RHS = (0, 7, 15)


@pytest.fixture(scope="module")
def c4(gbx, orc):
    dims, Ls, nrhs = (24, 24, 24, 24), 12, 16
    dbl = gbx.gen_doubled_gauge(dims, 401)
    g = gbx.Geometry(dims)
    U64 = gbx.GaugeField(g, "f64").upload(dbl)
    x64 = gbx.SpinorField(g, Ls, nrhs, "f64").fill_gaussian(402)
    y64 = gbx.SpinorField(g, Ls, nrhs, "f64")
    gbx.dhop(U64, y64, x64)
    o = LatticeOracle(orc, dims, Ls, dbl)
    psi = {r: x64.download_rhs(r) for r in C4_CHECK_RHS}
    want = {r: o.dhop(psi[r]) for r in C4_CHECK_RHS}
    yield {"dims": dims, "Ls": Ls, "nrhs": nrhs, "dbl": dbl, "geom": g, "x64": x64, "y64": y64, "psi": psi,
           "want": want}
    for f in (x64, y64, U64):
        f.close()


def test_f64_kernel_24x24x24x24_ls12_n16(gpu, c4):
    """The FP64 kernel the multi-RHS forms are compared with, on the oracle."""
    for r in C4_CHECK_RHS:
        assert rel_l2(c4["y64"].download_rhs(r), c4["want"][r]) <= 1e-14


@pytest.mark.parametrize("prec,form", [("f32", "valu"), ("f32", "tile"), ("f32", "mfma"), ("bf16", "valu"),
                                       ("bf16", "mfma")])
def test_multi_rhs_24x24x24x24_ls12_n16(gpu, gbx, c4, tune_cache, prec, form):
    g, Ls, nrhs = c4["geom"], c4["Ls"], c4["nrhs"]
    U = gbx.GaugeField(g, prec).upload(c4["dbl"])
    x = gbx.SpinorField(g, Ls, nrhs, prec).copy_from(c4["x64"])
    y = gbx.SpinorField(g, Ls, nrhs, prec)
    assert y.nbytes > 1 << 32 or prec != "f32"
    cfg, _ = gbx.tune(U, y, x, kinds=form, cache_path=tune_cache)
    gbx.dhop(U, y, x, cfg=cfg)
    tol = 1e-6 if prec == "f32" else storage_tol(prec, c4["psi"][0])
    for r in C4_CHECK_RHS:
        err = rel_l2(y.download_rhs(r), c4["want"][r])
        print(f"{prec} {form} cfg={cfg.as_dict()} rhs {r}: rel L2 vs FP64 oracle {err:.3e} (bound {tol:.3e})")
        assert err <= tol
    # every right-hand side against the FP64 kernel on the unrounded input
    errs = [rel_l2(y.download_rhs(r), c4["y64"].download_rhs(r)) for r in range(nrhs)]
    print(f"{prec} {form}: max rel L2 over 16 rhs vs FP64 kernel {max(errs):.3e}")
    assert max(errs) <= tol
    for f in (U, x, y):
        f.close()


# ------------------------------------------------ the uneven batch of 40
def batch_specs():
    rng = np.random.default_rng(505)
    return [(int(rng.choice([8, 12, 16, 20, 24])), int(rng.choice([8, 12, 16])), int(rng.integers(1, 9)))
            for _ in range(40)]


@pytest.fixture(scope="module")
def c5(gbx):
    specs = batch_specs()
    lats = []
    for i, (L, Ls, nrhs) in enumerate(specs):
        g = gbx.Geometry((L, L, L, L))
        U = gbx.GaugeField(g, "f32").upload(gbx.gen_doubled_gauge((L, L, L, L), 5050 + i))
        x0 = gbx.SpinorField(g, Ls, nrhs, "f32").fill_gaussian(6060 + i)  # the input's values, chiral basis
        x = gbx.SpinorField(g, Ls, nrhs, "f32").copy_from(x0).set_basis(gbx.BASIS_TDIAG)
        y = gbx.SpinorField(g, Ls, nrhs, "f32")
        lats.append((g, U, x0, x, y))
    b = gbx.Batch([l[1] for l in lats], [l[4] for l in lats], [l[3] for l in lats])
    b.run()
    for l in lats:
        l[4].set_basis(gbx.BASIS_CHIRAL)
    yield specs, lats
    b.close()
    for l in lats:
        for f in l[1:]:
            f.close()


@pytest.mark.parametrize("part", range(4))
def test_uneven_batch_of_40_vs_oracle(gpu, gbx, orc, c5, part):
    """Lattices 10*part .. 10*part+9 of the one-launch batch, every right-hand
    side against the oracle on the FP32 input values (<= 1e-6)."""
    specs, lats = c5
    worst = 0.0
    for i in range(10 * part, 10 * part + 10):
        L, Ls, nrhs = specs[i]
        g, U, x0, x, y = lats[i]
        o = LatticeOracle(orc, (L, L, L, L), Ls, gbx.gen_doubled_gauge((L, L, L, L), 5050 + i))
        for r in range(nrhs):
            err = rel_l2(y.download_rhs(r), o.dhop(x0.download_rhs(r)))
            worst = max(worst, err)
            assert err <= 1e-6, (i, specs[i], r, err)
    print(f"lattices {10 * part}..{10 * part + 9}: max rel L2 vs FP64 oracle {worst:.3e}")
