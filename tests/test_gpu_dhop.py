"""GPU parity of the hopping kernels against the oracle (the CPU restatement
of the reference Hand kernel, WilsonKernelsHandCpu.h:385-451).

Tolerances (relative L2 over the whole output, vs the FP64 oracle):
  FP64 storage           <= 1e-14   (FMA contraction on the GPU only)
  FP32 storage           <= 1e-6
  FP16 / BF16 storage    <= 2.5 x the relative error of storing the input
                            field in that format (conftest.storage_tol: one
                            rounding each of spinor, links and output gives
                            about 1.7 x; 16-bit sums would exceed 2.5 x)
"""
import numpy as np
import pytest

from conftest import rel_l2, storage_tol

pytestmark = pytest.mark.gpu

TOL = {"f64": 1e-14, "f32": 1e-6}


def run_dhop(gbx, p, prec, dag=False, comp12=False, block=None, cfg=None, psi=None):
    geom = gbx.Geometry(p.dims, block)
    U = gbx.GaugeField(geom, prec, comp12).upload(p.dbl)
    src = gbx.SpinorField(geom, p.Ls, p.nrhs, prec).upload(p.psi if psi is None else psi)
    dst = gbx.SpinorField(geom, p.Ls, p.nrhs, prec)
    gbx.dhop(U, dst, src, dag=dag, cfg=cfg)
    return dst.download()


@pytest.mark.parametrize("dims,Ls", [((4, 4, 4, 4), 4), ((8, 8, 8, 8), 8), ((4, 6, 8, 2), 3), ((2, 2, 2, 2), 1),
                                     ((6, 4, 2, 10), 5)])
@pytest.mark.parametrize("dag", [False, True])
def test_f64_vs_oracle(gpu, gbx, problem, dims, Ls, dag):
    p = problem(dims, Ls)
    got = run_dhop(gbx, p, "f64", dag=dag)[0]
    want = p.oracle_lex(dag=dag)
    err = rel_l2(got, want)
    print(f"dims={dims} Ls={Ls} dag={dag} rel_l2={err:.3e}")
    assert err <= TOL["f64"]


@pytest.mark.parametrize("prec", ["f32", "f16", "bf16"])
@pytest.mark.parametrize("comp12", [False, True])
@pytest.mark.parametrize("dag", [False, True])
def test_low_precision_vs_f64_oracle(gpu, gbx, problem, prec, comp12, dag):
    p = problem((8, 4, 4, 8), 6)
    got = run_dhop(gbx, p, prec, dag=dag, comp12=comp12)[0]
    want = p.oracle_lex(dag=dag)
    err = rel_l2(got, want)
    tol = storage_tol(prec, p.psi)
    print(f"prec={prec} comp12={comp12} dag={dag} rel_l2_vs_f64_oracle={err:.3e} bound={tol:.3e}")
    assert err <= tol


def test_f32_vs_f32_oracle(gpu, gbx, problem):
    p = problem((4, 4, 4, 8), 4)
    got = run_dhop(gbx, p, "f32")[0]
    want = p.oracle_lex(dtype=np.float32)
    assert rel_l2(got, want) <= 1e-6


def test_f64_comp12(gpu, gbx, problem):
    p = problem((4, 4, 6, 4), 4)
    got = run_dhop(gbx, p, "f64", comp12=True)[0]
    assert rel_l2(got, p.oracle_lex()) <= 1e-14


@pytest.mark.parametrize("nrhs", [2, 5])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_multi_rhs(gpu, gbx, problem, nrhs, prec):
    p = problem((4, 4, 4, 4), 4, nrhs=nrhs)
    got = run_dhop(gbx, p, prec)
    for r in range(nrhs):
        err = rel_l2(got[r], p.oracle_lex(rhs=r))
        print(f"rhs={r} rel_l2={err:.3e}")
        assert err <= TOL[prec]


@pytest.mark.parametrize("block", [(1, 1, 1, 1), (4, 2, 2, 2), (8, 4, 4, 4), (2, 4, 1, 8)])
def test_traversal_blocks(gpu, gbx, problem, block):
    p = problem((8, 4, 4, 8), 4)
    got = run_dhop(gbx, p, "f64", block=block)[0]
    assert rel_l2(got, p.oracle_lex()) <= 1e-14


@pytest.mark.parametrize("spw,threads,xcd,nt", [(1, 64, 0, 0), (3, 64, 1, 1), (16, 256, 1, 0), (7, 128, 0, 1),
                                                (64, 1024, 1, 1), (5, 512, 1, 1)])
def test_launch_configs_identical(gpu, gbx, problem, spw, threads, xcd, nt):
    """Every launch configuration computes the same bits (the work split
    changes, the per-point arithmetic does not)."""
    p = problem((4, 6, 4, 6), 4)
    base = run_dhop(gbx, p, "f64")
    got = run_dhop(gbx, p, "f64", cfg=gbx.kcfg(spw, threads, xcd, nt))
    assert np.array_equal(base, got)
    if threads <= 256:  # the gather pipelines (launch bound 256)
        for pf in (1, 2):
            got = run_dhop(gbx, p, "f64", cfg=gbx.kcfg(spw, threads, xcd, nt, pf=pf))
            assert np.array_equal(base, got), pf


def test_axpby(gpu, gbx, problem):
    """out = a*in + b*Dhop(in) (the M = 1 - kappa Dhop form)."""
    p = problem((4, 4, 4, 4), 4)
    geom = gbx.Geometry(p.dims)
    U = gbx.GaugeField(geom, "f64").upload(p.dbl)
    src = gbx.SpinorField(geom, p.Ls, 1, "f64").upload(p.psi)
    dst = gbx.SpinorField(geom, p.Ls, 1, "f64")
    for dag in (False, True):
        gbx.dhop_axpby(U, dst, src, 1.0, -0.12, dag=dag)
        want = p.psi[0] - 0.12 * p.oracle_lex(dag=dag)
        got = dst.download()[0]
        assert rel_l2(got, want) <= 1e-14
        # the fused form on every gather pipeline depth: the same bits
        for pf in (0, 1, 2):
            gbx.dhop_xpay(U, dst, src, src, 1.0, -0.12, dag=dag, cfg=gbx.kcfg(4, 64, 1, 0, pf=pf))
            assert np.array_equal(dst.download()[0], got), pf
    with pytest.raises(gbx.GbxError) as e:
        gbx.dhop_xpay(U, dst, src, src, 1.0, -0.12, cfg=gbx.kcfg(4, 64, kind=2, tile=(2, 2, 2, 2)))
    assert e.value.status == 8


def test_errors_on_device_objects(gpu, gbx, problem):
    p = problem((4, 4, 4, 4), 2)
    geom = gbx.Geometry(p.dims)
    U = gbx.GaugeField(geom, "f64").upload(p.dbl)
    a = gbx.SpinorField(geom, 2, 1, "f64").upload(p.psi)
    b32 = gbx.SpinorField(geom, 2, 1, "f32")
    with pytest.raises(gbx.GbxError) as e:
        gbx.dhop(U, a, a)
    assert e.value.status == 3  # GBX_EALIAS
    with pytest.raises(gbx.GbxError) as e:
        gbx.dhop(U, b32, a)
    assert e.value.status == 9  # GBX_EMISMATCH
    other = gbx.SpinorField(gbx.Geometry((4, 4, 4, 4)), 2, 1, "f64")
    with pytest.raises(gbx.GbxError) as e:
        gbx.dhop(U, other, a)
    assert e.value.status == 9
    with pytest.raises(gbx.GbxError):
        gbx.dhop(U, gbx.SpinorField(geom, 2, 1, "f64"), a, cfg=gbx.kcfg(1, 100))  # threads not a multiple of 64


def test_upload_download_roundtrip(gpu, gbx, problem):
    p = problem((4, 6, 2, 4), 3, nrhs=2)
    geom = gbx.Geometry(p.dims)
    x = gbx.SpinorField(geom, 3, 2, "f64").upload(p.psi)
    assert np.array_equal(x.download(), p.psi)
    for prec, tol in (("f32", 1e-7), ("f16", 1e-3), ("bf16", 8e-3)):
        y = gbx.SpinorField(geom, 3, 2, prec).copy_from(x)
        assert rel_l2(y.download(), p.psi) <= tol
        z = gbx.SpinorField(geom, 3, 2, prec).upload(p.psi)
        assert np.array_equal(y.download(), z.download())


def test_fill_gaussian(gpu, gbx):
    """Device Gaussian fill: the field is a function of the lexicographic 5D
    index (same bits under two site orders), each precision is the FP64 field
    rounded, seeds differ, and the sample moments are those of N(0,1)."""
    dims, Ls, nrhs = (8, 4, 6, 4), 4, 3
    ga, gb = gbx.Geometry(dims), gbx.Geometry(dims, block=(1, 1, 1, 1))
    a = gbx.SpinorField(ga, Ls, nrhs, "f64").fill_gaussian(77).download()
    b = gbx.SpinorField(gb, Ls, nrhs, "f64").fill_gaussian(77).download()
    assert np.array_equal(a, b)
    c = gbx.SpinorField(ga, Ls, nrhs, "f64").fill_gaussian(78).download()
    assert not np.any(a == c)
    lo = gbx.SpinorField(ga, Ls, nrhs, "f32").fill_gaussian(77).download()
    assert np.array_equal(lo, a.astype(np.float32).astype(np.float64))
    h = gbx.SpinorField(ga, Ls, nrhs, "f16").fill_gaussian(77).download()
    assert np.all(np.abs(h - a) <= 2.0 ** -10 * np.abs(a) + 2.0 ** -24)
    v = a.reshape(-1)
    n = v.size
    assert abs(v.mean()) < 5 / np.sqrt(n)
    assert abs(v.var() - 1.0) < 5 * np.sqrt(2.0 / n)
    assert abs(np.mean(v ** 4) - 3.0) < 0.1


# ----------------------------------------------------------- drop-in entry
def ref_case(gbx, orc, dims, Ls, nsimd, dtype):
    links = gbx.gen_su3_links(dims, 101)
    dbl = gbx.double_links(dims, links)
    psi = gbx.gen_spinor(dims, Ls, 1, 102)[0]
    U = orc.pack_gauge(dims, nsimd, dbl).astype(dtype)
    inp = orc.pack_spinor(dims, nsimd, Ls, psi).astype(dtype)
    nbr, prm = orc.tables(dims, nsimd, Ls)
    return U, inp, nbr, prm, int(np.prod(dims)) // nsimd


@pytest.mark.parametrize("nsimd", [1, 2, 4, 8, 16])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("dag", [False, True])
def test_dslash_kernel_dropin_canonical(gpu, gbx, orc, nsimd, dtype, dag):
    dims, Ls = (4, 4, 4, 8), 4
    U, inp, nbr, prm, nsite = ref_case(gbx, orc, dims, Ls, nsimd, dtype)
    out = np.zeros_like(inp)
    usec = gbx.dslash_kernel(3, U, out, inp, nbr, nsite, Ls, prm, nsimd, dag=dag)
    assert gbx.dslash_last_path() == 1
    want = orc.hand(U, inp, nbr, prm, nsite, Ls, nsimd, dag=dag)
    # the drop-in computes with the reference's rounding: the oracle's bits
    assert np.array_equal(out, want)
    assert usec >= 0.0


@pytest.mark.parametrize("nsimd", [1, 4, 8])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_dslash_kernel_dropin_arbitrary_tables(gpu, gbx, orc, nsimd, dtype):
    """Non-canonical tables (outer sites relabelled by a permutation) take
    the table-driven kernel and must still match the oracle on the same
    tables and arrays."""
    dims, Ls = (4, 4, 4, 8), 3
    U, inp, nbr, prm, nsite = ref_case(gbx, orc, dims, Ls, nsimd, dtype)
    rng = np.random.default_rng(7)
    perm = rng.permutation(nsite)  # new index of old outer site
    inv = np.argsort(perm)
    Ur = U.reshape(nsite, -1)[inv].reshape(-1).copy()
    inr = inp.reshape(nsite, Ls, -1)[inv].reshape(-1).copy()
    nb = nbr.reshape(nsite, Ls, 8)[inv]
    nb = (perm[nb // Ls] * Ls + nb % Ls).astype(np.uint64).reshape(-1).copy()
    pr = prm.reshape(nsite, Ls, 8)[inv].reshape(-1).copy()
    out = np.zeros_like(inr)
    gbx.dslash_kernel(2, Ur, out, inr, nb, nsite, Ls, pr, nsimd)
    assert gbx.dslash_last_path() == 2
    want = orc.hand(orc.aligned_copy(Ur), orc.aligned_copy(inr), orc.aligned_copy(nb), orc.aligned_copy(pr), nsite, Ls,
                    nsimd)
    assert np.array_equal(out, want)


def relabel_outer_sites(U, inp, nbr, prm, nsite, Ls, seed=7):
    """The same problem with its outer sites permuted: non-canonical tables."""
    rng = np.random.default_rng(seed)
    perm = rng.permutation(nsite)  # new index of old outer site
    inv = np.argsort(perm)
    Ur = U.reshape(nsite, -1)[inv].reshape(-1).copy()
    inr = inp.reshape(nsite, Ls, -1)[inv].reshape(-1).copy()
    nb = nbr.reshape(nsite, Ls, 8)[inv]
    nb = (perm[nb // Ls] * Ls + nb % Ls).astype(np.uint64).reshape(-1).copy()
    pr = prm.reshape(nsite, Ls, 8)[inv].reshape(-1).copy()
    return Ur, inr, nb, pr


@pytest.mark.parametrize("nsimd", [1, 2, 8, 16])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("tables", ["canonical", "relabelled"])
def test_dslash_kernel_rrii_layout(gpu, gbx, orc, nsimd, dtype, tables):
    """RRII host arrays (bench.cc:105-106): the result is the RIRI result in
    RRII order, bit for bit, on both device paths, and matches the oracle."""
    dims, Ls = (4, 4, 4, 8), 3
    U, inp, nbr, prm, nsite = ref_case(gbx, orc, dims, Ls, nsimd, dtype)
    if tables == "relabelled":
        U, inp, nbr, prm = relabel_outer_sites(U, inp, nbr, prm, nsite, Ls)
    for dag in (False, True):
        out = np.zeros_like(inp)
        gbx.dslash_kernel(2, U, out, inp, nbr, nsite, Ls, prm, nsimd, dag=dag)
        path = gbx.dslash_last_path()
        out_r = np.zeros_like(inp)
        gbx.dslash_kernel(2, orc.to_rrii(U, nsimd), out_r, orc.to_rrii(inp, nsimd), nbr, nsite, Ls, prm, nsimd, dag=dag,
                          rrii=True)
        assert gbx.dslash_last_path() == path == (1 if tables == "canonical" else 2)
        assert np.array_equal(out_r, orc.to_rrii(out, nsimd))
        want = orc.hand(orc.aligned_copy(U), orc.aligned_copy(inp), orc.aligned_copy(nbr), orc.aligned_copy(prm), nsite,
                        Ls, nsimd, dag=dag)
        assert np.array_equal(orc.from_rrii(out_r, nsimd), want)


@pytest.mark.parametrize("data_nsimd,width", [(1, 4), (2, 8), (4, 16), (8, 8)])
@pytest.mark.parametrize("rrii", [False, True])
@pytest.mark.parametrize("tables", ["canonical", "relabelled"])
def test_dslash_kernel_expanded_simd(gpu, gbx, orc, data_nsimd, width, rrii, tables):
    """Vectors of `width` lanes holding width/data_nsimd replicas of a
    data_nsimd-lane problem (bench.cc:1-3, 98-138); tables and permutes are
    those of data_nsimd (Simd_riri_vec.h:466), so replica nn's output is the
    oracle's Dhop of replica nn.  The replicas here hold different fields."""
    dims, Ls = (4, 4, 4, 8), 2
    R = width // data_nsimd
    nsite = int(np.prod(dims)) // data_nsimd
    nbr, prm = orc.tables(dims, data_nsimd, Ls)
    Us, ins = [], []
    for nn in range(R):
        dbl = gbx.double_links(dims, gbx.gen_su3_links(dims, 600 + nn))
        Us.append(orc.pack_gauge(dims, data_nsimd, dbl))
        ins.append(orc.pack_spinor(dims, data_nsimd, Ls, gbx.gen_spinor(dims, Ls, 1, 700 + nn)[0]))
    if tables == "relabelled":
        rel = [relabel_outer_sites(u, i, nbr, prm, nsite, Ls) for u, i in zip(Us, ins)]
        Us, ins = [r[0] for r in rel], [r[1] for r in rel]
        nbr, prm = rel[0][2], rel[0][3]
    U, inp = orc.expand_lanes(Us, data_nsimd), orc.expand_lanes(ins, data_nsimd)
    if rrii:
        U, inp = orc.to_rrii(U, width), orc.to_rrii(inp, width)
    out = np.zeros_like(inp)
    gbx.dslash_kernel(2, U, out, inp, nbr, nsite, Ls, prm, width, rrii=rrii, data_nsimd=data_nsimd)
    assert gbx.dslash_last_path() == (1 if tables == "canonical" else 2)
    if rrii:
        out = orc.from_rrii(out, width)
    for nn in range(R):
        want = orc.hand(orc.aligned_copy(Us[nn]), orc.aligned_copy(ins[nn]), orc.aligned_copy(nbr),
                        orc.aligned_copy(prm), nsite, Ls, data_nsimd)
        assert np.array_equal(orc.replica_lanes(out, width, data_nsimd, nn), want)


def test_dslash_contexts_from_threads(gpu, gbx, orc):
    """Two contexts driven from two host threads at once (one on the SoA
    path, one on the table path, each with its own buffers and stream) give
    the same bits as the process-wide context, and keep their own last path."""
    import threading
    dims, Ls = (4, 4, 4, 8), 3
    A = ref_case(gbx, orc, dims, Ls, 4, np.float64)
    B = ref_case(gbx, orc, dims, Ls, 2, np.float32)
    Br = relabel_outer_sites(*B[:4], B[4], Ls)
    jobs = [("A", 4, A[0], A[1], A[2], A[3], A[4], 1), ("B", 2, Br[0], Br[1], Br[2], Br[3], B[4], 2)]
    want = {}
    for name, ns, U, inp, nbr, prm, nsite, path in jobs:
        out = np.zeros_like(inp)
        gbx.dslash_kernel(2, U, out, inp, nbr, nsite, Ls, prm, ns, dag=True)
        want[name] = out
    got, errors = {}, []

    def work(ctx, job):
        name, ns, U, inp, nbr, prm, nsite, path = job
        try:
            for rep in range(5):
                out = np.zeros_like(inp)
                ctx.kernel(2, U, out, inp, nbr, nsite, Ls, prm, ns, dag=True)
                assert np.array_equal(out, want[name]), (name, rep)
                assert ctx.last_path() == path
            got[name] = True
        except Exception as e:  # reported below, in the main thread
            errors.append(e)

    ctxs = [gbx.DslashContext(), gbx.DslashContext()]
    threads = [threading.Thread(target=work, args=(c, j)) for c, j in zip(ctxs, jobs)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not errors, errors
    assert got == {"A": True, "B": True}
    # a context reused with another shape and precision
    ctxs[0].kernel(2, Br[0], np.zeros_like(Br[1]), Br[1], Br[2], B[4], Ls, Br[3], 2)
    assert ctxs[0].last_path() == 2
    for c in ctxs:
        c.close()


def test_dslash_kernel_layout_argument_errors(gpu, gbx, orc):
    dims, Ls = (4, 4, 4, 4), 2
    U, inp, nbr, prm, nsite = ref_case(gbx, orc, dims, Ls, 2, np.float64)
    out = np.zeros_like(inp)
    for width, data in ((2, 4), (6, 2), (128, 2)):  # data wider than vector, not a power of two, too wide
        with pytest.raises(gbx.GbxError) as e:
            gbx.dslash_kernel(1, U, out, inp, nbr, nsite, Ls, prm, width, data_nsimd=data)
        assert e.value.status == 1


def test_dslash_kernel_rejects_bad_tables(gpu, gbx, orc):
    dims, Ls, nsimd = (4, 4, 4, 4), 2, 1
    U, inp, nbr, prm, nsite = ref_case(gbx, orc, dims, Ls, nsimd, np.float64)
    out = np.zeros_like(inp)
    bad = nbr.copy()
    bad[17] = nsite * Ls
    with pytest.raises(gbx.GbxError) as e:
        gbx.dslash_kernel(1, U, out, inp, bad, nsite, Ls, prm, nsimd)
    assert e.value.status == 4  # GBX_ETABLE
    badp = prm.copy()
    badp[3] = 2
    with pytest.raises(gbx.GbxError) as e:
        gbx.dslash_kernel(1, U, out, inp, nbr, nsite, Ls, badp, nsimd)
    assert e.value.status == 4
    with pytest.raises(gbx.GbxError) as e:
        gbx.dslash_kernel(1, U, inp, inp, nbr, nsite, Ls, prm, nsimd)
    assert e.value.status == 3  # GBX_EALIAS


# ----------------------------------------------------------- full size
def test_full_size_24x4_ls16_f64(gpu, gbx, problem):
    """24^4 x Ls16 FP64 (the benchmark workload) against the oracle run on the
    host cores, plus size-independent properties."""
    p = problem((24, 24, 24, 24), 16, gseed=201, sseed=202)
    geom = gbx.Geometry(p.dims)
    U = gbx.GaugeField(geom, "f64").upload(p.dbl)
    src = gbx.SpinorField(geom, 16, 1, "f64").upload(p.psi)
    dst = gbx.SpinorField(geom, 16, 1, "f64")
    gbx.dhop(U, dst, src)
    got = dst.download()[0]
    want = p.oracle_lex(nthreads=16)
    err = rel_l2(got, want)
    print(f"24^4 x 16 FP64 rel_l2 vs oracle = {err:.3e}")
    assert err <= 1e-14
    # linearity: Dhop(2 psi) = 2 Dhop(psi) exactly (scaling by 2 is exact)
    src.scale(2.0)
    gbx.dhop(U, dst, src)
    assert np.array_equal(dst.download()[0], 2.0 * got)


@pytest.mark.parametrize("dtype,dag", [(np.float64, False), (np.float64, True), (np.float32, False)])
def test_dropin_full_size_24x4_ls16_bitexact(gpu, gbx, orc, dtype, dag):
    """The drop-in entry at the benchmark size (24^4 x Ls16, Nsimd 4, the
    reference's 64-byte vectors) against the oracle on the host cores: the
    same bits, FP64 and FP32, Dhop and DhopDag."""
    dims, Ls, nsimd = (24, 24, 24, 24), 16, 4
    dbl = gbx.double_links(dims, gbx.gen_su3_links(dims, 201))
    U = orc.pack_gauge(dims, nsimd, dbl).astype(dtype)
    inp = orc.pack_spinor(dims, nsimd, Ls, gbx.gen_spinor(dims, Ls, 1, 202)[0]).astype(dtype)
    nbr, prm = orc.tables(dims, nsimd, Ls)
    nsite = int(np.prod(dims)) // nsimd
    out = np.zeros_like(inp)
    gbx.dslash_kernel(2, U, out, inp, nbr, nsite, Ls, prm, nsimd, dag=dag)
    assert gbx.dslash_last_path() == 1
    want = orc.hand(orc.aligned_copy(U), orc.aligned_copy(inp), nbr, prm, nsite, Ls, nsimd, dag=dag, nthreads=16)
    assert np.array_equal(out, want)


# ----------------------------------------------------------- tuner, orders
def test_tuner_picks_a_config_with_identical_results(gpu, gbx, problem, tmp_path):
    p = problem((8, 8, 8, 8), 8)
    geom = gbx.Geometry(p.dims)
    U = gbx.GaugeField(geom, "f32").upload(p.dbl)
    x = gbx.SpinorField(geom, 8, 1, "f32").upload(p.psi)
    y = gbx.SpinorField(geom, 8, 1, "f32")
    gbx.dhop(U, y, x, cfg=gbx.kcfg(1, 64, 0, 0))
    base = y.download()
    cache = str(tmp_path / "tune.jsonl")
    best, us = gbx.tune(U, y, x, cache_path=cache)
    assert us > 0 and best.threads % 64 == 0
    assert gbx.dhop_get_cfg(U, x).as_dict() == best.as_dict()
    y.zero()
    gbx.dhop(U, y, x)  # now runs the tuned configuration
    assert np.array_equal(y.download(), base)
    best2, _ = gbx.tune(U, y, x, cache_path=cache)  # served from the cache file
    assert best2.as_dict() == best.as_dict()


@pytest.mark.parametrize("levels", [[(2, 2, 2, 2)], [(2, 2, 2, 2), (4, 4, 4, 4)], [(1, 2, 1, 2), (2, 4, 4, 4), (4, 4, 8, 8)]])
def test_multilevel_site_orders(gpu, gbx, problem, levels):
    p = problem((4, 4, 8, 8), 4)
    geom = gbx.Geometry(p.dims, levels=levels)
    U = gbx.GaugeField(geom, "f64").upload(p.dbl)
    x = gbx.SpinorField(geom, 4, 1, "f64").upload(p.psi)
    y = gbx.SpinorField(geom, 4, 1, "f64")
    gbx.dhop(U, y, x)
    assert rel_l2(y.download()[0], p.oracle_lex()) <= 1e-14


@pytest.mark.parametrize("mc,cblk,pf", [(2, 0, 0), (4, 3, 1), (8, 7, 0), (4, 1000, 1), (8, 0, 2), (2, 5, 2)])
def test_minor_chunk_configs(gpu, gbx, problem, mc, cblk, pf):
    p = problem((4, 4, 4, 6), 4, nrhs=2)
    geom = gbx.Geometry(p.dims)
    U = gbx.GaugeField(geom, "f32").upload(p.dbl)
    x = gbx.SpinorField(geom, 4, 2, "f32").upload(p.psi)
    y = gbx.SpinorField(geom, 4, 2, "f32")
    gbx.dhop(U, y, x, cfg=gbx.kcfg(1, 64, 0, 0))
    base = y.download()
    y.zero()
    gbx.dhop(U, y, x, cfg=gbx.kcfg(8, 64, 1, 1, 0, mc, pf, 4, cblk))
    assert np.array_equal(y.download(), base)


def test_cpp_host_bench(gpu):
    """The C++ host program drives dslash_kernel<Simd> through gbx_dslash.hpp
    and checks the result against the device-resident path."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "gridbench_amd", "dslash_bench")
    for args in (["8", "8", "20", "d", "1"], ["8", "4", "5", "d", "8"], ["8", "8", "10", "f", "16"]):
        r = subprocess.run([exe] + args, capture_output=True, text=True, timeout=120)
        print(r.stdout)
        assert r.returncode == 0, r.stdout + r.stderr
        assert "path 1" in r.stdout


def test_full_size_16rhs_f32_vs_f64(gpu, gbx):
    """24^4 x Ls16 x 16 right-hand sides in FP32 storage: every right-hand side
    against the FP64 kernel on the same input (<= 1e-6 relative)."""
    dims, Ls, nrhs = (24, 24, 24, 24), 16, 16
    V = 24 ** 4
    geom = gbx.Geometry(dims)
    dbl = gbx.gen_doubled_gauge(dims, 201)
    U32 = gbx.GaugeField(geom, "f32").upload(dbl)
    U64 = gbx.GaugeField(geom, "f64").upload(dbl)
    x = gbx.SpinorField(geom, Ls, nrhs, "f32")
    x64 = gbx.SpinorField(geom, Ls, 1, "f64")
    y64 = gbx.SpinorField(geom, Ls, 1, "f64")
    rhs = [0, 7, 15]
    for r in range(nrhs):
        x.upload_rhs(r, gbx.gen_gaussian(202, r, 0, V * Ls * 24))
    y = gbx.SpinorField(geom, Ls, nrhs, "f32")
    gbx.dhop(U32, y, x)
    for r in rhs:
        x64.upload_rhs(0, gbx.gen_gaussian(202, r, 0, V * Ls * 24))
        gbx.dhop(U64, y64, x64)
        err = rel_l2(y.download_rhs(r), y64.download_rhs(0))
        print(f"rhs {r}: rel_l2(FP32 16-RHS vs FP64) = {err:.2e}")
        assert err <= 1e-6


def test_caller_defined_site_order(gpu, gbx, problem):
    """gbx_geom_create_order: any permutation of the sites gives the same
    Dhop bit for bit (the order changes only which slot a site occupies)."""
    p = problem((4, 4, 4, 8), 4)
    want = run_dhop(gbx, p, "f64", cfg=gbx.kcfg(4, 64))[0]
    rng = np.random.default_rng(7)
    for perm in (rng.permutation(p.V), np.arange(p.V)[::-1].copy()):
        g = gbx.Geometry.with_order(p.dims, perm)
        U = gbx.GaugeField(g, "f64").upload(p.dbl)
        x = gbx.SpinorField(g, p.Ls, 1, "f64").upload(p.psi)
        y = gbx.SpinorField(g, p.Ls, 1, "f64")
        gbx.dhop(U, y, x, cfg=gbx.kcfg(4, 64))
        assert np.array_equal(y.download()[0], want)
    with pytest.raises(gbx.GbxError):
        gbx.Geometry.with_order(p.dims, np.zeros(p.V, np.uint32))  # not a permutation
