"""Matrix-core form for FP32 storage (kernel kind 1 on FP32 fields,
k_dhop_mfma32): v_mfma_f32_16x16x4f32 on the projected half spinors.  Every
right-hand side against the FP64 oracle (WilsonKernelsHandCpu.h:426-446, the
link shared across s and right-hand sides as in MULT_2SPIN, lines 62-90) and
the whole output against the VALU kernel on the same FP32 inputs, both
within 1e-6 (the FP32 bound of the VALU kernel); chiral and GBX_BASIS_TDIAG
fields (whose +-t legs feed one spin half straight to the MFMAs), Dhop and
DhopDag, 18- and 12-real links.  The two kernels sum the 8 legs in different
orders, so they agree to FP32 rounding, not bit for bit; across launch
shapes the MFMA form computes every output element by the same MFMA
sequence and is bit-identical."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-6
CASES = [((4, 4, 4, 4), 8, 2), ((8, 4, 4, 8), 4, 8), ((4, 6, 2, 4), 16, 1), ((2, 2, 2, 2), 16, 2), ((4, 4, 2, 6), 12, 4)]


def fields(gbx, p, comp12=False, tdiag=False):
    geom = gbx.Geometry(p.dims)
    U = gbx.GaugeField(geom, "f32", comp12).upload(p.dbl)
    x = gbx.SpinorField(geom, p.Ls, p.nrhs, "f32").upload(p.psi)
    if tdiag:
        x.set_basis(gbx.BASIS_TDIAG)
    y = gbx.SpinorField(geom, p.Ls, p.nrhs, "f32")
    return geom, U, x, y


def chiral(gbx, f):
    f.set_basis(gbx.BASIS_CHIRAL)
    return f.download()


@pytest.mark.parametrize("dims,Ls,nrhs", CASES)
@pytest.mark.parametrize("tdiag", [False, True])
@pytest.mark.parametrize("dag", [False, True])
def test_mfma32_vs_oracle_and_valu(gpu, gbx, problem, dims, Ls, nrhs, tdiag, dag):
    p = problem(dims, Ls, nrhs=nrhs)
    geom, U, x, y = fields(gbx, p, tdiag=tdiag)
    print(f"mfma32 dims={dims} Ls={Ls} nrhs={nrhs} tdiag={tdiag} dag={dag}", flush=True)
    gbx.dhop(U, y, x, dag=dag, cfg=gbx.kcfg(2, 128, 1, 1, kind=1, mc=16, pf=1))
    assert y.basis == x.basis
    mf = chiral(gbx, y)
    gbx.dhop(U, y, x, dag=dag, cfg=gbx.kcfg(1, 64, 0, 0))
    va = chiral(gbx, y)
    d = rel_l2(mf, va)
    print(f"  rel_l2(MFMA - VALU) = {d:.2e}")
    assert d <= TOL
    for r in range(nrhs):
        err = rel_l2(mf[r], p.oracle_lex(rhs=r, dag=dag))
        print(f"  rhs {r}: rel_l2 vs FP64 oracle = {err:.2e}")
        assert err <= TOL


@pytest.mark.parametrize("tdiag", [False, True])
@pytest.mark.parametrize("dag", [False, True])
def test_mfma32_compressed_links(gpu, gbx, problem, tdiag, dag):
    p = problem((4, 4, 8, 4), 8, nrhs=4)
    geom, U, x, y = fields(gbx, p, comp12=True, tdiag=tdiag)
    gbx.dhop(U, y, x, dag=dag, cfg=gbx.kcfg(4, 256, 1, 0, kind=1, mc=0, pf=0))
    mf = chiral(gbx, y)
    for r in range(4):
        err = rel_l2(mf[r], p.oracle_lex(rhs=r, dag=dag))
        print(f"comp12 tdiag={tdiag} dag={dag} rhs {r}: {err:.2e}")
        assert err <= TOL


@pytest.mark.parametrize("tdiag", [False, True])
@pytest.mark.parametrize("spw,threads,xcd,nt,mc,pf,wgpc,cblk", [
    (1, 64, 0, 0, 0, 0, 0, 0), (1, 256, 1, 1, 16, 1, 0, 0), (3, 192, 1, 1, 32, 1, 0, 0), (7, 128, 0, 1, 64, 0, 4, 0),
    (16, 256, 1, 0, 16, 1, 2, 0), (5, 64, 1, 1, 32, 0, 0, 0), (4, 256, 1, 1, 16, 1, 0, 32), (3, 128, 1, 0, 16, 0, 0, 16)])
def test_mfma32_launch_shapes_identical(gpu, gbx, problem, tdiag, spw, threads, xcd, nt, mc, pf, wgpc, cblk):
    p = problem((6, 4, 4, 4), 16, nrhs=4)
    geom, U, x, y = fields(gbx, p, tdiag=tdiag)
    gbx.dhop(U, y, x, cfg=gbx.kcfg(1, 64, 0, 0, kind=1))
    want = chiral(gbx, y)  # the basis change is deterministic: equal bits stay equal
    y.zero()
    gbx.dhop(U, y, x, cfg=gbx.kcfg(spw, threads, xcd, nt, kind=1, mc=mc, pf=pf, wgpc=wgpc, cblk=cblk))
    assert np.array_equal(chiral(gbx, y), want)


def test_mfma32_full_size_24x24x24x24_ls16_n16(gpu, gbx):
    """The FP32 16-RHS benchmark shape (8.2 GB per field, TDIAG basis): the
    MFMA form against the VALU kernel on right-hand sides 0, 9 and 15."""
    dims = (24, 24, 24, 24)
    geom = gbx.Geometry(dims)
    U = gbx.GaugeField(geom, "f32").upload(gbx.gen_doubled_gauge(dims, 201))
    x = gbx.SpinorField(geom, 16, 16, "f32").fill_gaussian(1202).set_basis(gbx.BASIS_TDIAG)
    y = gbx.SpinorField(geom, 16, 16, "f32")
    gbx.dhop(U, y, x, cfg=gbx.kcfg(4, 256, 1, 1, kind=1, pf=1))
    y.set_basis(gbx.BASIS_CHIRAL)
    got = [y.download_rhs(r) for r in (0, 9, 15)]
    gbx.dhop(U, y, x, cfg=gbx.kcfg(4, 256, 1, 0, mc=64, pf=1))
    y.set_basis(gbx.BASIS_CHIRAL)
    for r, g in zip((0, 9, 15), got):
        d = rel_l2(g, y.download_rhs(r))
        print(f"rhs {r}: rel_l2(MFMA - VALU) = {d:.2e}")
        assert d <= TOL
    for f in (U, x, y):
        f.close()


def test_mfma32_rejects_what_it_does_not_cover(gpu, gbx, problem):
    p = problem((4, 4, 4, 4), 6, nrhs=1)
    geom, U, x, y = fields(gbx, p)
    with pytest.raises(gbx.GbxError) as e:  # Ls*nrhs = 6 is not a multiple of 16
        gbx.dhop(U, y, x, cfg=gbx.kcfg(1, 64, kind=1))
    assert e.value.status == 8
    p = problem((4, 4, 4, 4), 16, nrhs=1)
    geom, U, x, y = fields(gbx, p)
    for bad in (gbx.kcfg(1, 64, kind=1, mc=8), gbx.kcfg(1, 320, kind=1), gbx.kcfg(1, 64, kind=1, mc=16, cblk=24)):
        with pytest.raises(gbx.GbxError) as e:
            gbx.dhop(U, y, x, cfg=bad)
        assert e.value.status == 1


def test_tuner_searches_the_fp32_mfma_form(gpu, gbx, problem, tmp_path):
    p = problem((8, 8, 4, 8), 8, nrhs=4)
    geom, U, x, y = fields(gbx, p, tdiag=True)
    cfg, us = gbx.tune(U, y, x, kinds="mfma", cache_path=str(tmp_path / "m.jsonl"))
    assert cfg.kind == 1 and us > 0
    gbx.dhop(U, y, x, cfg=cfg)
    got = chiral(gbx, y)
    for r in range(4):
        assert rel_l2(got[r], p.oracle_lex(rhs=r)) <= TOL
