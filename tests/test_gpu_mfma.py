"""Matrix-core (MFMA) form of the multi-right-hand-side hopping term (kernel
kind 1, FP16/BF16 storage, FP32 accumulation; the FP32-storage form is
tested in test_gpu_mfma32.py): every right-hand side against
the FP64 oracle (same bounds as the VALU kernel on 16-bit storage), and the
whole output against the VALU kernel on the same 16-bit inputs.

Both bounds are conftest.storage_tol: 2.5 x the relative error of storing
the input field in the 16-bit format.  The two kernels sum in different
orders and the MFMA form rounds the rebuilt third link row of 12-real links
to 16 bits, so they agree to a fraction of the last stored bit rather than
exactly.
"""
import numpy as np
import pytest

from conftest import rel_l2, storage_tol

pytestmark = pytest.mark.gpu


CASES = [((4, 4, 4, 4), 8, 2), ((8, 4, 4, 8), 4, 8), ((4, 6, 2, 4), 16, 4), ((2, 2, 2, 2), 16, 1)]


def fields(gbx, p, prec, comp12=False):
    geom = gbx.Geometry(p.dims)
    U = gbx.GaugeField(geom, prec, comp12).upload(p.dbl)
    x = gbx.SpinorField(geom, p.Ls, p.nrhs, prec).upload(p.psi)
    y = gbx.SpinorField(geom, p.Ls, p.nrhs, prec)
    return geom, U, x, y


@pytest.mark.parametrize("dims,Ls,nrhs", CASES)
@pytest.mark.parametrize("prec", ["f16", "bf16"])
@pytest.mark.parametrize("dag", [False, True])
def test_mfma_vs_oracle_and_valu(gpu, gbx, problem, dims, Ls, nrhs, prec, dag):
    p = problem(dims, Ls, nrhs=nrhs)
    geom, U, x, y = fields(gbx, p, prec)
    gbx.dhop(U, y, x, dag=dag, cfg=gbx.kcfg(2, 128, 1, 1, kind=1, mc=16, pf=1))
    mf = y.download()
    gbx.dhop(U, y, x, dag=dag, cfg=gbx.kcfg(1, 64, 0, 0))
    va = y.download()
    d = rel_l2(mf, va)
    print(f"{prec} dag={dag} dims={dims} Ls={Ls} nrhs={nrhs}: rel_l2(MFMA - VALU) = {d:.2e}")
    assert d <= storage_tol(prec, p.psi)
    for r in range(nrhs):
        err = rel_l2(mf[r], p.oracle_lex(rhs=r, dag=dag))
        print(f"  rhs {r}: rel_l2 vs FP64 oracle = {err:.2e}")
        assert err <= storage_tol(prec, p.psi)


@pytest.mark.parametrize("prec", ["f16", "bf16"])
@pytest.mark.parametrize("dag", [False, True])
def test_mfma_compressed_links(gpu, gbx, problem, prec, dag):
    p = problem((4, 4, 8, 4), 8, nrhs=4)
    geom, U, x, y = fields(gbx, p, prec, comp12=True)
    gbx.dhop(U, y, x, dag=dag, cfg=gbx.kcfg(4, 256, 1, 0, kind=1, mc=0, pf=0))
    mf = y.download()
    gbx.dhop(U, y, x, dag=dag, cfg=gbx.kcfg(1, 64, 0, 0))
    d = rel_l2(mf, y.download())
    print(f"{prec} comp12 dag={dag}: rel_l2(MFMA - VALU) = {d:.2e}")
    assert d <= storage_tol(prec, p.psi)
    for r in range(4):
        assert rel_l2(mf[r], p.oracle_lex(rhs=r, dag=dag)) <= storage_tol(prec, p.psi)


@pytest.mark.parametrize("spw,threads,xcd,nt,mc,pf,wgpc,cblk", [
    (1, 64, 0, 0, 0, 0, 0, 0), (1, 256, 1, 1, 16, 1, 0, 0), (3, 192, 1, 1, 32, 1, 0, 0), (7, 128, 0, 1, 64, 0, 4, 0),
    (16, 256, 1, 0, 16, 1, 2, 0), (5, 64, 1, 1, 32, 0, 0, 0), (4, 256, 1, 1, 16, 1, 0, 32), (3, 128, 1, 0, 16, 0, 0, 16),
    (2, 64, 0, 1, 32, 1, 0, 32)])
def test_mfma_launch_shapes_identical(gpu, gbx, problem, spw, threads, xcd, nt, mc, pf, wgpc, cblk):
    """Every launch shape -- including lattice sweeps over column chunks
    (cblk) -- computes each output element by the same sequence of MFMAs, so
    the results are bit-identical across shapes."""
    p = problem((6, 4, 4, 4), 16, nrhs=4)
    geom, U, x, y = fields(gbx, p, "f16")
    gbx.dhop(U, y, x, cfg=gbx.kcfg(1, 64, 0, 0, kind=1))
    base = y.download()
    y.zero()
    gbx.dhop(U, y, x, cfg=gbx.kcfg(spw, threads, xcd, nt, kind=1, mc=mc, pf=pf, wgpc=wgpc, cblk=cblk))
    assert np.array_equal(y.download(), base)


def test_mfma_rejects_unsupported(gpu, gbx, problem):
    p = problem((4, 4, 4, 4), 6, nrhs=1)
    geom, U, x, y = fields(gbx, p, "f16")
    with pytest.raises(gbx.GbxError) as e:  # Ls*nrhs = 6 is not a multiple of 16
        gbx.dhop(U, y, x, cfg=gbx.kcfg(1, 64, kind=1))
    assert e.value.status == 8
    p = problem((4, 4, 4, 4), 16, nrhs=1)
    geom, U, x, y = fields(gbx, p, "f64")
    with pytest.raises(gbx.GbxError) as e:  # FP64 storage has no MFMA form
        gbx.dhop(U, y, x, cfg=gbx.kcfg(1, 64, kind=1))
    assert e.value.status == 8
    geom, U, x, y = fields(gbx, p, "f16")
    for bad in (gbx.kcfg(1, 64, kind=1, mc=8), gbx.kcfg(1, 320, kind=1), gbx.kcfg(1, 64, kind=1, mc=48),
                gbx.kcfg(1, 64, kind=1, mc=16, cblk=24)):
        with pytest.raises(gbx.GbxError) as e:
            gbx.dhop(U, y, x, cfg=bad)
        assert e.value.status == 1


def test_tuner_considers_mfma(gpu, gbx, problem, tmp_path):
    """The tuner times the MFMA form next to the VALU kernel on 16-bit
    multi-RHS fields; whichever it picks agrees with the oracle."""
    p = problem((8, 8, 8, 8), 8, nrhs=4)
    geom, U, x, y = fields(gbx, p, "f16")
    cfg, us = gbx.tune(U, y, x, cache_path=str(tmp_path / "t.jsonl"))
    print(f"tuned: {cfg.as_dict()} {us:.1f} us")
    gbx.dhop(U, y, x)
    got = y.download()
    for r in range(4):
        assert rel_l2(got[r], p.oracle_lex(rhs=r)) <= storage_tol("f16", p.psi)


def test_tune_restricted_to_one_kind(gpu, gbx, problem, tmp_path):
    """gbx_tune_kinds: each restricted search returns a configuration of its
    own kind, both agree with the oracle, neither changes what gbx_dhop uses;
    FP64 storage has no MFMA candidates (GBX_EUNSUPPORTED)."""
    p = problem((8, 4, 4, 8), 8, nrhs=2)
    geom, U, x, y = fields(gbx, p, "bf16")
    before = gbx.dhop_get_cfg(U, x).as_dict()
    for kind, code in (("valu", 0), ("mfma", 1)):
        cfg, us = gbx.tune(U, y, x, kinds=kind, cache_path=str(tmp_path / "k.jsonl"))
        assert cfg.kind == code and us > 0
        gbx.dhop(U, y, x, cfg=cfg)
        got = y.download()
        for r in range(2):
            assert rel_l2(got[r], p.oracle_lex(rhs=r)) <= storage_tol("bf16", p.psi)
    assert gbx.dhop_get_cfg(U, x).as_dict() == before
    geom, U, x, y = fields(gbx, p, "f64")
    with pytest.raises(gbx.GbxError) as e:
        gbx.tune(U, y, x, kinds="mfma")
    assert e.value.status == 8
